"""Pins of the oracle's reference mode by identities the reference's formulas imply (see
ref_pins.py for the derivations), and a check that each pin catches a planted mistake: a copy of
oracle/lbm_oracle.c with one formula changed is built in a temporary directory, and the pins run
against it in a subprocess must fail.  CPU only."""
import json
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

import ref_pins as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")


@pytest.mark.parametrize("Q", [9, 27])
def test_body_link_mirror_identity(orc, Q):
    n, bad = P.body_link_mirror(orc, Q)
    assert n >= (70 if Q == 9 else 1000) and bad == 0


@pytest.mark.parametrize("Q", [9, 27])
def test_wall_rows_mirror_identity(orc, Q):
    n, bad = P.wall_mirror(orc, Q)
    assert n >= (200 if Q == 9 else 10000) and bad == 0


@pytest.mark.parametrize("Q", [9, 27])
def test_inlet_copy_and_outlet_extrapolation_identities(orc, Q):
    n, bad = P.inlet_outlet_copies(orc, Q)
    assert n >= (100 if Q == 9 else 3000) and bad == 0


@pytest.mark.parametrize("Q", [9, 27])
def test_reference_bulk_equals_plain_bulk(orc, Q):
    """Collision, stream and Update of the reference mode agree with the plain-mode step (pinned
    by the numpy step and the Taylor-Green rate) away from all boundaries.  Observed bit for bit;
    the bound allows the moment sums a different rounding."""
    dpop, dmac, n = P.bulk_vs_plain(orc, Q)
    assert n > 1000 and dpop <= 1e-15 and dmac <= 1e-15


@pytest.mark.parametrize("Q", [9, 27])
def test_reference_momentum_exchange_step1_closed_form(orc, Q):
    F, Fc, links = P.momentum_exchange_step1(orc, Q)
    assert links >= 70
    assert abs(F[0] - Fc[0]) <= 1e-14 * abs(Fc[0]) and Fc[0] > 0
    # symmetric initial state (disk/sphere on the channel axis, ran = 0): no lift
    assert np.abs(F[1:]).max() <= 1e-15 * abs(F[0])


# ------------------------------------------------------------------- planted mistakes
# (name, exact line fragment of oracle/lbm_oracle.c, replacement, pins that must fail)
MUTANTS = [
    ("usrbc reads direction ii instead of bb(ii)",
     "RF(s, ii, i, j, k) = RF(s, L->bb[ii], i + 2 * e[0], j + 2 * e[1], k + 2 * e[2]);",
     "RF(s, ii, i, j, k) = RF(s, ii, i + 2 * e[0], j + 2 * e[1], k + 2 * e[2]);", ["body9", "body27"]),
    ("collision relaxes with 1/tau0",
     "(RF(s, ii, i, j, k) - feq) / (oreal)(s->tau0 + 0.5);", "(RF(s, ii, i, j, k) - feq) / (oreal)s->tau0;",
     ["bulk9", "bulk27"]),
    ("momentum exchange with the sign of e_x flipped",
     "fx += m * (double)e[0];", "fx -= m * (double)e[0];", ["mem9", "mem27"]),
    ("momentum exchange reads f_ii at the solid node",
     "double m = RF(s, ii, i, j, k) + RF(s, L->bb[ii], i + e[0], j + e[1], k + e[2]);",
     "double m = RF(s, ii, i, j, k) + RF(s, ii, i + e[0], j + e[1], k + e[2]);", ["mem9", "mem27"]),
    ("3D y-wall fill reads the first interior row",
     "RF(s, ii, i, ls - 1, k) = RF(s, L->bb[ii], i + 2 * L->e[ii][0], ls + 1, k + 2 * L->e[ii][2]);",
     "RF(s, ii, i, ls - 1, k) = RF(s, L->bb[ii], i + 2 * L->e[ii][0], ls, k + 2 * L->e[ii][2]);", ["wall27"]),
    ("2D bottom wall fill shifts x by one instead of two",
     "RF(s, 5, i, 1, 0) = RF(s, 7, i + 2, 3, 0);", "RF(s, 5, i, 1, 0) = RF(s, 7, i + 1, 3, 0);", ["wall9"]),
    ("3D inlet ghost copies the second interior column",
     "RF(s, ii, g - 1, j, k) = RF(s, ii, g, j, k);", "RF(s, ii, g - 1, j, k) = RF(s, ii, g + 1, j, k);", ["io27"]),
    ("3D outlet extrapolation drops the lateral shift",
     "RF(s, ii, le0, j, k) = RF(s, ii, le0 + L->e[ii][0], j + L->e[ii][1], k + L->e[ii][2]);",
     "RF(s, ii, le0, j, k) = RF(s, ii, le0 + L->e[ii][0], j, k);", ["io27"]),
    ("2D outlet extrapolation of direction 7 from the same row",
     "RF(s, 7, lx - 2, j, 0) = RF(s, 7, lx - 3, j - 1, 0);", "RF(s, 7, lx - 2, j, 0) = RF(s, 7, lx - 3, j, 0);",
     ["io9"]),
    ("stream1 pulls along +e",
     "RFT(s, ii, i, j, k) = RF(s, ii, i - e[0], j - e[1], k - e[2]);",
     "RFT(s, ii, i, j, k) = RF(s, ii, i + e[0], j + e[1], k + e[2]);", ["bulk9", "bulk27"]),
    ("Update divides by 2 instead of 3",
     "s->p[N] = s->p[N] + fv / (oreal)3;", "s->p[N] = s->p[N] + fv / (oreal)2;", ["bulk9", "bulk27"]),
]

RUNNER = r"""
import json, sys
sys.path.insert(0, sys.argv[1]); sys.path.insert(0, sys.argv[2])
import numpy as np
import oracle as O
import ref_pins as P
out = {}
for Q in (9, 27):
    n, bad = P.body_link_mirror(O, Q); out["body%d" % Q] = bad == 0
    n, bad = P.wall_mirror(O, Q); out["wall%d" % Q] = bad == 0
    n, bad = P.inlet_outlet_copies(O, Q); out["io%d" % Q] = bad == 0
    dpop, dmac, n = P.bulk_vs_plain(O, Q); out["bulk%d" % Q] = bool(dpop <= 1e-15 and dmac <= 1e-15)
    F, Fc, links = P.momentum_exchange_step1(O, Q)
    out["mem%d" % Q] = bool(abs(F[0] - Fc[0]) <= 1e-14 * abs(Fc[0]) and Fc[0] > 0 and
                            np.abs(F[1:]).max() <= 1e-15 * abs(F[0]))
print(json.dumps(out))
"""


def _run_pins(lib_path):
    env = dict(os.environ, LBM_ORACLE_LIB=lib_path, OMP_NUM_THREADS="2")
    here = os.path.dirname(os.path.abspath(__file__))
    r = subprocess.run([sys.executable, "-c", RUNNER, ORACLE_DIR, here], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def _build_mutant(tmp, old, new):
    src = open(os.path.join(ORACLE_DIR, "lbm_oracle.c")).read()
    assert src.count(old) == 1, old
    d = os.path.join(tmp, "m")
    os.makedirs(d, exist_ok=True)
    shutil.copy(os.path.join(ORACLE_DIR, "lbm_oracle.h"), d)
    with open(os.path.join(d, "lbm_oracle.c"), "w") as fh:
        fh.write(src.replace(old, new))
    so = os.path.join(d, "liboracle_mutant.so")
    subprocess.run(["gcc", "-O2", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-shared", "-o", so,
                    os.path.join(d, "lbm_oracle.c"), "-lm"], check=True)
    return so


def test_pins_pass_on_the_oracle_in_a_subprocess(orc):
    res = _run_pins(os.path.join(ORACLE_DIR, "liboracle.so"))
    assert all(res.values()), res


@pytest.mark.parametrize("name,old,new,must_fail", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(tmp_path, name, old, new, must_fail):
    res = _run_pins(_build_mutant(str(tmp_path), old, new))
    assert not any(res[k] for k in must_fail), (name, res)
