"""The oracle against the committed golden fixtures (tests/golden/*.npz, written by
tests/golden/make_golden.py): a change to the oracle that moves any fixture's fields, forces,
populations or adjoint gradients fails here, on the CPU."""
import glob
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = sorted(glob.glob(os.path.join(ROOT, "tests", "golden", "*.npz")))


def load(path):
    z = np.load(path)                      # allow_pickle stays False: plain arrays and a str
    meta = json.loads(str(z["meta"]))
    return meta, {k[3:]: z[k] for k in z.files if k.startswith("in_")}, {k[4:]: z[k] for k in z.files if k.startswith("out_")}


def oracle_cfg(orc, cfgkw):
    kw = dict(cfgkw)
    Q, nx, ny, nz = kw.pop("Q"), kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
    for k in ("center", "periodic", "coef"):
        if k in kw:
            kw[k] = tuple(kw[k])
    return orc.make_cfg(Q, nx, ny, nz, **kw)


def test_fixtures_are_present():
    assert len(GOLDEN) >= 7


@pytest.mark.parametrize("path", GOLDEN, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_reproduces_fixture(orc, path):
    meta, inp, out = load(path)
    cfg = oracle_cfg(orc, meta["cfg"])
    if meta["mode"] == "plain":
        s = orc.PlainState(cfg, mask=inp.get("mask"), sfrac=inp.get("sfrac"))
        s.step(meta["steps"])
        assert np.array_equal(np.stack(s.fields()), out["fields"])
        assert np.array_equal(s.forces(), out["forces"])
        f = s.populations()
        if "population_columns" in out:
            f = f[..., out["population_columns"]]
        assert np.array_equal(f, out["populations"])
    elif meta["mode"] == "ref":
        s = orc.RefState(cfg, ran=inp.get("ran"))
        s.step(meta["steps"])
        assert np.array_equal(np.stack(s.fields()), out["fields"])
        assert np.array_equal(s.forces(), out["forces"])
    else:
        J, ds, du = orc.adjoint(cfg, mask=inp.get("mask"), sfrac=inp.get("sfrac"), n0=meta["n0"], n1=meta["steps"],
                                scale=1.0, K=meta["K"])
        assert J == float(out["J"])
        assert np.array_equal(ds, out["dJds"]) and np.array_equal(du, out["dJdu"])
