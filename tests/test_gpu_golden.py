"""The HIP path against the committed golden fixtures (tests/golden/*.npz, oracle FP64 results
on seeded inputs), through the C ABI.  FP64 at round-off level (fields and populations 1e-13
relative L2, forces 1e-11 of their scale); FP32 2e-5 and FP16-shifted 3e-3 on the velocity
(plain mode); reference mode FP64 only; the adjoint fixture at 1e-9 (FP64) and 2e-3 (FP32)."""
import numpy as np
import pytest

from conftest import rel_l2
from test_golden import GOLDEN, load

pytestmark = pytest.mark.gpu


def gpu_solver(pkg, meta, inp, precision, bc_mode):
    L = pkg.lbm
    kw = dict(meta["cfg"])
    Q, nx, ny, nz = kw.pop("Q"), kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
    if kw.pop("variant", 0) == 1:   # the single-node 2D variant of the reference rules
        bc_mode = L.BC_REFERENCE_SN
    for k in ("center", "periodic", "coef"):
        if k in kw:
            kw[k] = tuple(kw[k])
    s = pkg.Solver(L.make_config(Q, nx, ny, nz, precision=precision, bc_mode=bc_mode, **kw))
    if "mask" in inp:
        s.set_mask(inp["mask"])
    if "sfrac" in inp:
        sf = inp["sfrac"]
        s.set_solid_fraction(sf if precision == L.FP64 else sf.astype(np.float32).astype(np.float64))
    if "ran" in inp:
        s.set_ran(inp["ran"])
    return s


def check_forces(Fg, Fo, tol):
    assert Fg.shape == Fo.shape and np.array_equal(Fg[:, 0], Fo[:, 0])
    scale = max(np.abs(Fo[:, 1:]).max(), 1e-300)
    assert np.abs(Fg[:, 1:] - Fo[:, 1:]).max() <= tol * scale


CASES = []
for path in GOLDEN:
    mode = load(path)[0]["mode"]
    precs = {"plain": ("fp64", "fp32", "fp16s"), "ref": ("fp64",), "adjoint": ("fp64", "fp32")}[mode]
    CASES += [pytest.param(path, p, id=f"{path.split('/')[-1][:-4]}-{p}") for p in precs]


@pytest.mark.parametrize("path,prec", CASES)
def test_hip_matches_fixture(pkg, path, prec):
    L = pkg.lbm
    meta, inp, out = load(path)
    precision = L.PRECISIONS[prec]
    print(f"[golden] {path.split('/')[-1]} {prec}", flush=True)
    if meta["mode"] == "adjoint":
        with gpu_solver(pkg, meta, inp, precision, L.BC_PLAIN) as s:
            s.initialize()
            J, ds, du, _, _ = s.adjoint(meta["n0"], meta["steps"], 1.0, meta["K"])
        tol = 1e-9 if prec == "fp64" else 2e-3
        assert abs(J - float(out["J"])) <= tol * abs(float(out["J"]))
        assert rel_l2(ds, out["dJds"]) <= tol and rel_l2(du, out["dJdu"]) <= tol
        return
    bc = L.BC_PLAIN if meta["mode"] == "plain" else L.BC_REFERENCE
    with gpu_solver(pkg, meta, inp, precision, bc) as s:
        s.initialize()
        s.step(meta["steps"])
        fields = np.stack(s.fields())
        forces = s.forces()
        pops = s.populations()
    ref = out["fields"]
    if prec == "fp64":
        assert rel_l2(fields, ref) <= 1e-13
        for a, b in zip(fields, ref):
            assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())
        check_forces(forces, out["forces"], 1e-11)
        if "populations" in out:
            if "population_columns" in out:
                pops = pops[..., out["population_columns"]]
            assert rel_l2(pops, out["populations"]) <= 1e-13
    else:
        budget = 2e-5 if prec == "fp32" else 3e-3
        err = rel_l2(fields[1], ref[1])
        print(f"[golden] velocity u rel-L2 {err:.2e}")
        assert err <= budget
