"""Writes the golden fixtures of this directory from the FP64 CPU oracle (test infrastructure).

Each fixture is one compressed .npz holding the full input -- the solver configuration as JSON
(keyword arguments of oracle.make_cfg / cylinder_flow_amd.lbm.make_config), and the mask,
solid-fraction or injected ran arrays when the case has them -- and the expected output after
`steps` steps: the macroscopic fields (p, u, v[, w]) on the physical nodes, the force record
(step, Fx, Fy, Fz) and the populations (D3Q27: on three x columns, listed in
out_population_columns).  The reference ships no test
vectors for this path, so the fixtures freeze the oracle's results on seeded inputs; the
oracle itself is pinned by tests/test_oracle.py.

    python tests/golden/make_golden.py          # rewrites every fixture
    python tests/golden/make_golden.py NAME...  # rewrites the named ones
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))

import oracle as O  # noqa: E402

LEGENDRE = (0.04, 0.004, -0.003, 0.002, 0.001, -0.002, 0.001, 0.0005)

# name: (mode, cfg kwargs, steps, extras)
CASES = {
    "plain_d2q9_disk_walls": ("plain", dict(Q=9, nx=48, ny=24, nz=1, tau0=0.06, u0=0.05, body=O.BODY_SPHERE,
                                            center=(14.0, 11.5, 0.0), r2=16.0, perturb=1e-3, force_interval=5), 60, {}),
    "plain_d2q9_periodic_sfrac": ("plain", dict(Q=9, nx=40, ny=20, nz=1, tau0=0.05, u0=0.05, perturb=1e-3,
                                                periodic=(True, True, False), force_interval=4), 50, {"sfrac": 9}),
    "plain_d3q27_sphere_parabolic": ("plain", dict(Q=27, nx=24, ny=14, nz=12, tau0=0.06, u0=0.05,
                                                   body=O.BODY_SPHERE, center=(8.0, 6.5, 5.5), r2=9.0,
                                                   inlet=O.INLET_PARABOLIC, perturb=1e-3, force_interval=3), 30, {}),
    "plain_d3q27_mask_legendre": ("plain", dict(Q=27, nx=20, ny=12, nz=10, tau0=0.05, u0=0.05, body=O.BODY_MASK,
                                                inlet=O.INLET_LEGENDRE, coef=LEGENDRE, perturb=1e-3,
                                                periodic=(False, False, True), force_interval=2), 24,
                                  {"mask": 3, "sfrac": 4}),
    "ref_d2q9_disk": ("ref", dict(Q=9, nx=64, ny=32, nz=1, tau0=0.036, u0=0.1, body=O.BODY_SPHERE,
                                  center=(20.0, 15.5, 0.0), r2=16.0, force_interval=1), 40, {}),
    "ref_d3q27_sphere": ("ref", dict(Q=27, nx=32, ny=16, nz=14, tau0=0.024, u0=0.1, body=O.BODY_SPHERE,
                                     center=(10.0, 7.0, 6.0), r2=9.0, force_interval=1), 20, {"ran": 17}),
    "ref_d2q9_single_node": ("ref", dict(Q=9, nx=49, ny=21, nz=1, tau0=0.036, u0=0.1, body=O.BODY_SPHERE,
                                         center=(0.25 * 48 - 2.0, 0.5 * 20, 0.0), r2=0.01 * 20 * 20,
                                         force_interval=5, variant=1), 40, {"ran": 23}),
    "adjoint_d2q9_mask_sfrac": ("adjoint", dict(Q=9, nx=24, ny=12, nz=1, tau0=0.1, u0=0.05, body=O.BODY_MASK,
                                                inlet=O.INLET_PARABOLIC, perturb=1e-3), 20,
                                {"mask": 5, "sfrac": 6, "n0": 8, "K": 7}),
}


def inputs(cfgkw, extras):
    nz, ny, nx = cfgkw["nz"], cfgkw["ny"], cfgkw["nx"]
    shape = (nz, ny, nx)
    out = {}
    if "mask" in extras:
        rng = np.random.default_rng(extras["mask"])
        m = np.zeros(shape, np.uint8)
        m[:, ny // 3:ny // 3 + 3, nx // 3:nx // 3 + 2] = 1
        m[tuple(rng.integers(0, n, 4) for n in shape)] = 1
        m[:, :, 0] = 0  # keep the inlet column fluid
        out["mask"] = m
    if "sfrac" in extras:
        rng = np.random.default_rng(extras["sfrac"])
        zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
        s = np.exp(-((xx - nx / 2.0) ** 2 + (yy - ny / 2.0) ** 2) / 6.0)
        out["sfrac"] = np.clip(s + 0.01 * rng.standard_normal(shape), 0.0, 1.0)
    if "ran" in extras:
        g = 3 if cfgkw["Q"] == 27 else 2
        n = (nx + 2 * g) * (ny + 2 * g) * ((nz + 2 * g) if cfgkw["Q"] == 27 else 1)
        out["ran"] = np.random.default_rng(extras["ran"]).integers(0, 100, n).astype(np.int32)
    return out


def make_cfg(cfgkw):
    kw = dict(cfgkw)
    Q, nx, ny, nz = kw.pop("Q"), kw.pop("nx"), kw.pop("ny"), kw.pop("nz")
    return O.make_cfg(Q, nx, ny, nz, **kw)


def run(name):
    mode, cfgkw, steps, extras = CASES[name]
    arrs = inputs(cfgkw, extras)
    cfg = make_cfg(cfgkw)
    res = {}
    if mode == "plain":
        s = O.PlainState(cfg, mask=arrs.get("mask"), sfrac=arrs.get("sfrac"))
        s.step(steps)
        res["fields"] = np.stack(s.fields())
        res["forces"] = s.forces()
        f = s.populations()
        if cfgkw["Q"] == 27:   # three x columns: inlet, through the body, outlet
            cols = np.array([0, cfgkw["nx"] // 3, cfgkw["nx"] - 1])
            res["population_columns"] = cols
            f = f[..., cols]
        res["populations"] = f
    elif mode == "ref":
        s = O.RefState(cfg, ran=arrs.get("ran"))
        s.step(steps)
        res["fields"] = np.stack(s.fields())
        res["forces"] = s.forces()
    else:
        J, ds, du = O.adjoint(cfg, mask=arrs.get("mask"), sfrac=arrs.get("sfrac"), n0=extras["n0"], n1=steps,
                              scale=1.0, K=extras["K"])
        res["J"] = np.array(J)
        res["dJds"] = ds
        res["dJdu"] = du
    meta = {"mode": mode, "cfg": cfgkw, "steps": steps,
            "n0": extras.get("n0"), "K": extras.get("K")}
    return meta, arrs, res


def main():
    for name in (sys.argv[1:] or CASES):
        meta, arrs, res = run(name)
        path = os.path.join(HERE, name + ".npz")
        np.savez_compressed(path, meta=np.array(json.dumps(meta)), **{"in_" + k: v for k, v in arrs.items()},
                            **{"out_" + k: v for k, v in res.items()})
        print(f"{path}: {os.path.getsize(path)} bytes")


if __name__ == "__main__":
    main()
