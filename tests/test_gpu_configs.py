"""The workloads of the two configurations with a CPU-checkable FP64 run, at their full sizes.

  config 0  D2Q9 400x100 cylinder (D 20, Re 100), FP64, the whole 20 000 steps against the CPU
            oracle: fields after the run and the drag/lift record of every 10th step
            (MPI Kokkos/lbm.cpp:356-407 logs Cd/Cl every step; here every 10th);
  config 1  D3Q27 256x128x128 sphere (D 24, Re 200): 100 FP64 steps against the CPU oracle, and
            10 000 steps in FP32 and FP16-shifted storage against the FP64 GPU run (itself
            bit-identical to the oracle wherever both run): mean Cd, RMS Cl, Strouhal number
            (none: the Re 200 sphere wake is steady) and the velocity's relative L2 error.
The measured figures are printed for the record.
"""
import time

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _oracle_cfg(orc, W):
    c = W.cfg
    Q = c.lattice
    return orc.make_cfg(Q, c.nx, c.ny, c.nz if Q == 27 else 1, tau0=c.tau0, u0=c.u0, body=c.body,
                        center=(c.cx, c.cy, c.cz), r2=c.r2, seed=c.seed, perturb=c.perturb,
                        force_interval=c.force_interval)


def test_config0_fp64_full_run_matches_the_oracle(pkg, orc):
    W = pkg.configs.cylinder2d_400x100()
    print(f"[config0] GPU FP64 {W.steps} steps", flush=True)
    with pkg.Solver(W.cfg) as s:
        s.initialize()
        t0 = time.perf_counter()
        s.step(W.steps)
        s.sync()
        wall = time.perf_counter() - t0
        Fg = s.forces()
        pg, ug, vg, _ = s.fields()
    print(f"[config0] CPU oracle {W.steps} steps", flush=True)
    o = orc.PlainState(_oracle_cfg(orc, W))
    t0 = time.perf_counter()
    o.step(W.steps)
    cpu = time.perf_counter() - t0
    Fo = o.forces()
    po, uo, vo, _ = o.fields()
    eu = rel_l2(np.stack([ug, vg]), np.stack([uo, vo]))
    ep = rel_l2(pg, po)
    assert Fg.shape == Fo.shape == (W.steps // 10, 4)
    assert np.array_equal(Fg[:, 0], Fo[:, 0])
    eF = np.abs(Fg[:, 1:3] - Fo[:, 1:3]).max() / np.abs(Fo[:, 1:3]).max()
    cd, cl = W.cd_scale() * Fg[:, 1], W.cd_scale() * Fg[:, 2]
    half = Fg[:, 0] > W.steps // 2
    # the lift also carries the channel's transverse acoustic mode (period 2 ny / c_s = 346 steps,
    # St 1.16): band-pass to St in [0.05, 0.5] (St 0.237 here, D/H = 0.2 blockage)
    st = pkg.configs.strouhal(cl[half], 10.0, W.D, W.u0, band=(0.05, 0.5))
    print(f"[config0] velocity rel-L2 {eu:.2e}, pressure rel-L2 {ep:.2e}, force records max rel diff {eF:.2e}; "
          f"mean Cd {cd[half].mean():.5f}, RMS Cl {cl[half].std():.5f}, St {st:.4f}; "
          f"GPU {W.cfg.nx * W.cfg.ny * W.steps / wall / 1e6:.0f} MLUPS (wall), CPU oracle {cpu:.1f} s", flush=True)
    assert eu <= 1e-12 and ep <= 1e-12
    assert eF <= 1e-12
    assert 1.5 < cd[half].mean() < 3.0 and cl[half].std() > 0.1 and 0.1 < st < 0.3   # shedding wake (D/H = 0.2)


def test_config1_fp64_full_size_matches_the_oracle(pkg, orc):
    W = pkg.configs.sphere3d_256x128x128(precision=pkg.lbm.FP64)
    n = 100
    print(f"[config1] GPU FP64 {n} steps", flush=True)
    with pkg.Solver(W.cfg) as s:
        s.initialize()
        s.step(n)
        Fg = s.forces()
        fg = s.fields()
    print(f"[config1] CPU oracle {n} steps", flush=True)
    o = orc.PlainState(_oracle_cfg(orc, W))
    o.step(n)
    Fo = o.forces()
    fo = o.fields()
    e = max(rel_l2(a, b) for a, b in zip(fg[:3], fo[:3]))
    eF = np.abs(Fg[:, 1:] - Fo[:, 1:]).max() / np.abs(Fo[:, 1:]).max()
    print(f"[config1] FP64 {n} steps: fields max rel-L2 {e:.2e}, force records max rel diff {eF:.2e}", flush=True)
    assert e <= 1e-12 and eF <= 1e-12


def test_config1_storage_formats_against_fp64(pkg):
    L = pkg.lbm
    res = {}
    for name, prec in (("fp64", L.FP64), ("fp32", L.FP32), ("fp16s", L.FP16S)):
        W = pkg.configs.sphere3d_256x128x128(precision=prec)
        print(f"[config1] {name} {W.steps} steps", flush=True)
        with pkg.Solver(W.cfg) as s:
            s.initialize()
            t0 = time.perf_counter()
            s.step(W.steps)
            s.sync()
            wall = time.perf_counter() - t0
            F = s.forces()
            _, u, v, w = s.fields()
        sel = F[:, 0] > W.steps // 2
        cd, cl = W.cd_scale() * F[sel, 1], W.cd_scale() * F[sel, 2]
        res[name] = dict(cd=cd.mean(), cl=np.sqrt(np.mean(cl ** 2)),
                         st=pkg.configs.strouhal(cl, 10.0, W.D, W.u0, band=(0.05, 0.5)), u=np.stack([u, v, w]),
                         mlups=W.cfg.nx * W.cfg.ny * W.cfg.nz * W.steps / wall / 1e6)
    r = res["fp64"]
    print(f"[config1] fp64: mean Cd {r['cd']:.6f}, RMS Cl {r['cl']:.2e}, St {r['st']}, {r['mlups']:.0f} MLUPS", flush=True)
    for name in ("fp32", "fp16s"):
        x = res[name]
        ecd = abs(x["cd"] - r["cd"]) / abs(r["cd"])
        eu = rel_l2(x["u"], r["u"])
        print(f"[config1] {name}: mean Cd {x['cd']:.6f} (rel err {ecd:.2e}), RMS Cl {x['cl']:.2e}, "
              f"velocity rel-L2 {eu:.2e}, {x['mlups']:.0f} MLUPS", flush=True)
        # about 3x the observed errors (fp32: Cd 3.6e-5, velocity 1.4e-5; fp16s: 3.8e-4, 3.7e-4);
        # fp16s drag also within the 1e-3 target
        assert ecd <= (1e-4 if name == "fp32" else 1e-3)
        assert eu <= (5e-5 if name == "fp32" else 1e-3)
