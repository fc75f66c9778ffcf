"""Storage-format accuracy against FP64 on a reduced copy of the 512x256x256 cylinder workload.

D3Q27 192x96x64, z-spanning cylinder D = 24 at (48, 48), Re 300, u 0.05, walls on the four
sides, 4000 steps, drag/lift every 10 steps.  The FP64 run is the oracle's algorithm on the GPU
(bit-identical to the CPU oracle on the parity cases).  Compared: mean drag coefficient over
steps 2000-4000 and the velocity field after 1000 steps (relative L2).  Observed on the MI355X:
FP32 2.0e-5 / 1.2e-5, FP16-shifted 2.2e-4 / 3.5e-4, BF16-shifted 1.6e-3 / 3.0e-3; the budgets
below leave a factor of 3-5 and hold FP16 to the 1e-3 drag target.  The lift is not compared:
the wake has not started to shed by step 4000 (its RMS is ~2 % of the mean drag and noise-led).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

BUDGET = {"fp32": (1e-4, 1e-4), "fp16s": (1e-3, 2e-3), "bf16s": (1e-2, 1.5e-2)}


def run(pkg, prec, steps=4000, early=1000):
    L = pkg.lbm
    nx, ny, nz, D, Re, u = 192, 96, 64, 24.0, 300.0, 0.05
    c = L.make_config(27, nx, ny, nz, tau0=3 * u * D / Re, u0=u, precision=L.PRECISIONS[prec], body=L.BODY_ZCYL,
                      center=(nx / 4.0, ny / 2.0, 0.0), r2=(D / 2) ** 2, seed=pkg.configs.SEED, perturb=1e-3,
                      force_interval=10)
    with L.Solver(c) as s:
        s.initialize()
        s.step(early)
        vel = np.stack(s.fields()[1:])
        s.step(steps - early)
        F = s.forces()
    scale = 2.0 / (u * u * D * nz)
    cd = scale * F[F[:, 0] > steps // 2, 1]
    return cd.mean(), vel


def test_storage_formats_against_fp64(pkg):
    print("[accuracy] 192x96x64 cylinder Re 300, FP64 reference", flush=True)
    cd64, v64 = run(pkg, "fp64")
    assert 1.0 < cd64 < 4.0
    for prec, (cd_tol, vel_tol) in BUDGET.items():
        cd, vel = run(pkg, prec)
        cd_err = abs(cd - cd64) / abs(cd64)
        vel_err = np.sqrt(np.sum((vel - v64) ** 2) / np.sum(v64 ** 2))
        print(f"[accuracy] {prec}: Cd {cd:.6f} (FP64 {cd64:.6f}) rel err {cd_err:.2e}; velocity rel-L2 {vel_err:.2e}",
              flush=True)
        assert cd_err <= cd_tol
        assert vel_err <= vel_tol
