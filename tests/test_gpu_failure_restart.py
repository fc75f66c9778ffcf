"""Failure detection and checkpoint/restart through the C ABI (no reference counterpart: the
reference aborts on errors and has no restart).

  * lbm_sync reports LBM_ERR_UNSTABLE for a non-finite population in the state, whether or not
    forces are recorded: a NaN injected through lbm_set_populations (plain mode), and a run
    driven unstable by its parameters (tau0 -> 0 at a high inlet velocity; reference mode and
    plain mode), each forced from the host, so every kernel still exits normally;
  * a run split in two (populations and step counter saved, a fresh context restored with
    lbm_set_populations + lbm_set_step) continues bit for bit as the uninterrupted run:
    populations, fields and the force records after the split.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _cfg(L, prec, bc=None, **kw):
    base = dict(tau0=0.06, u0=0.05, precision=L.PRECISIONS[prec], bc_mode=L.BC_PLAIN if bc is None else bc,
                body=L.BODY_SPHERE, center=(14.0, 8.5, 6.5), r2=9.0, force_interval=0)
    if base["bc_mode"] == L.BC_PLAIN:
        base["perturb"] = 1e-3
    base.update(kw)
    return L.make_config(27, 40, 18, 14, **base)


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16s", "bf16s"])
def test_sync_reports_an_injected_nan_with_forces_off(pkg, prec):
    L = pkg.lbm
    print(f"[failure] injected NaN {prec}", flush=True)
    with pkg.Solver(_cfg(L, prec)) as s:
        s.initialize()
        s.step(4)
        s.sync()
        f = s.populations()
        f[5, 7, 3, 30] = np.nan           # a fluid node far from the body
        s.set_populations(f)
        s.step(2)
        with pytest.raises(L.LbmError) as ei:
            s.sync()
        assert ei.value.code == 5
        assert "non-finite" in str(ei.value)


@pytest.mark.parametrize("bc", ["reference", "plain"])
def test_sync_reports_divergence_with_forces_off(pkg, bc):
    """tau0 = 1e-4 (relaxation time 0.5001) at inlet velocity 0.5 diverges within ~20 steps (the
    CPU oracle's reference mode: step 20 on this box); after 60 steps lbm_sync must say so."""
    L = pkg.lbm
    print(f"[failure] divergence {bc}", flush=True)
    mode = L.BC_REFERENCE if bc == "reference" else L.BC_PLAIN
    with pkg.Solver(_cfg(L, "fp64", bc=mode, tau0=1e-4, u0=0.5)) as s:
        s.initialize()
        s.step(60)
        with pytest.raises(L.LbmError) as ei:
            s.sync()
        assert ei.value.code == 5


def test_sync_passes_a_healthy_run(pkg):
    L = pkg.lbm
    for bc in (L.BC_REFERENCE, L.BC_PLAIN):
        with pkg.Solver(_cfg(L, "fp64", bc=bc, force_interval=5)) as s:
            s.initialize()
            s.step(30)
            s.sync()


@pytest.mark.parametrize("prec,dense", [("fp64", "0"), ("fp32", "0"), ("fp32", "1"), ("fp16s", "0"), ("bf16s", "0")])
def test_split_run_restarts_bit_for_bit(pkg, monkeypatch, prec, dense):
    L = pkg.lbm
    monkeypatch.setenv("LBM_DENSE", dense)
    g = _cfg(L, prec, force_interval=3)
    print(f"[restart] {prec} dense={dense}", flush=True)
    with pkg.Solver(g) as a:
        a.initialize()
        a.step(20)
        fa, ma, Fa = a.populations(), a.fields(), a.forces()
    with pkg.Solver(g) as b:
        b.initialize()
        b.step(10)
        saved, step = b.populations(), b.info().step
    assert step == 10
    with pkg.Solver(g) as c:
        c.initialize()
        c.set_populations(saved)
        c.set_step(step)
        c.step(10)
        fc, mc, Fc = c.populations(), c.fields(), c.forces()
        assert c.info().step == 20
    assert np.array_equal(fa, fc)
    for x, y in zip(ma, mc):
        assert np.array_equal(x, y)
    assert np.array_equal(Fc, Fa[Fa[:, 0] > 10])     # records of steps 12, 15, 18


@pytest.mark.parametrize("tb", ["1", "0"])
def test_split_run_restarts_bit_for_bit_d2q9_fp64(pkg, monkeypatch, tb):
    """The D2Q9 FP64 plain path of small domains, with several steps per launch (whose force
    steps within a launch follow from the restored step counter) and with one launch per step:
    a run split at step 11 and restored into a new context equals the uninterrupted 25 steps."""
    L = pkg.lbm
    monkeypatch.setenv("LBM_TEMPORAL_BLOCKING", tb)
    g = L.make_config(9, 100, 40, 1, tau0=0.06, u0=0.05, precision=L.FP64, bc_mode=L.BC_PLAIN, perturb=1e-3,
                      force_interval=3, body=L.BODY_SPHERE, center=(25.0, 19.5, 0.0), r2=25.0)
    print(f"[restart] d2q9 fp64 temporal blocking {tb}", flush=True)
    with pkg.Solver(g) as a:
        a.initialize()
        a.step(25)
        fa, ma, Fa = a.populations(), a.fields(), a.forces()
    with pkg.Solver(g) as b:
        b.initialize()
        b.step(11)
        saved, step = b.populations(), b.info().step
    assert step == 11
    with pkg.Solver(g) as c:
        c.initialize()
        c.set_populations(saved)
        c.set_step(step)
        c.step(14)
        fc, mc, Fc = c.populations(), c.fields(), c.forces()
        assert c.info().step == 25
    assert np.array_equal(fa, fc)
    for x, y in zip(ma, mc):
        assert np.array_equal(x, y)
    assert np.array_equal(Fc, Fa[Fa[:, 0] > 11])     # records of steps 12, 15, 18, 21, 24
