"""GPU (HIP) vs CPU oracle parity on seeded small inputs, through the C ABI.

FP64 runs are compared at round-off level (the kernels follow the oracle's rounding order,
both built without FMA contraction); forces differ only by summation order.  FP32 and the
16-bit storage formats are compared to the FP64 oracle with stated error budgets.
"""
import os

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def plain_pair(pkg, orc, Q, nx, ny, nz=1, *, precision, body=0, center=(0.0, 0.0, 0.0), r2=0.0,
               periodic=(False, False, False), inlet=0, coef=(), perturb=1e-3, tau0=0.06, u0=0.05,
               force_interval=1, store_uref=0.0):
    L = pkg.lbm
    g = L.make_config(Q, nx, ny, nz, tau0=tau0, u0=u0, precision=precision, bc_mode=L.BC_PLAIN,
                      periodic=periodic, body=body, center=center, r2=r2, inlet=inlet, coef=coef,
                      perturb=perturb, force_interval=force_interval, store_uref=store_uref)
    o = orc.make_cfg(Q, nx, ny, nz, tau0=tau0, u0=u0, body=body, center=center, r2=r2, inlet=inlet,
                     coef=coef, perturb=perturb, periodic=periodic, force_interval=force_interval)
    return g, o


PLAIN_CASES = {
    # name: (Q, nx, ny, nz, kwargs, steps)
    "d2q9_disk_walls_uniform": (9, 72, 40, 1, dict(body=1, center=(20.0, 19.5, 0.0), r2=25.0), 80),
    "d2q9_channel_parabolic": (9, 70, 33, 1, dict(inlet=1), 60),
    "d2q9_periodic_disk": (9, 48, 36, 1, dict(body=1, center=(24.0, 17.0, 0.0), r2=20.0,
                                              periodic=(True, True, False)), 60),
    "d3q27_sphere_walls_parabolic": (27, 36, 20, 18, dict(body=1, center=(10.0, 9.5, 8.7), r2=16.0, inlet=1), 40),
    "d3q27_zcyl_legendre": (27, 40, 22, 14, dict(body=2, center=(12.0, 11.0, 0.0), r2=12.0, inlet=2,
                                                  coef=(0.04, 0.005, -0.004, 0.003, 0.002, -0.001, 0.001, 0.0005)), 40),
    "d3q27_periodic_yz_sphere": (27, 32, 18, 16, dict(body=1, center=(12.0, 9.0, 8.0), r2=10.0,
                                                       periodic=(False, True, True)), 40),
    "d3q27_fully_periodic": (27, 24, 20, 18, dict(body=1, center=(12.0, 10.0, 9.0), r2=9.0,
                                                   periodic=(True, True, True)), 30),
    # ragged and minimal extents: odd nx (the last lane reaches into the padding), thin
    # channels, the smallest accepted box
    "d3q27_odd_sizes_sphere": (27, 37, 13, 11, dict(body=1, center=(11.0, 6.3, 5.1), r2=9.0, inlet=1), 30),
    "d2q9_odd_thin_disk": (9, 101, 7, 1, dict(body=1, center=(30.0, 3.2, 0.0), r2=2.0), 50),
    "d3q27_minimal_3x3x3": (27, 3, 3, 3, dict(), 20),
    "d2q9_minimal_3x3_periodic_y": (9, 3, 3, 1, dict(periodic=(False, True, False)), 20),
}


def run_pair(pkg, orc, name, precision, sfrac=False, store_uref=0.0):
    Q, nx, ny, nz, kw, steps = PLAIN_CASES[name]
    g, o = plain_pair(pkg, orc, Q, nx, ny, nz, precision=precision, store_uref=store_uref, **kw)
    s = None
    if sfrac:
        shape = (nz, ny, nx)
        rng = np.random.default_rng(7)
        s = np.zeros(shape)
        idx = tuple(rng.integers(0, n, 12) for n in shape)
        s[idx] = rng.uniform(0.0, 1.0, 12)
        if precision != pkg.lbm.FP64:
            s = s.astype(np.float32).astype(np.float64)
    gs = pkg.Solver(g)
    if s is not None:
        gs.set_solid_fraction(s)
    gs.initialize()
    os_ = orc.PlainState(o, sfrac=s)
    print(f"[parity] {name} prec={precision} sfrac={sfrac} steps={steps}", flush=True)
    gs.step(steps)
    os_.step(steps)
    return gs, os_


@pytest.mark.parametrize("name", sorted(PLAIN_CASES))
def test_plain_fp64_matches_oracle(pkg, orc, name):
    gs, os_ = run_pair(pkg, orc, name, pkg.lbm.FP64, sfrac=name.startswith("d2q9_periodic") or "periodic_yz" in name)
    fg, fo = gs.populations(), os_.populations()
    assert rel_l2(fg, fo) <= 1e-13, (rel_l2(fg, fo), np.abs(fg - fo).max())
    for a, b in zip(gs.fields(), os_.fields()):
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())
    Fg, Fo = gs.forces(), os_.forces()
    assert Fg.shape == Fo.shape and np.array_equal(Fg[:, 0], Fo[:, 0])
    scale = max(np.abs(Fo[:, 1:]).max(), 1e-300)
    assert np.abs(Fg[:, 1:] - Fo[:, 1:]).max() <= 1e-11 * scale


@pytest.mark.parametrize("name", sorted(PLAIN_CASES))
def test_plain_fp32_within_budget(pkg, orc, name):
    gs, os_ = run_pair(pkg, orc, name, pkg.lbm.FP32)
    # FP32 collision and storage vs the FP64 oracle: error ~ steps * 2^-24 relative
    p_g, u_g, v_g, w_g = gs.fields()
    p_o, u_o, v_o, w_o = os_.fields()
    assert rel_l2(u_g, u_o) <= 2e-5, rel_l2(u_g, u_o)
    assert rel_l2(gs.populations(), os_.populations()) <= 2e-5


@pytest.mark.parametrize("store_uref", [0.0, 0.05])
@pytest.mark.parametrize("precision,budget", [("fp16s", 3e-3), ("bf16s", 3e-2)])
@pytest.mark.parametrize("name", ["d2q9_disk_walls_uniform", "d3q27_sphere_walls_parabolic"])
def test_plain_16bit_within_budget(pkg, orc, name, precision, budget, store_uref):
    """16-bit storage against the FP64 oracle; store_uref = u0 stores deviations from the free
    stream's equilibrium (opposite directions then have different offsets, so the nodes next
    to the body take the per-node path instead of the bounce-back copies)."""
    gs, os_ = run_pair(pkg, orc, name, pkg.lbm.PRECISIONS[precision], store_uref=store_uref)
    u_g = gs.fields()[1]
    u_o = os_.fields()[1]
    err = rel_l2(u_g, u_o)
    print(f"[parity] {name} {precision} velocity rel-L2 {err:.3e}")
    assert err <= budget


REF_CASES = {
    # reference geometry formula (tiny body) and enlarged bodies; (Q, nx, ny, nz, body kwargs, u0, tau0, steps)
    "ref2d_default_disk": (9, 64, 32, 1, None, 0.1, 0.036, 60),
    "ref2d_disk": (9, 80, 36, 1, dict(center=(22.0, 17.5, 0.0), r2=20.0), 0.1, 0.036, 60),
    "ref3d_default_sphere": (27, 40, 16, 14, None, 0.1, 0.024, 30),
    "ref3d_sphere": (27, 36, 18, 16, dict(center=(10.0, 8.0, 7.0), r2=9.0), 0.1, 0.024, 30),
    # population buffers above 256 MB (359 MB / 303 MB): the fused kernel's non-temporal stores
    "ref3d_large_sphere": (27, 128, 96, 96, dict(center=(30.0, 47.5, 47.5), r2=64.0), 0.1, 0.024, 3),
    "ref2d_large_disk": (9, 2048, 2048, 1, dict(center=(500.0, 1023.5, 0.0), r2=400.0), 0.1, 0.036, 3),
}


@pytest.mark.parametrize("name", sorted(REF_CASES))
def test_reference_mode_matches_oracle(pkg, orc, name):
    L = pkg.lbm
    Q, nx, ny, nz, body, u0, tau0, steps = REF_CASES[name]
    if body is None:  # the reference's own body formula
        if Q == 27:
            c = float((nx - 1) // 8)
            body = dict(center=(c, c, c), r2=((nx - 1) / 64.0) ** 2)
        else:
            r = 0.03125 * (nx - 1) / 2.0
            body = dict(center=(0.5 * (nx - 1), 0.5 * (ny - 1), 0.0), r2=r * r)
    g = L.make_config(Q, nx, ny, nz, tau0=tau0, u0=u0, precision=L.FP64, bc_mode=L.BC_REFERENCE,
                      body=L.BODY_SPHERE, force_interval=1, **body)
    o = orc.make_cfg(Q, nx, ny, nz, tau0=tau0, u0=u0, body=orc.BODY_SPHERE, force_interval=1, **body)
    print(f"[parity] {name} reference mode, {steps} steps", flush=True)
    gs = pkg.Solver(g).initialize()
    rs = orc.RefState(o)
    gs.step(steps)
    rs.step(steps)
    for a, b in zip(gs.fields(), rs.fields()):
        assert rel_l2(a, b) <= 1e-12 or np.abs(b).max() == 0.0
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())
    assert rel_l2(gs.populations(), rs.populations()) <= 1e-13
    Fg, Fo = gs.forces(), rs.forces()
    assert Fg.shape == Fo.shape
    assert np.abs(Fg[:, 1:] - Fo[:, 1:]).max() <= 1e-11 * max(np.abs(Fo[:, 1:]).max(), 1e-300)


@pytest.mark.parametrize("name,warm", [("ref2d_disk", 4), ("ref3d_sphere", 4), ("ref3d_sphere", 0)])
def test_staged_step_matches_oracle_stages(pkg, orc, name, warm):
    """lbm_collide / lbm_boundary / lbm_stream / lbm_update against the oracle's stages of the
    reference step: the populations after every stage and the fields after every Update to
    round-off (1e-13), forces to 1e-11; starting right after lbm_initialize or after fused
    steps, and fused steps again after the staged ones."""
    L = pkg.lbm
    Q, nx, ny, nz, body, u0, tau0, _ = REF_CASES[name]
    g = L.make_config(Q, nx, ny, nz, tau0=tau0, u0=u0, precision=L.FP64, bc_mode=L.BC_REFERENCE,
                      body=L.BODY_SPHERE, force_interval=2, **body)
    o = orc.make_cfg(Q, nx, ny, nz, tau0=tau0, u0=u0, body=orc.BODY_SPHERE, force_interval=2, **body)
    print(f"[parity] {name} staged step after {warm} fused steps", flush=True)
    with pkg.Solver(g) as gs:
        gs.initialize()
        rs = orc.RefState(o)
        if warm:
            gs.step(warm)
            rs.step(warm)
        for cycle in range(3):
            for k, call in enumerate((gs.collide, gs.boundary, gs.stream, gs.update)):
                call()
                rs.stage(k)
                a, b = gs.populations(), rs.populations()
                assert rel_l2(a, b) <= 1e-13, (cycle, k)
                assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max()), (cycle, k)
            for a, b in zip(gs.fields(), rs.fields()):
                assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())
        gs.step(3)
        rs.step(3)
        assert rel_l2(gs.populations(), rs.populations()) <= 1e-13
        Fg, Fo = gs.forces(), rs.forces()
    assert Fg.shape == Fo.shape == ((warm + 6) // 2, 4)
    assert np.array_equal(Fg[:, 0], Fo[:, 0])
    assert np.abs(Fg[:, 1:] - Fo[:, 1:]).max() <= 1e-11 * max(np.abs(Fo[:, 1:]).max(), 1e-300)


def test_staged_step_call_order(pkg):
    """Out of order -> LBM_ERR_STATE (4), lbm_step inside an unfinished step -> LBM_ERR_STATE, the
    plain mode -> LBM_ERR_INVALID (1); the step then completes normally."""
    L = pkg.lbm
    g = L.make_config(9, 40, 20, tau0=0.036, u0=0.1, precision=L.FP64, bc_mode=L.BC_REFERENCE,
                      body=L.BODY_SPHERE, center=(10.0, 9.5, 0.0), r2=9.0)
    with pkg.Solver(g) as s:
        s.initialize()
        for bad in (s.boundary, s.stream, s.update):
            with pytest.raises(L.LbmError) as e:
                bad()
            assert e.value.code == 4
        s.collide()
        with pytest.raises(L.LbmError) as e:
            s.step(1)
        assert e.value.code == 4
        with pytest.raises(L.LbmError) as e:
            s.collide()
        assert e.value.code == 4
        s.boundary()
        s.stream()
        s.update()
        s.step(2)
        assert np.isfinite(s.fields()[1]).all()
    p = L.make_config(9, 40, 20, tau0=0.036, u0=0.1, precision=L.FP64, bc_mode=L.BC_PLAIN)
    with pkg.Solver(p) as s:
        s.initialize()
        with pytest.raises(L.LbmError) as e:
            s.collide()
        assert e.value.code == 1


def test_reference_mode_injected_ran(pkg, orc):
    """An injected ran field (in place of the Kokkos pool draw) reaches both sides identically."""
    L = pkg.lbm
    nx, ny = 50, 24
    g = L.make_config(9, nx, ny, tau0=0.036, u0=0.1, precision=L.FP64, bc_mode=L.BC_REFERENCE,
                      body=L.BODY_SPHERE, center=(14.0, 11.5, 0.0), r2=12.0, force_interval=5)
    o = orc.make_cfg(9, nx, ny, tau0=0.036, u0=0.1, body=orc.BODY_SPHERE, center=(14.0, 11.5, 0.0), r2=12.0,
                     force_interval=5)
    ran = np.random.default_rng(3).integers(0, 100, (ny + 4) * (nx + 4)).astype(np.int32)
    gs = pkg.Solver(g)
    gs.set_ran(ran)
    gs.initialize()
    rs = orc.RefState(o, ran=ran)
    gs.step(40)
    rs.step(40)
    assert rel_l2(gs.fields()[1], rs.fields()[1]) <= 1e-13


@pytest.mark.parametrize("sx,sy,steps", [(48, 20, 60), (64, 31, 40)])
def test_single_node_2d_variant_matches_oracle(pkg, orc, sx, sy, steps):
    """The single-node 2D variant (LBM(sx, sy) holds sx+1 by sy+1 nodes, its disk at
    (0.25 sx, 0.5 sy) of radius 0.1 sy, perturbation 1 + ran*0.001, wall fills over
    i in [2, lx-2)) against the oracle's restatement with the same injected ran; and it is
    not the MPI variant: the same inputs in LBM_BC_REFERENCE give different fields."""
    L = pkg.lbm
    nx, ny = sx + 1, sy + 1
    body = dict(center=(0.25 * sx - 2.0, 0.5 * sy, 0.0), r2=0.01 * sy * sy)
    ran = np.random.default_rng(sx).integers(0, 100, (ny + 4) * (nx + 4)).astype(np.int32)
    out = {}
    for bc in (L.BC_REFERENCE_SN, L.BC_REFERENCE):
        g = L.make_config(9, nx, ny, tau0=0.036, u0=0.1, precision=L.FP64, bc_mode=bc, body=L.BODY_SPHERE,
                          force_interval=0, **body)
        print(f"[parity] single-node 2D variant {sx}x{sy} bc={bc}, {steps} steps", flush=True)
        with pkg.Solver(g) as gs:
            gs.set_ran(ran)
            gs.initialize()
            gs.step(steps)
            out[bc] = (np.stack(gs.fields()), gs.populations())
    o = orc.make_cfg(9, nx, ny, tau0=0.036, u0=0.1, body=orc.BODY_SPHERE, variant=1, **body)
    rs = orc.RefState(o, ran=ran)
    rs.step(steps)
    fg, pg = out[L.BC_REFERENCE_SN]
    for a, b in zip(fg, rs.fields()):
        assert np.abs(a - b).max() <= 1e-13 * max(1.0, np.abs(b).max())
    assert rel_l2(pg, rs.populations()) <= 1e-13
    assert rel_l2(out[L.BC_REFERENCE][0][1], fg[1]) > 1e-6


@pytest.mark.parametrize("Q,shape,prec,tol", [(9, (1, 20, 40), "fp64", 1e-9), (27, (10, 10, 16), "fp64", 1e-9),
                                              (9, (1, 20, 40), "fp32", 2e-3)])
def test_adjoint_matches_cpu_adjoint(pkg, orc, Q, shape, prec, tol):
    L = pkg.lbm
    nz, ny, nx = shape
    rng = np.random.default_rng(11)
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
    s = np.exp(-((xx - nx / 3) ** 2 + (yy - ny / 2) ** 2 + (zz - nz / 2) ** 2 * (Q == 27)) / 6.0)
    s = np.clip(s + 0.01 * rng.standard_normal(s.shape), 0.0, 1.0)
    if prec == "fp32":
        s = s.astype(np.float32).astype(np.float64)
    mask = np.zeros(shape, np.uint8)
    mask[nz // 2, ny // 4, (2 * nx) // 3] = 1
    n0, n1, K = 8, 30, 7
    g = L.make_config(Q, nx, ny, nz, tau0=0.1, u0=0.05, precision=L.PRECISIONS[prec], body=L.BODY_MASK,
                      perturb=1e-3, inlet=L.INLET_PARABOLIC)
    o = orc.make_cfg(Q, nx, ny, nz, tau0=0.1, u0=0.05, body=orc.BODY_MASK, perturb=1e-3, inlet=orc.INLET_PARABOLIC)
    print(f"[parity] adjoint Q={Q} {shape} {prec}", flush=True)
    gs = pkg.Solver(g)
    gs.set_mask(mask)
    gs.set_solid_fraction(s)
    gs.initialize()
    Jg, dsg, dug, tf, ta = gs.adjoint(n0, n1, 1.0, K)
    Jo, dso, duo = orc.adjoint(o, mask=mask, sfrac=s, n0=n0, n1=n1, scale=1.0, K=K)
    assert abs(Jg - Jo) <= tol * abs(Jo)
    assert rel_l2(dsg, dso) <= tol, rel_l2(dsg, dso)
    assert rel_l2(dug, duo) <= tol, rel_l2(dug, duo)


@pytest.mark.parametrize("prec,V", [("fp32", 1), ("fp32", 2), ("fp16s", 2), ("fp16s", 4), ("bf16s", 2)])
@pytest.mark.parametrize("Q,shape,kw", [
    (27, (14, 20, 64), dict(body=1, center=(20.0, 9.5, 6.5), r2=9.0, inlet=1)),
    (27, (10, 12, 256), dict(body=2, center=(120.0, 6.0, 0.0), r2=8.0, periodic=(False, True, False))),
    (9, (1, 30, 200), dict(body=1, center=(40.0, 14.5, 0.0), r2=25.0)),
    (27, (8, 10, 256), dict(body=3, mask="edges")),
    (9, (1, 16, 128), dict(body=3, mask="edges")),
    (27, (5, 7, 37), dict(body=1, center=(11.0, 3.5, 2.5), r2=4.0)),
    (9, (1, 9, 101), dict(body=1, center=(30.0, 4.0, 0.0), r2=4.0, periodic=(True, False, False))),
    (27, (3, 3, 3), dict()),
])
def test_fast_path_matches_per_node_path(pkg, orc, monkeypatch, prec, V, Q, shape, kw):
    _fast_vs_per_node(pkg, monkeypatch, prec, V, Q, shape, kw, 0.0)


@pytest.mark.parametrize("prec,V", [("fp16s", 2), ("fp16s", 4), ("bf16s", 2)])
@pytest.mark.parametrize("Q,shape,kw", [
    (27, (14, 20, 64), dict(body=1, center=(20.0, 9.5, 6.5), r2=9.0, inlet=1)),
    (9, (1, 16, 128), dict(body=3, mask="edges")),
])
def test_fast_path_matches_per_node_path_with_storage_offsets(pkg, monkeypatch, prec, V, Q, shape, kw):
    _fast_vs_per_node(pkg, monkeypatch, prec, V, Q, shape, kw, 0.05)


def _fast_vs_per_node(pkg, monkeypatch, prec, V, Q, shape, kw, store_uref):
    """The packed path (V nodes per lane) followed by the per-node recomputation of the nodes
    it gets wrong, and the per-node path on every fluid node, give bit-identical populations,
    fields and forces.  The "edges" mask puts solid nodes on the inlet and outlet lanes,
    against the walls and as isolated points, so lanes hold both solid and fluid nodes; the
    shapes include rows whose last lane reaches into the padding and periodic x edges."""
    L = pkg.lbm
    nz, ny, nx = shape
    kw = dict(kw)
    mask = None
    if kw.pop("mask", None) == "edges":
        mask = np.zeros(shape, dtype=np.uint8)
        mask[:, 2:5, 0] = 1                  # inlet column
        mask[:, ny // 2, 1] = 1
        mask[:, 3:6, nx - 1] = 1             # outlet column
        mask[:, ny - 4, nx - 2] = 1
        mask[:, 0, 40:43] = 1                # against the y = 0 wall
        mask[:, ny // 2, 75] = 1             # isolated node, odd x
        mask[:, ny // 2 + 1, 100:104] = 1
    g = L.make_config(Q, nx, ny, nz, tau0=0.05, u0=0.05, precision=L.PRECISIONS[prec], bc_mode=L.BC_PLAIN,
                      perturb=1e-3, force_interval=3, store_uref=store_uref, **kw)
    monkeypatch.setenv("LBM_NODES_PER_LANE", str(V))
    monkeypatch.setenv("LBM_DENSE", "0")     # the packed path, not the dense pass of small domains
    out = []
    for slow in ("0", "1"):
        monkeypatch.setenv("LBM_FORCE_SLOW_PATH", slow)
        print(f"[parity] fast/per-node {prec} V={V} Q={Q} {shape} store_uref={store_uref} slow={slow}", flush=True)
        s = pkg.Solver(g)
        if mask is not None:
            s.set_mask(mask)
        s.initialize()
        s.step(25)
        out.append((s.populations(), s.fields(), s.forces()))
        s.close()
    (fa, ma, Fa), (fb, mb, Fb) = out
    assert np.array_equal(fa, fb)
    for a_, b_ in zip(ma, mb):
        assert np.array_equal(a_, b_)
    assert np.array_equal(Fa, Fb)


PLANES = ("LBM_DIRECTION_PLANES", (None, "1"))


@pytest.mark.parametrize("prec,env,values,periodic", [
    ("fp16s", "LBM_ALIGNED_PULL", (None, "1"), False),
    ("bf16s", "LBM_ALIGNED_PULL", (None, "1"), False),
    ("fp16s", "LBM_ALIGNED_PULL", (None, "1"), True),
    ("fp32", "LBM_FAST_BLOCKS_PER_CU", ("3", "0"), False),
    ("fp16s", "LBM_FAST_BLOCKS_PER_CU", ("5", "0"), False),
    ("fp64", *PLANES, False), ("fp32", *PLANES, False), ("fp16s", *PLANES, False), ("fp16s", *PLANES, True)])
def test_launch_variants_give_identical_results(pkg, monkeypatch, prec, env, values, periodic):
    """The aligned 16-bit pull (aligned loads, lane shifts, LDS hand-over at wave edges) equals
    the unaligned loads bit for bit on a 4-wave row with a cylinder and walls (or periodic y
    and z); the blocks-per-CU cap changes only residency; the populations stored one plane per
    direction instead of row-interleaved change only addresses."""
    L = pkg.lbm
    g = L.make_config(27, 512, 12, 10, tau0=0.05, u0=0.05, precision=L.PRECISIONS[prec], body=L.BODY_ZCYL,
                      center=(100.0, 5.5, 0.0), r2=9.0, perturb=1e-3, force_interval=4,
                      periodic=(False, periodic, periodic))
    monkeypatch.setenv("LBM_DENSE", "0")     # the packed path, not the dense pass of small domains
    out = []
    for v in values:
        if v is None:
            monkeypatch.delenv(env, raising=False)
        else:
            monkeypatch.setenv(env, v)
        print(f"[parity] {prec} {env}={v}", flush=True)
        with L.Solver(g) as s:
            s.initialize()
            s.step(17)
            out.append((s.populations(), s.forces()))
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("prec", ["fp64", "fp32", "fp16s"])
def test_kernel_timed_steps_equal_plain_steps(pkg, prec):
    """lbm_step_kernel_timed (events around each step's bulk kernel) steps the same state as
    lbm_step, and the bulk kernel's summed time lies inside the steps' time."""
    L = pkg.lbm
    g = L.make_config(27, 64, 20, 16, tau0=0.05, u0=0.05, precision=L.PRECISIONS[prec], body=L.BODY_SPHERE,
                      center=(16.0, 9.5, 7.5), r2=9.0, perturb=1e-3, force_interval=4)
    out = []
    for timed in (False, True):
        with L.Solver(g) as s:
            s.initialize()
            if timed:
                t, k = s.step_kernel_timed(12)
                print(f"[parity] {prec} 12 steps {t:.3f} ms, bulk kernel {k:.3f} ms", flush=True)
                assert 0.0 < k < t
            else:
                s.step(12)
            out.append((s.populations(), s.forces()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


@pytest.mark.parametrize("prec,Q,periodic", [("fp32", 27, False), ("fp16s", 27, True), ("bf16s", 9, False),
                                             ("fp32", 9, True)])
def test_dense_mode_equals_list_path(pkg, monkeypatch, prec, Q, periodic):
    """Dense mode (one per-node pass over every node, chosen when more than a quarter of the
    fluid nodes need the per-node path) against the packed pass + bounce-back copies + fix-up
    list on a mask with a noisy solid fraction: populations and fields bit for bit, forces up to
    their summation order (per block instead of per node, 1e-12)."""
    L = pkg.lbm
    shape = (12, 20, 70) if Q == 27 else (1, 40, 90)
    nz, ny, nx = shape
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
    r2 = (xx - 20.0) ** 2 + (yy - ny / 2 + 0.5) ** 2 + ((zz - nz / 2 + 0.5) ** 2 if Q == 27 else 0)
    mask = (r2 <= 9.0).astype(np.uint8)
    sf = np.clip(np.exp(-r2 / 40.0) + 0.01 * np.random.default_rng(7).standard_normal(shape), 0.0, 1.0)
    sf = np.where(mask == 1, 0.0, sf).astype(np.float32).astype(np.float64)
    g = L.make_config(Q, nx, ny, nz, tau0=0.05, u0=0.05, precision=L.PRECISIONS[prec], body=L.BODY_MASK,
                      perturb=1e-3, force_interval=3, inlet=L.INLET_PARABOLIC,
                      periodic=(False, periodic, periodic and Q == 27))
    out = []
    for dense in ("0", "1"):
        monkeypatch.setenv("LBM_DENSE", dense)
        print(f"[parity] dense={dense} {prec} Q={Q} {shape} periodic={periodic}", flush=True)
        with pkg.Solver(g) as s:
            s.set_mask(mask)
            s.set_solid_fraction(sf)
            s.initialize()
            s.step(13)
            out.append((s.populations(), np.stack(s.fields()), s.forces()))
    (pa, fa, Fa), (pb, fb, Fb) = out
    assert np.array_equal(pa, pb)
    assert np.array_equal(fa, fb)
    assert Fa.shape == Fb.shape == (4, 4) and np.array_equal(Fa[:, 0], Fb[:, 0])
    assert np.abs(Fa[:, 1:] - Fb[:, 1:]).max() <= 1e-12 * np.abs(Fa[:, 1:]).max()


@pytest.mark.parametrize("prec,bc,Q", [("fp64", "plain", 27), ("fp32", "plain", 27), ("fp16s", "plain", 27),
                                       ("fp16s", "plain", 9), ("bf16s", "plain", 27), ("fp64", "ref", 27),
                                       ("fp64", "ref", 9)])
def test_fields_on_request_leave_the_state_unchanged(pkg, prec, bc, Q):
    """Steps write only populations; reading the fields reruns the last step from its input with
    the moments on.  Reading them in the middle of a run changes nothing that follows (populations,
    fields and forces bit for bit against an uninterrupted run), and the fields read equal those of
    a run stopped at that step."""
    L = pkg.lbm
    nz = 12 if Q == 27 else 1
    kw = dict(tau0=0.06, u0=0.05, precision=L.PRECISIONS[prec], force_interval=3, body=L.BODY_SPHERE,
              center=(14.0, 8.5, 5.5), r2=9.0)
    if bc == "plain":
        g = L.make_config(Q, 48, 18, nz, bc_mode=L.BC_PLAIN, perturb=1e-3, **kw)
    else:
        g = L.make_config(Q, 48, 18, nz, bc_mode=L.BC_REFERENCE, **kw)
    print(f"[parity] fields on request {prec} {bc} Q={Q}", flush=True)
    runs = []
    for read_at in (7, None):
        s = pkg.Solver(g).initialize()
        mid = None
        if read_at:
            s.step(read_at)
            mid = s.fields()
            s.step(12 - read_at)
        else:
            s.step(12)
        runs.append((s.populations(), s.fields(), s.forces(), mid))
        s.close()
    (fa, ma, Fa, mid), (fb, mb, Fb, _) = runs
    assert np.array_equal(fa, fb) and np.array_equal(Fa, Fb)
    for x, y in zip(ma, mb):
        assert np.array_equal(x, y)
    s = pkg.Solver(g).initialize()
    s.step(7)
    ref7 = s.fields()
    s.close()
    for x, y in zip(mid, ref7):
        assert np.array_equal(x, y)
    assert not np.array_equal(mid[1], ma[1])


@pytest.mark.parametrize("fi", [0, 1, 3, 10])
@pytest.mark.parametrize("case", ["disk", "edges"])
def test_temporal_blocking_equals_one_launch_per_step(pkg, monkeypatch, fi, case):
    """Several D2Q9 FP64 steps per launch (small domains) against one launch per step: the
    populations and fields bit for bit, the force records to rounding (summed per tile instead
    of per row block), for step counts that leave several remainders of the launches (7 steps
    by default), on a disk and on a mask that touches the inlet and outlet columns and the
    walls; with a record every step, one call of 700 steps holds more records than the
    per-tile partials' slots, so they are summed in several batches within the call."""
    _tb_vs_one_launch(pkg, monkeypatch, fi, case)


@pytest.mark.parametrize("tile,depth", [("24", "5"), ("32", "5"), ("24", "3"), ("32", "8"), ("24", "1"), ("32", "10"),
                                        ("27", "7"), ("25", "4")])
def test_temporal_blocking_tile_and_depth_variants(pkg, monkeypatch, tile, depth):
    """The same comparison for the other tile sides and depths the host can select (the 32 x 32
    tile is the default where 24 x 24 tiles would not fit the CUs once): each step computes
    only the rings of the tile that later steps read, so every (tile, depth) pair changes which
    waves run which step."""
    monkeypatch.setenv("LBM_TB_TILE", tile)
    monkeypatch.setenv("LBM_TB_DEPTH", depth)
    _tb_vs_one_launch(pkg, monkeypatch, 3, "edges")


@pytest.mark.parametrize("nx,ny", [(12, 6), (30, 9), (15, 3)])
def test_temporal_blocking_on_domains_smaller_than_a_core(pkg, monkeypatch, nx, ny):
    """Domains smaller than one tile's core: a single tile column holds the inlet and the
    outlet (the inlet tiles' thread order), one or two tile rows hold both walls; a mask node
    next to the inlet, one against a wall and one at the outlet corner.  Several steps per launch
    (forced on) against one launch per step, bit for bit, forces to rounding."""
    L = pkg.lbm
    kw = dict(tau0=0.06, u0=0.05, precision=L.FP64, bc_mode=L.BC_PLAIN, perturb=1e-3, force_interval=3)
    g = L.make_config(9, nx, ny, 1, body=L.BODY_MASK, **kw)
    mask = np.zeros((1, ny, nx), np.uint8)
    mask[0, ny // 2, 1] = 1
    mask[0, 0, nx // 2] = 1
    mask[0, ny - 1, nx - 1] = 1
    out = []
    for tb in ("0", "1"):
        monkeypatch.setenv("LBM_TEMPORAL_BLOCKING", tb)
        print(f"[parity] temporal blocking {tb} on {nx}x{ny}", flush=True)
        s = pkg.Solver(g)
        s.set_mask(mask)
        s.initialize()
        res = []
        for n in (2, 9, 16):
            s.step(n)
            res.append((s.populations(), s.fields()))
        out.append((res, s.forces()))
        s.close()
    (ra, Fa), (rb, Fb) = out
    for (fa, ma), (fb, mb) in zip(ra, rb):
        assert np.array_equal(fa, fb)
        for x, y in zip(ma, mb):
            assert np.array_equal(x, y)
    assert Fa.shape == Fb.shape and np.array_equal(Fa[:, 0], Fb[:, 0])
    assert np.abs(Fa[:, 1:] - Fb[:, 1:]).max() <= 1e-13 * max(np.abs(Fa[:, 1:]).max(), 1e-300)


def _tb_vs_one_launch(pkg, monkeypatch, fi, case):
    L = pkg.lbm
    nx, ny = 100, 40
    kw = dict(tau0=0.06, u0=0.05, precision=L.FP64, bc_mode=L.BC_PLAIN, perturb=1e-3, force_interval=fi)
    mask = None
    if case == "disk":
        g = L.make_config(9, nx, ny, 1, body=L.BODY_SPHERE, center=(25.0, 19.5, 0.0), r2=25.0, **kw)
    else:
        g = L.make_config(9, nx, ny, 1, body=L.BODY_MASK, **kw)
        mask = np.zeros((1, ny, nx), np.uint8)
        mask[0, 2:5, 0] = 1
        mask[0, 20, 1] = 1
        mask[0, 3:6, nx - 1] = 1
        mask[0, ny - 4, nx - 2] = 1
        mask[0, 0, 40:43] = 1
        mask[0, 20, 75] = 1
        mask[0, 21, 50:54] = 1
    out = []
    for tb in ("0", "1"):
        monkeypatch.setenv("LBM_TEMPORAL_BLOCKING", tb)
        print(f"[parity] temporal blocking {tb} case={case} fi={fi} tile={os.environ.get('LBM_TB_TILE')} "
              f"depth={os.environ.get('LBM_TB_DEPTH')}", flush=True)
        s = pkg.Solver(g)
        if mask is not None:
            s.set_mask(mask)
        s.initialize()
        res = []
        for n in (1, 2, 7, 23, 11, 700 if fi == 1 else 3):
            s.step(n)
            res.append((s.populations(), s.fields()))
        out.append((res, s.forces()))
        s.close()
    (ra, Fa), (rb, Fb) = out
    for (fa, ma), (fb, mb) in zip(ra, rb):
        assert np.array_equal(fa, fb)
        for x, y in zip(ma, mb):
            assert np.array_equal(x, y)
    assert Fa.shape == Fb.shape and np.array_equal(Fa[:, 0], Fb[:, 0])
    if fi:
        assert np.abs(Fa[:, 1:] - Fb[:, 1:]).max() <= 1e-13 * np.abs(Fa[:, 1:]).max()
