"""Pins of the CPU FP64 oracle (no GPU).

The reference ships no tests, fixtures or golden vectors, so each part of the oracle is
pinned by what the mathematics fixes or by an independent restatement:
  lattice tables        isotropy identities, opposites, the reference's listed entries
  equilibrium           its first three moments (closed form of the incompressible model)
  reference fills       provenance replay reproducing the counts of an independent analysis
                        of the reference's fill order (172 lagged / 228 never-refreshed pulls
                        on a 10x8x8 D3Q27 box; none for the 2D MPI variant)
  reference step        moment/override rules of Update, first Collision is the identity
  plain streaming+BGK   an independent numpy pull implementation on periodic boxes
  walls + Zou-He        exact inlet momentum; Poiseuille with the wall half a node outside
  collision/viscosity   Taylor-Green decay rate exp(-2 nu k^2 t)
  momentum exchange     total-momentum balance in a closed periodic box (exact), symmetry
  adjoint               central finite differences
"""
import numpy as np
import pytest

from conftest import rel_l2


# ---------------------------------------------------------------------------------- lattice
@pytest.mark.parametrize("Q", [9, 27])
def test_lattice_isotropy_and_opposites(orc, Q):
    e, w, bb = orc.lattice(Q)
    D = 2 if Q == 9 else 3
    assert abs(w.sum() - 1.0) < 1e-15
    I = np.eye(3)[:D, :D]
    M2 = np.einsum("q,qa,qb->ab", w, e[:, :D], e[:, :D])
    assert np.allclose(M2, I / 3.0, atol=1e-15)
    M4 = np.einsum("q,qa,qb,qc,qd->abcd", w, *(e[:, :D],) * 4)
    iso = (np.einsum("ab,cd->abcd", I, I) + np.einsum("ac,bd->abcd", I, I) + np.einsum("ad,bc->abcd", I, I)) / 9.0
    assert np.allclose(M4, iso, atol=1e-15)
    assert np.all(e[bb] == -e) and np.all(bb[bb] == np.arange(Q))
    assert len({tuple(v) for v in e}) == Q


def test_lattice_reference_ordering(orc):
    e9, w9, bb9 = orc.lattice(9)
    # MPI Kokkos/lbm.cpp:61-90: opposites 1<->3, 2<->4, 5<->7, 6<->8; e5=(1,1), e6=(-1,1), e8=(1,-1)
    assert list(bb9) == [0, 3, 4, 1, 2, 7, 8, 5, 6]
    assert tuple(e9[5, :2]) == (1, 1) and tuple(e9[6, :2]) == (-1, 1) and tuple(e9[8, :2]) == (1, -1)
    e27, w27, bb27 = orc.lattice(27)
    # 3D MPI+KOKKOS+MPIIO/lbm.cpp:82-217: pairs (2k-1, 2k); e9=(1,-1,0), e13=(1,0,-1), e17=(0,1,-1),
    # e21=(1,-1,1), e25=(1,-1,-1); weights 8/27, 2/27 (1-6), 1/54 (7-18), 1/216 (19-26)
    assert all(bb27[2 * k - 1] == 2 * k for k in range(1, 14))
    for i, v in {9: (1, -1, 0), 13: (1, 0, -1), 17: (0, 1, -1), 21: (1, -1, 1), 25: (1, -1, -1)}.items():
        assert tuple(e27[i]) == v
    assert np.allclose(w27 * 216, [64] + [16] * 6 + [4] * 12 + [1] * 8)


# ------------------------------------------------------------------------------ equilibrium
@pytest.mark.parametrize("Q", [9, 27])
def test_equilibrium_moments(orc, Q):
    e, w, _ = orc.lattice(Q)
    D = 2 if Q == 9 else 3
    rng = np.random.default_rng(1)
    for _ in range(5):
        p = rng.uniform(-0.01, 0.01)
        u = rng.uniform(-0.1, 0.1, 3)
        if D == 2:
            u[2] = 0.0
        f = orc.feq(Q, p, *u)
        assert abs(f.sum() - 3 * p) < 1e-15
        assert np.allclose(e.T @ f, u, atol=1e-16)
        M2 = np.einsum("q,qa,qb->ab", f, e, e)[:D, :D]
        assert np.allclose(M2, p * np.eye(D) + np.outer(u[:D], u[:D]), atol=1e-16)


# ------------------------------------------------------------------ reference fill order
def _fill_provenance(orc, Q, nx, ny, nz, gens=3, variant=0, where=False):
    r = orc.RefState(orc.make_cfg(Q, nx, ny, nz, tau0=0.1, u0=0.1, variant=variant))
    lx, ly, lz, g = r.lx, r.ly, r.lz, r.ghost
    e, _, _ = orc.lattice(Q)
    ids = np.arange(lx * ly * lz * Q, dtype=np.float64)
    f = ids.copy()
    B = 1e7
    kr = np.arange(g, lz - g) if Q == 27 else np.arange(1)
    K, J, I = np.meshgrid(kr, np.arange(g, ly - g), np.arange(g, lx - g), indexing="ij")
    node = I + lx * (J + ly * K)
    interior = (Q * node[..., None] + np.arange(Q)).ravel()
    counts = None
    for gen in range(1, gens + 1):
        f[interior] = gen * B + ids[interior]
        r.set_raw(f)
        orc.lib().orc_ref_boundary_fills(r.h)
        f = r.raw()
        src = (Q * (node[..., None] - (e[:, 0] + lx * (e[:, 1] + ly * e[:, 2])))) + np.arange(Q)
        gv = (f[src.ravel()] // B).astype(int)
        counts = dict(current=int((gv == gen).sum()), lagged=int((gv == gen - 1).sum()),
                      initial=int(((gv == 0) & (gen > 1)).sum()))
    if where:   # lagged pulls as (i, j, direction) in padded coordinates
        lag = np.argwhere(gv.reshape(node.shape + (Q,)) == gens - 1)
        counts["lagged_at"] = sorted((int(I[tuple(t[:-1])]), int(J[tuple(t[:-1])]), int(t[-1])) for t in lag)
    return counts


def test_reference_3d_fill_order_provenance(orc):
    c = _fill_provenance(orc, 27, 10, 8, 8)
    assert c == dict(current=16880, lagged=172, initial=228)


def test_reference_2d_fill_order_has_no_stale_pulls(orc):
    c = _fill_provenance(orc, 9, 12, 10, 1)
    assert c["lagged"] == 0 and c["initial"] == 0


def test_single_node_2d_variant_lags_two_inlet_corner_pulls(orc):
    """The single-node 2D variant fills the bottom/top walls over i in [2, lx-2) only and runs
    the inlet copy first (Kokkos/lbm.cpp:147-167), so exactly two pulls per step read the
    previous step's wall fill: node (2, 2) direction 5 and node (2, ly-3) direction 8 (an
    independent analysis of that fill order finds the same two)."""
    nx, ny = 12, 10
    c = _fill_provenance(orc, 9, nx, ny, 1, variant=1, where=True)
    ly = ny + 4
    assert c["lagged"] == 2 and c["initial"] == 0
    assert c["lagged_at"] == [(2, 2, 5), (2, ly - 3, 8)]


def test_single_node_variant_initial_velocity(orc):
    """Initial u of the 2D variants (injected ran): the parabola times 0.9 + 0.002 ran (MPI
    variant, MPI Kokkos/lbm.cpp:109) or 1 + 0.001 ran (single-node variant, Kokkos/lbm.cpp:123),
    zero in the body."""
    nx, ny = 21, 9
    ran = np.random.default_rng(1).integers(0, 100, (ny + 4) * (nx + 4)).astype(np.int32)
    y = np.arange(ny)[:, None]
    prof = 0.1 * 4.0 * y * (ny - 1 - y) / (ny - 1) ** 2
    r = ran.reshape(ny + 4, nx + 4)[2:-2, 2:-2]
    for variant, factor in ((0, 0.9 + 0.002 * r), (1, 1.0 + 0.001 * r)):
        s = orc.RefState(orc.make_cfg(9, nx, ny, tau0=0.036, u0=0.1, body=orc.BODY_SPHERE, center=(6.0, 4.0, 0.0),
                                      r2=2.0, variant=variant), ran=ran)
        u = s.fields()[1][0]
        xx, yy = np.meshgrid(np.arange(nx), np.arange(ny))
        solid = (xx - 6.0) ** 2 + (yy - 4.0) ** 2 <= 2.0
        assert np.allclose(u, np.where(solid, 0.0, prof * factor), rtol=1e-14, atol=1e-17)


def test_reference_first_collision_is_identity(orc):
    for Q, shp in ((9, (1, 14, 30)), (27, (8, 10, 14))):
        nz, ny, nx = shp
        r = orc.RefState(orc.make_cfg(Q, nx, ny, nz, tau0=0.03, u0=0.1, body=orc.BODY_SPHERE,
                                      center=(5.0, 4.0, 3.0), r2=4.0))
        before = r.raw()
        orc.lib().orc_ref_collision(r.h)
        assert np.array_equal(r.raw(), before)


def test_reference_stream_is_pull(orc):
    Q, nx, ny, nz = 27, 7, 6, 5
    r = orc.RefState(orc.make_cfg(Q, nx, ny, nz, tau0=0.1, u0=0.1))
    lx, ly, g = r.lx, r.ly, r.ghost
    tags = np.arange(r.lx * r.ly * r.lz * Q, dtype=np.float64)
    r.set_raw(tags)
    orc.lib().orc_ref_stream(r.h)
    out = r.populations()
    e, _, _ = orc.lattice(Q)
    for ii in (0, 1, 8, 19, 26):
        for (k, j, i) in ((0, 0, 0), (2, 3, 4), (nz - 1, ny - 1, nx - 1)):
            src = (i + g - e[ii, 0]) + lx * ((j + g - e[ii, 1]) + ly * (k + g - e[ii, 2]))
            assert out[ii, k, j, i] == tags[ii + Q * src]


@pytest.mark.parametrize("Q", [9, 27])
def test_reference_update_overrides(orc, Q):
    nx, ny, nz = (30, 14, 1) if Q == 9 else (16, 10, 12)
    r = orc.RefState(orc.make_cfg(Q, nx, ny, nz, tau0=0.03, u0=0.1, body=orc.BODY_SPHERE,
                                  center=(8.0, 5.0, 6.0), r2=4.0, force_interval=1))
    r.step(3)
    p, u, v, w = r.fields()
    n = ny if Q == 9 else nz
    c = np.arange(n)
    prof = 0.1 * 4.0 * c * (n - 1 - c) / (n - 1) ** 2
    inlet = u[:, :, 0]
    expect = np.broadcast_to(prof[None, :] if Q == 9 else prof[:, None], inlet.shape)
    assert np.array_equal(inlet, expect)                 # lbm.cpp:403-408
    assert np.all(v[:, :, 0] == 0) and np.all(p[:, :, -1] == 0)   # lbm.cpp:409-412
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
    solid = (xx - 8.0) ** 2 + (yy - 5.0) ** 2 + ((zz - 6.0) ** 2 if Q == 27 else 0) <= 4.0
    assert np.all(u[solid] == 0) and np.all(p[solid] == 0)
    # the reference's Cd/Cl record of every step (MPI Kokkos/lbm.cpp:398-407)
    F = r.forces()
    assert F.shape == (3, 4) and np.array_equal(F[:, 0], [1, 2, 3]) and np.all(np.isfinite(F))


@pytest.mark.parametrize("Q,variant", [(9, 0), (9, 1), (27, 0)])
def test_reference_stages_compose_to_the_step(orc, Q, variant):
    """The four stages of the staged C ABI (Collision; fills + usrbc + momentum exchange;
    stream1 + stream2; Update), called in the reference main's order, reproduce orc_ref_step
    bit for bit, forces included; a stage order with stream before the fills does not."""
    nx, ny, nz = (30, 14, 1) if Q == 9 else (16, 10, 12)
    kw = dict(tau0=0.03, u0=0.1, body=orc.BODY_SPHERE, center=(8.0, 5.0, 6.0), r2=4.0, force_interval=2)
    if variant:
        kw["variant"] = 1
    a = orc.RefState(orc.make_cfg(Q, nx, ny, nz, **kw))
    b = orc.RefState(orc.make_cfg(Q, nx, ny, nz, **kw))
    a.step(5)
    for _ in range(5):
        for k in range(4):
            b.stage(k)
    assert np.array_equal(a.raw(), b.raw())
    for x, y in zip(a.fields(), b.fields()):
        assert np.array_equal(x, y)
    assert np.array_equal(a.forces(), b.forces()) and a.forces().shape == (2, 4)
    c = orc.RefState(orc.make_cfg(Q, nx, ny, nz, **kw))
    for _ in range(5):
        for k in (0, 2, 1, 3):
            c.stage(k)
    assert not np.array_equal(a.populations(), c.populations())


# ------------------------------------------------------- plain mode vs an independent numpy
def numpy_periodic_step(f, e, w, tau0, D):
    """Textbook pull streaming on a periodic box + the shifted incompressible BGK."""
    Q = len(w)
    g = np.stack([np.roll(f[i], shift=(e[i, 2], e[i, 1], e[i, 0]), axis=(0, 1, 2)) for i in range(Q)])
    p = g.sum(0) / 3.0
    u = np.einsum("qzyx,qd->dzyx", g, e.astype(float))
    eu = np.einsum("qd,dzyx->qzyx", e.astype(float), u)
    usq = (u ** 2).sum(0)
    feq = w[:, None, None, None] * (3 * p + 3 * eu + 4.5 * eu ** 2 - 1.5 * usq)
    return g - (g - feq) / (tau0 + 0.5)


@pytest.mark.parametrize("Q,shape", [(9, (1, 12, 16)), (27, (6, 7, 8))])
def test_plain_periodic_matches_numpy(orc, Q, shape):
    nz, ny, nx = shape
    cfg = orc.make_cfg(Q, nx, ny, nz, tau0=0.07, u0=0.04, perturb=5e-3, periodic=(True, True, True))
    s = orc.PlainState(cfg)
    e, w, _ = orc.lattice(Q)
    rng = np.random.default_rng(2)
    f = s.populations() + 1e-3 * rng.standard_normal((Q,) + shape)
    s.set_populations(f)
    for _ in range(5):
        f = numpy_periodic_step(f, e, w, 0.07, 2 if Q == 9 else 3)
    s.step(5)
    assert rel_l2(s.populations(), f) < 1e-14


@pytest.mark.parametrize("Q,shape", [(9, (1, 20, 30)), (27, (8, 10, 12))])
def test_zou_he_inlet_momentum_is_exact(orc, Q, shape):
    nz, ny, nx = shape
    s = orc.PlainState(orc.make_cfg(Q, nx, ny, nz, tau0=0.1, u0=0.05, inlet=orc.INLET_PARABOLIC, perturb=1e-3))
    s.step(4)
    g = s.last_g()
    e, _, _ = orc.lattice(Q)
    mom = np.einsum("qzyx,qd->dzyx", g[:, :, :, :1], e.astype(float))[..., 0]
    assert np.abs(mom[0] - s.inlet()).max() < 1e-16
    assert np.abs(mom[1:]).max() < 1e-16
    assert not np.isnan(g).any()


def textbook_regularize(orc, Q, g, U):
    """Latt & Chopard's regularisation from its definition: g_i = feq_i(p, U) +
    w_i / (2 cs^4) (e_i e_i - cs^2 I) : Pi_neq, Pi_neq = sum_k e_k e_k (g_k - feq_k), cs^2 = 1/3,
    p = sum g / 3 (shifted incompressible model).  g: (Q, n) columns of populations."""
    e, w, _ = orc.lattice(Q)
    D = 2 if Q == 9 else 3
    e = e[:, :D].astype(float)
    out = np.empty_like(g)
    for c in range(g.shape[1]):
        p = g[:, c].sum() / 3.0
        fe = orc.feq(Q, p, U[c])
        pi = np.einsum("k,ka,kb->ab", g[:, c] - fe, e, e)
        qi = np.einsum("ia,ib->iab", e, e) - np.eye(D) / 3.0
        out[:, c] = fe + w * 4.5 * np.einsum("iab,ab->i", qi, pi)
    return out


def test_regularized_inlet_matches_textbook_d2q9(orc):
    """Zou & He (1997) D2Q9 velocity inlet (rho u -> u for the incompressible model) followed by
    the textbook regularisation, computed here from the pulled populations of a random state,
    equals the oracle's inlet column after one step."""
    nx, ny, U = 6, 5, 0.05
    s = orc.PlainState(orc.make_cfg(9, nx, ny, 1, tau0=0.1, u0=U, periodic=(False, True, False)))
    rng = np.random.default_rng(11)
    H = np.stack([orc.feq(9, 0.0, U) for _ in range(nx * ny)], 1).reshape(9, 1, ny, nx)
    H = H + 1e-3 * rng.standard_normal(H.shape)
    s.set_populations(H)
    s.step(1)
    out = s.last_g()[:, 0, :, 0]
    e, w, _ = orc.lattice(9)
    g = np.zeros((9, ny))
    for i in range(9):
        ex, ey = e[i, 0], e[i, 1]
        if ex <= 0:   # knowns: pulled from (0 - ex, y - ey), periodic in y
            g[i] = np.roll(H[i, 0, :, -ex], ey)
    # classic numbering 1:(1,0) 2:(0,1) 3:(-1,0) 4:(0,-1) 5:(1,1) 6:(-1,1) 7:(-1,-1) 8:(1,-1)
    assert [tuple(e[i, :2]) for i in range(9)] == [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1),
                                                  (-1, -1), (1, -1)]
    g[1] = g[3] + 2.0 / 3.0 * U
    g[5] = g[7] - 0.5 * (g[2] - g[4]) + U / 6.0
    g[8] = g[6] + 0.5 * (g[2] - g[4]) + U / 6.0
    ref = textbook_regularize(orc, 9, g, np.full(ny, U))
    assert np.abs(out - ref).max() < 1e-15


@pytest.mark.parametrize("inlet", [0, 1])
def test_regularized_inlet_properties_d3q27(orc, inlet):
    """D3Q27: the inlet column carries exactly the prescribed momentum and is a fixed point of
    the textbook regularisation (its non-hydrodynamic moments are those of feq + Pi_neq)."""
    nz, ny, nx = 5, 6, 7
    s = orc.PlainState(orc.make_cfg(27, nx, ny, nz, tau0=0.1, u0=0.05, inlet=inlet, perturb=1e-3))
    rng = np.random.default_rng(3)
    s.set_populations(s.populations() + 1e-3 * rng.standard_normal((27, nz, ny, nx)))
    s.step(1)
    out = s.last_g()[:, :, :, 0].reshape(27, -1)
    U = s.inlet().reshape(-1)
    e, _, _ = orc.lattice(27)
    mom = e.astype(float).T @ out
    assert np.abs(mom[0] - U).max() < 1e-16 and np.abs(mom[1:]).max() < 1e-16
    assert np.abs(textbook_regularize(orc, 27, out, U) - out).max() < 1e-15


@pytest.mark.parametrize("Q,shape,tau0,steps", [(9, (1, 20, 100), 0.03, 2000), (27, (6, 6, 40), 0.018, 1500)])
def test_velocity_inlet_is_stable_near_tau_half(orc, Q, shape, tau0, steps):
    """Uniform flow with a 1e-3 perturbation through the inlet and outlet at tau = 0.53 (D2Q9,
    the Re 100 cylinder workload) and 0.518 (D3Q27, the Re 200 sphere workload) decays.  Plain
    Zou-He (no regularisation) diverges from the inlet column within ~1000 steps here."""
    nz, ny, nx = shape
    per = (False, True, Q == 27)
    s = orc.PlainState(orc.make_cfg(Q, nx, ny, nz, tau0=tau0, u0=0.05, periodic=per, perturb=1e-3))
    s.step(steps)
    u = s.fields()[1]
    assert np.isfinite(u).all()
    assert np.abs(u - 0.05).max() < 1e-4


@pytest.mark.parametrize("Q", [9, 27])
def test_shear_mode_between_walls_decays_with_wall_half_node_outside(orc, Q):
    """u_x = U sin(pi (y + 1/2)/n) between no-slip walls at y = -1/2 and y = n - 1/2 decays as
    exp(-nu (pi/n)^2 t) with nu = tau0/3.  A wall displaced by d changes the rate by ~4d/n, so
    a 1 % match pins the half-way bounce-back wall to within ~0.03 of a spacing."""
    n, T, U = 16, 300, 1e-3
    tau0 = np.sqrt(3.0 / 16.0)
    if Q == 9:
        shape, per = (1, n, 8), (True, False, False)
    else:
        shape, per = (n, 3, 8), (True, True, False)   # walls in z
    nz, ny, nx = shape
    y = np.arange(n) + 0.5
    mode = U * np.sin(np.pi * y / n)
    ux = np.broadcast_to((mode[:, None] if Q == 9 else mode[:, None, None]), (ny, nx) if Q == 9 else shape)
    ux = np.asarray(ux).reshape(shape)
    s = orc.PlainState(orc.make_cfg(Q, nx, ny, nz, tau0=tau0, u0=0.0, periodic=per))
    f = np.stack([orc.feq(Q, 0.0, a) for a in ux.ravel()], 1).reshape((Q,) + shape)
    s.set_populations(f)
    s.step(T)
    u = s.fields()[1]
    decay = np.sum(u * ux) / np.sum(ux * ux)
    assert abs(decay / np.exp(-(tau0 / 3) * (np.pi / n) ** 2 * T) - 1) < 1e-2


@pytest.mark.parametrize("Q,shape,per", [(9, (1, 9, 24), (False, True, False)),
                                         (27, (5, 6, 16), (False, True, True))])
def test_uniform_flow_through_inlet_and_outlet_is_a_fixed_point(orc, Q, shape, per):
    nz, ny, nx = shape
    s = orc.PlainState(orc.make_cfg(Q, nx, ny, nz, tau0=0.05, u0=0.06, periodic=per))
    f0 = s.populations()
    s.step(50)
    assert np.abs(s.populations() - f0).max() < 1e-13 * np.abs(f0).max()


def test_taylor_green_decay(orc):
    # n = 64 keeps the lattice-dispersion correction to the rate (~k^2/12) below 0.1 %
    n = 64
    tau0 = 0.3
    nu = tau0 / 3.0
    U = 0.01
    k = 2 * np.pi / n
    y, x = np.mgrid[0:n, 0:n] + 0.0
    ux = U * np.sin(k * x) * np.cos(k * y)
    uy = -U * np.cos(k * x) * np.sin(k * y)
    cfg = orc.make_cfg(9, n, n, 1, tau0=tau0, u0=0.0, periodic=(True, True, False))
    s = orc.PlainState(cfg)
    f = np.stack([orc.feq(9, 0.0, a, b) for a, b in zip(ux.ravel(), uy.ravel())], 1).reshape(9, 1, n, n)
    s.set_populations(f)
    T = 400
    s.step(T)
    u = s.fields()[1][0]
    decay = np.sum(u * ux) / np.sum(ux * ux)
    assert abs(decay / np.exp(-2 * nu * k * k * T) - 1) < 5e-3


@pytest.mark.parametrize("Q,shape", [(9, (1, 24, 40)), (27, (12, 14, 16))])
def test_momentum_exchange_balances_fluid_momentum(orc, Q, shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(4)
    s = np.clip(rng.uniform(-0.5, 1.0, shape), 0, 1) * (rng.uniform(size=shape) < 0.05)
    cfg = orc.make_cfg(Q, nx, ny, nz, tau0=0.1, u0=0.05, body=orc.BODY_SPHERE, center=(nx / 2, ny / 2, nz / 2),
                       r2=9.0, perturb=1e-3, periodic=(True, True, True), force_interval=1)
    st = orc.PlainState(cfg, sfrac=s)
    e, _, _ = orc.lattice(Q)
    mom = lambda h: np.einsum("qzyx,qd->d", h, e.astype(float))
    h0 = st.populations()
    P0, M0 = mom(h0), h0.sum()
    st.step(6)
    h1 = st.populations()
    F = st.forces()[:, 1:].sum(0)
    assert np.abs(mom(h1) - P0 + F).max() < 1e-12 * max(1.0, np.abs(P0).max())
    assert abs(h1.sum() - M0) < 1e-12


def test_symmetric_body_has_no_lift(orc):
    cfg = orc.make_cfg(9, 60, 30, 1, tau0=0.06, u0=0.05, body=orc.BODY_SPHERE, center=(15.0, 14.5, 0.0), r2=16.0,
                       force_interval=1)
    s = orc.PlainState(cfg)
    s.step(50)
    F = s.forces()
    assert np.abs(F[:, 2]).max() < 1e-14 * np.abs(F[:, 1]).max()
    assert F[-1, 1] > 0


def test_partial_bounce_back_limits(orc):
    shape = (1, 12, 20)
    base = dict(tau0=0.1, u0=0.05, perturb=1e-3, periodic=(True, True, False))
    a = orc.PlainState(orc.make_cfg(9, 20, 12, 1, **base))
    b = orc.PlainState(orc.make_cfg(9, 20, 12, 1, **base), sfrac=np.zeros(shape))
    a.step(5)
    b.step(5)
    assert np.array_equal(a.populations(), b.populations())
    s = np.zeros(shape)
    s[0, 6, 10] = 1.0
    c = orc.PlainState(orc.make_cfg(9, 20, 12, 1, **base), sfrac=s)
    c.step(3)
    g = c.last_g()[:, 0, 6, 10]
    h = c.populations()[:, 0, 6, 10]
    _, _, bb = orc.lattice(9)
    assert np.array_equal(h, g[bb])   # s = 1: every population leaves reversed


# ----------------------------------------------------------------------------- adjoint
@pytest.mark.parametrize("Q,shape", [(9, (1, 12, 24)), (27, (6, 6, 10))])
def test_cpu_adjoint_matches_finite_differences(orc, Q, shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(5)
    zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
    s = np.exp(-((xx - nx / 3) ** 2 + (yy - ny / 2) ** 2 + (Q == 27) * (zz - nz / 2) ** 2) / 5.0)
    s = np.clip(s + 0.01 * rng.standard_normal(shape), 0.0, 1.0)
    mask = np.zeros(shape, np.uint8)
    mask[nz // 2, ny // 4, (2 * nx) // 3] = 1
    cfg = orc.make_cfg(Q, nx, ny, nz, tau0=0.1, u0=0.05, body=orc.BODY_MASK, perturb=1e-3, inlet=orc.INLET_PARABOLIC)
    n0, n1 = 6, 20
    J, ds, du = orc.adjoint(cfg, mask=mask, sfrac=s, n0=n0, n1=n1, scale=1.0, K=6)
    assert abs(J - orc.objective(cfg, mask=mask, sfrac=s, n0=n0, n1=n1)) < 1e-14 * abs(J)
    eps = 1e-6
    idx = [tuple(rng.integers(0, n) for n in shape) for _ in range(4)] + [(nz // 2, ny // 2, nx // 3)]
    for i in idx:
        sp, sm = s.copy(), s.copy()
        sp[i] += eps
        sm[i] -= eps
        fd = (orc.objective(cfg, mask=mask, sfrac=sp, n0=n0, n1=n1) -
              orc.objective(cfg, mask=mask, sfrac=sm, n0=n0, n1=n1)) / (2 * eps)
        assert abs(ds[i] - fd) <= 1e-6 * max(abs(fd), 1e-3 * np.abs(ds).max())
    sh = (nz, ny)
    uin = orc.PlainState(cfg, mask=mask, sfrac=s).inlet()
    for row in ((0, ny // 2), (nz - 1, 1)):
        up, um = uin.copy(), uin.copy()
        up[row] += eps
        um[row] -= eps
        fd = (orc.objective(cfg, mask=mask, sfrac=s, uin=up, n0=n0, n1=n1) -
              orc.objective(cfg, mask=mask, sfrac=s, uin=um, n0=n0, n1=n1)) / (2 * eps)
        J2, ds2, du2 = orc.adjoint(cfg, mask=mask, sfrac=s, uin=uin, n0=n0, n1=n1, K=6)
        assert du2.shape == sh
        assert abs(du2[row] - fd) <= 1e-6 * max(abs(fd), 1e-3)


# -------------------------------------------------------------------- float build (baseline)
@pytest.mark.parametrize("Q,nx,ny,nz", [(9, 80, 24, 1), (27, 40, 16, 12)])
def test_float_build_tracks_the_fp64_oracle(orc, Q, nx, ny, nz):
    """liboracle_f32.so (the CPU baseline's FP32 leg) runs the same plain-mode step in float:
    after 60 steps with walls, inlet, outlet and a body its populations stay within FP32
    round-off of the FP64 oracle's, and differ from them (the arithmetic is float)."""
    import ctypes as C
    L32 = orc.f32_lib()
    cfg = orc.make_cfg(Q, nx, ny, nz, tau0=0.05, u0=0.05, body=orc.BODY_ZCYL if Q == 27 else orc.BODY_SPHERE,
                       center=(12.0, ny / 2 - 0.5, 0.0), r2=16.0, perturb=1e-3, force_interval=10)
    ref = orc.PlainState(cfg)
    ref.step(60)
    h = L32.orc_plain_create(C.byref(cfg), None, None, None, None)
    assert h
    try:
        L32.orc_plain_step(h, 60)
        f32 = np.zeros((Q, nz, ny, nx))
        L32.orc_plain_populations(h, f32.ctypes.data_as(C.c_void_p))
        n = L32.orc_plain_forces(h, None, 0)
        F32 = np.zeros((n, 4))
        L32.orc_plain_forces(h, F32.ctypes.data_as(C.c_void_p), n)
    finally:
        L32.orc_plain_destroy(h)
    f64 = ref.populations()
    e = rel_l2(f32, f64)
    assert 1e-9 < e < 1e-5, e
    F64 = ref.forces()
    assert np.array_equal(F32[:, 0], F64[:, 0])
    assert np.abs(F32[:, 1] - F64[:, 1]).max() < 2e-5 * np.abs(F64[:, 1]).max()   # observed 2.5e-6


@pytest.mark.parametrize("Q,nx,ny,nz", [(9, 60, 24, 1), (27, 24, 16, 12)])
def test_float_build_reference_mode_tracks_the_fp64_oracle(orc, Q, nx, ny, nz):
    """The float build's reference mode (the CPU baseline's FP32 reference-structured leg): 40
    steps with the reference's fills, usrbc and body stay within FP32 round-off of the FP64
    reference mode, and differ from it."""
    import ctypes as C
    L32 = orc.f32_lib()
    cfg = orc.make_cfg(Q, nx, ny, nz, tau0=0.06, u0=0.05, body=orc.BODY_SPHERE,
                       center=(nx / 4, ny / 2 - 0.5, nz / 2 - 0.5 if Q == 27 else 0.0), r2=9.0)
    ran = np.random.default_rng(3).integers(0, 100, size=200000).astype(np.int32)
    ref = orc.RefState(cfg, ran=ran)
    ref.step(40)
    h = L32.orc_ref_create(C.byref(cfg), ran.ctypes.data_as(C.c_void_p))
    assert h
    try:
        L32.orc_ref_step(h, 40)
        f32 = np.zeros((Q, nz, ny, nx))
        L32.orc_ref_populations(h, f32.ctypes.data_as(C.c_void_p))
    finally:
        L32.orc_ref_destroy(h)
    e = rel_l2(f32, ref.populations())
    assert 1e-9 < e < 1e-5, e
