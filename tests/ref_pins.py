"""Independent checks of the oracle's reference mode (helpers for test_oracle_ref_pins.py).

Each check follows from the reference's own formulas, not from the oracle's restatement of
them, and compares against something the oracle does not compute the same way:

  body links   usrbc sets f_i(X) <- f_bb(i)(X + 2 e_i) at a solid X with a fluid X + e_i
               (3D MPI+KOKKOS+MPIIO/lbm.cpp:350-356, MPI Kokkos/lbm.cpp:347-353), so after
               stream1/stream2 (lbm.cpp:359-370) a fluid Y whose upstream node Y - e_i is
               solid holds f_i(Y) = f_bb(i)(Y), bit for bit, whenever Y + e_i is not solid
               (its f_bb(i) is then untouched by usrbc);
  wall rows    the mirror fills f(ii, i, l_s - 1, k) = f(bb(ii), i + 2 e_x, l_s + 1, k + 2 e_z)
               (lbm.cpp:278-322, MPI Kokkos/lbm.cpp:331-345) give the same identity on the
               first and last interior rows of each wall, away from the edges where a later
               fill overwrites the ghost slot;
  inlet        the ghost copy f(ii, g-1, j, k) = f(ii, g, j, k) for e_x > 0 (lbm.cpp:323-333,
               MPI Kokkos/lbm.cpp:315-321) makes the first two interior columns agree in
               every e_x > 0 direction;
  outlet       the lateral extrapolation f(ii, l_e, j, k) = f(ii, l_e + e_x, j + e_y, k + e_z)
               for e_x < 0 (lbm.cpp:334-345, MPI Kokkos/lbm.cpp:323-329) gives
               f_i(nx-1, y, z) = f_i(nx-2, y + e_y, z + e_z);
  bulk         Collision (relaxation 1/(tau0 + 0.5), lbm.cpp:257-272), stream and Update
               (lbm.cpp:374-417) act away from every boundary exactly as the plain-mode step,
               whose streaming is pinned by an independent numpy step and whose relaxation
               rate is pinned by the Taylor-Green decay;
  momentum     at step 1 (f = feq of the initial macros; the first Collision changes nothing)
               the exchange (f_i(Y) + f_bb(i)(Y + e_i)) e_i summed over fluid Y with solid
               Y + e_i (MPI Kokkos/lbm.cpp:364-392) equals sum [feq_i(Y) + feq_i(Y - e_i)] e_i,
               computed here in numpy from the textbook equilibrium, and Fy = 0 for a disk on
               the channel axis (symmetric initial state).
Every function returns plain numbers so that test_oracle_ref_pins.py can also run them in a
subprocess against a deliberately broken oracle build.
"""
import numpy as np


def _solid(cfg, Q):
    nz = cfg.nz if Q == 27 else 1
    zz, yy, xx = np.mgrid[0:nz, 0:cfg.ny, 0:cfg.nx]
    r = (xx - cfg.cx) ** 2 + (yy - cfg.cy) ** 2
    if Q == 27 and cfg.body == 1:
        r = r + (zz - cfg.cz) ** 2
    return r <= cfg.r2 if cfg.body in (1, 2) else np.zeros(xx.shape, bool)


def body_cfg(O, Q, force_interval=0):
    if Q == 9:   # disk of radius 4 on the channel axis (MPI 2D variant)
        return O.make_cfg(9, 40, 20, 1, tau0=0.06, u0=0.1, body=O.BODY_SPHERE, center=(12.0, 9.5, 0.0), r2=16.0,
                          force_interval=force_interval)
    return O.make_cfg(27, 26, 16, 14, tau0=0.06, u0=0.1, body=O.BODY_SPHERE, center=(9.0, 7.5, 6.5), r2=12.0,
                      force_interval=force_interval)


def _shift(a, d):
    """b[Y] = a[Y - d] (d = (dz, dy, dx)), False outside the box."""
    out = np.zeros_like(a)
    src = tuple(slice(max(0, -k), n - max(0, k)) for k, n in zip(d, a.shape))
    dst = tuple(slice(max(0, k), n - max(0, -k)) for k, n in zip(d, a.shape))
    out[dst] = a[src]
    return out


def body_link_mirror(O, Q, steps=8):
    """(links checked, links where f_i(Y) != f_bb(i)(Y))."""
    cfg = body_cfg(O, Q)
    r = O.RefState(cfg)
    r.step(steps)
    f = r.populations()
    solid = _solid(cfg, Q)
    e, _, bb = O.lattice(Q)
    n = bad = 0
    for i in range(1, Q):
        d = (int(e[i, 2]), int(e[i, 1]), int(e[i, 0]))
        up_solid = _shift(solid, d)                               # Y - e_i solid
        down_solid = _shift(solid, tuple(-k for k in d))          # Y + e_i solid
        sel = ~solid & up_solid & ~down_solid
        n += int(sel.sum())
        bad += int((f[i][sel] != f[bb[i]][sel]).sum())
    return n, bad


def wall_mirror(O, Q, steps=7):
    """(pairs checked, pairs that differ) on the first/last interior rows of the y (and z) walls,
    away from the edges."""
    cfg = body_cfg(O, Q)
    r = O.RefState(cfg)
    r.step(steps)
    f = r.populations()
    e, _, bb = O.lattice(Q)
    nz, ny, nx = f.shape[1:]
    n = bad = 0
    xs = slice(1, nx - 1)
    zs = slice(1, nz - 1) if Q == 27 else slice(0, 1)
    for i in range(1, Q):
        for axis, comp in ((1, 1), (0, 2)):
            if Q == 9 and axis == 0:
                continue
            c = int(e[i, comp])
            if c == 0:
                continue
            row = 0 if c > 0 else (ny - 1 if axis == 1 else nz - 1)
            if axis == 1:
                a, b = f[i][zs, row, xs], f[bb[i]][zs, row, xs]
            else:
                a, b = f[i][row, 1:ny - 1, xs], f[bb[i]][row, 1:ny - 1, xs]
            n += a.size
            bad += int((a != b).sum())
    return n, bad


def inlet_outlet_copies(O, Q, steps=7):
    """(values checked, values that differ): e_x > 0 directions agree between x = 0 and x = 1;
    e_x < 0 directions satisfy f_i(nx-1, y, z) = f_i(nx-2, y + e_y, z + e_z)."""
    cfg = body_cfg(O, Q)
    r = O.RefState(cfg)
    r.step(steps)
    f = r.populations()
    e, _, _ = O.lattice(Q)
    nz, ny, nx = f.shape[1:]
    zs = slice(1, nz - 1) if Q == 27 else slice(0, 1)
    n = bad = 0
    for i in range(1, Q):
        ex, ey, ez = (int(v) for v in e[i])
        if ex > 0:
            a, b = f[i][zs, 1:ny - 1, 0], f[i][zs, 1:ny - 1, 1]
        elif ex < 0:
            z0, z1 = (1, nz - 1) if Q == 27 else (0, 1)
            a = f[i][z0:z1, 1:ny - 1, nx - 1]
            b = f[i][z0 + ez:z1 + ez, 1 + ey:ny - 1 + ey, nx - 2]
        else:
            continue
        n += a.size
        bad += int((a != b).sum())
    return n, bad


def bulk_vs_plain(O, Q):
    """Max relative difference, away from every boundary, between one reference-mode step and
    one plain-mode step from the same pre-collision state (populations, then the moments)."""
    nz, ny, nx = (1, 14, 22) if Q == 9 else (12, 11, 13)
    tau0 = 0.07
    kw = dict(tau0=tau0, u0=0.05)
    plain = O.PlainState(O.make_cfg(Q, nx, ny, nz, **kw))
    rng = np.random.default_rng(21)
    h0 = plain.populations() + 2e-3 * rng.standard_normal((Q, nz, ny, nx))
    plain.set_populations(h0)
    plain.step(1)
    g1 = plain.last_g()
    ref = O.RefState(O.make_cfg(Q, nx, ny, nz, **kw))
    raw = ref.raw().reshape(ref.lz, ref.ly, ref.lx, Q)
    gz = ref.ghost if Q == 27 else 0
    g = ref.ghost
    raw[gz:gz + nz, g:g + ny, g:g + nx, :] = np.moveaxis(g1, 0, -1)
    ref.set_raw(raw.ravel())
    O.lib().orc_ref_update(ref.h)   # the macros Collision uses, from the same state
    ref.step(1)
    plain.step(1)
    g2 = plain.last_g()
    f1 = ref.populations()
    m = 3
    core = (slice(m, nz - m) if Q == 27 else slice(0, 1), slice(m, ny - m), slice(m, nx - m))
    a, b = f1[(slice(None),) + core], g2[(slice(None),) + core]
    dpop = float(np.abs(a - b).max() / np.abs(b).max())
    dmac = 0.0
    for x, y in zip(ref.fields()[:3], plain.fields()[:3]):
        dmac = max(dmac, float(np.abs(x[core] - y[core]).max() / max(np.abs(y[core]).max(), 1e-300)))
    return dpop, dmac, int(a.size)


def feq_numpy(e, w, p, u):
    """Textbook shifted incompressible equilibrium w_i (3p + 3 e.u + 4.5 (e.u)^2 - 1.5 u.u);
    u: (3, ...) array."""
    eu = np.tensordot(e.astype(float), u, axes=(1, 0))
    usq = (u ** 2).sum(0)
    return w.reshape((-1,) + (1,) * p.ndim) * (3.0 * p + 3.0 * eu + 4.5 * eu ** 2 - 1.5 * usq)


def momentum_exchange_step1(O, Q):
    """(F from the oracle at step 1, F from the closed form, number of links)."""
    cfg = body_cfg(O, Q, force_interval=1)
    nz, ny, nx = (cfg.nz if Q == 27 else 1), cfg.ny, cfg.nx
    ref0 = O.RefState(cfg)
    ran = np.zeros(ref0.lx * ref0.ly * ref0.lz, np.int32)
    r = O.RefState(cfg, ran=ran)
    r.step(1)
    F = r.forces()[0, 1:]
    solid = _solid(cfg, Q)
    e, w, _ = O.lattice(Q)
    # initial macros with ran = 0: parabola in z (3D, lbm.cpp:239) or 0.9 x parabola in y
    # (MPI 2D, MPI Kokkos/lbm.cpp:109), zero in the body, p = v = w = 0
    n = nz if Q == 27 else ny
    c = np.arange(n)
    prof = cfg.u0 * 4.0 * c * (n - 1 - c) / (n - 1) ** 2
    ux = np.broadcast_to(prof[:, None, None] if Q == 27 else 0.9 * prof[None, :, None], (nz, ny, nx)).copy()
    ux[solid] = 0.0
    u = np.stack([ux, np.zeros_like(ux), np.zeros_like(ux)])
    feq = feq_numpy(e, w, np.zeros_like(ux), u)
    Fx = np.zeros(3)
    links = 0
    for i in range(1, Q):
        d = (int(e[i, 2]), int(e[i, 1]), int(e[i, 0]))
        down_solid = _shift(solid, tuple(-k for k in d))      # Y + e_i solid
        sel = ~solid & down_solid
        up = _shift(feq[i], d)                                # feq_i(Y - e_i)
        assert not (sel & _shift(solid, d)).any()             # no one-node fluid gaps
        m = (feq[i] + up)[sel].sum()
        Fx += m * e[i].astype(float)
        links += int(sel.sum())
    return F, Fx, links
