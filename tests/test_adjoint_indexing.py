"""Index arithmetic of the adjoint reverse kernel (k_adj_step), transcribed on the CPU: every
load and store of both its paths stays inside its buffer for every node and direction, on
shapes with walls, periodic axes, a mask and ragged extents (including the 24x10x8 D3Q27 case of
the torch-wrapper test).

The general path gathers mu_{m+1} only after classifying the source (wall -> the node's own
slot, beyond the outlet -> no load).  A gather issued for every direction without that
classification forms the source row y + e_y (or z + e_z) as an unsigned 32-bit value, which at a
wall is 2^32 - 1 + ny*(z + e_z): its element index lies gigabytes past the buffer.  The last test
shows that for the 24x10x8 case, which is why the committed kernel keeps the classification
before its loads (general path) or clamps inside the row (fast path)."""
import numpy as np
import pytest

E27 = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, -1, 0),
       (1, -1, 0), (-1, 1, 0), (1, 0, 1), (-1, 0, -1), (1, 0, -1), (-1, 0, 1), (0, 1, 1), (0, -1, -1), (0, 1, -1),
       (0, -1, 1), (1, 1, 1), (-1, -1, -1), (1, -1, 1), (-1, 1, -1), (1, 1, -1), (-1, -1, 1), (1, -1, -1), (-1, 1, 1)]
E9 = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (-1, 0, 0), (0, -1, 0), (1, 1, 0), (-1, 1, 0), (-1, -1, 0), (1, -1, 0)]
U32 = 1 << 32


def geometry(Q, nx, ny, nz):
    nxp = (nx + 63) // 64 * 64
    plane = nxp * ny * nz
    return dict(nx=nx, ny=ny, nz=nz, nxp=nxp, plane=plane, ds=nxp, rs=Q * nxp, size=Q * plane)


def opp(E, i):
    return E.index(tuple(-v for v in E[i]))


def kernel_indices(Q, nx, ny, nz, per, mask):
    """All element indices k_adj_step loads from mu_{m+1} and H and stores to mu_m, for every
    node (x < nx) and direction, following the kernel's two paths."""
    E = E27 if Q == 27 else E9
    g = geometry(Q, nx, ny, nz)
    xper, yper, zper = per
    solid = mask
    near = np.zeros_like(mask)
    for (ex, ey, ez) in E[1:]:
        for z in range(nz):
            for y in range(ny):
                for x in range(nx):
                    sx, sy, sz = x - ex, y - ey, z - ez
                    if xper: sx %= nx
                    if yper: sy %= ny
                    if zper: sz %= nz
                    if 0 <= sx < nx and 0 <= sy < ny and 0 <= sz < nz and solid[sz, sy, sx]:
                        near[z, y, x] = True
    clean = ~(solid | near).reshape(nz * ny, nx).any(1) if True else None
    idx = []

    def pel(i, x, row):
        return i * g["ds"] + row * g["rs"] + x

    for z in range(nz):
        for y in range(ny):
            row = y + ny * z
            ylo, yhi = y == 0, y == ny - 1
            zlo, zhi = Q == 27 and z == 0, Q == 27 and z == nz - 1
            fast = clean[row] and not (yper and (ylo or yhi)) and not (zper and (zlo or zhi))
            for x in range(nx):
                if not fast and solid[z, y, x]:
                    idx += [pel(i, x, row) for i in range(Q)]          # stores of a solid node
                    continue
                xhi = not xper and x == nx - 1
                for i, (ex, ey, ez) in enumerate(E):
                    o = opp(E, i)
                    if fast:
                        def wall(a, b):
                            return (a > 0 and yhi) or (a < 0 and ylo) or (b > 0 and zhi) or (b < 0 and zlo)
                        wm, wh = wall(ey, ez), wall(-ey, -ez)
                        mrow = (row + ((ey + ny * ez) % U32)) % U32
                        dx = x + ex
                        if ex > 0 and dx == nx: dx = 0 if xper else x
                        if ex < 0 and dx < 0: dx = nx - 1 if xper else x
                        idx.append(pel(o, x, row) if wm else pel(i, dx, mrow))
                        if ex < 0:                                  # the outlet term, loaded by every lane
                            idx.append(pel(o, x, row) if wm else pel(i, x, mrow))
                        hrow = (row - ((ey + ny * ez) % U32)) % U32
                        sx = x - ex
                        if ex > 0 and sx < 0: sx = nx - 1 if xper else 0
                        if ex < 0 and sx == nx: sx = 0 if xper else nx - 1
                        idx.append(pel(o, x, row) if wh else pel(i, sx, hrow))
                    else:
                        # lambda gather, general path
                        if not (ex < 0 and not xper and x == 0):
                            dx, dy, dz, wall, out = x + ex, y + ey, z + ez, False, False
                            if ey and not 0 <= dy < ny:
                                if yper: dy = 0 if ey > 0 else ny - 1
                                else: wall = True
                            if ez and not 0 <= dz < nz:
                                if zper: dz = 0 if ez > 0 else nz - 1
                                else: wall = True
                            if ex and not 0 <= dx < nx:
                                if xper: dx = 0 if ex > 0 else nx - 1
                                else: out = True
                            if wall:
                                idx.append(pel(o, x, row))
                            elif not out:
                                drow = dy + ny * dz
                                if near[z, y, x]:
                                    assert 0 <= dx + g["nxp"] * drow < g["plane"]   # flags[d]
                                idx.append(pel(i, dx, drow))
                        if ex < 0 and xhi:
                            dy, dz, ok = y + ey, z + ez, True
                            if ey and not 0 <= dy < ny:
                                if yper: dy = 0 if ey > 0 else ny - 1
                                else: ok = False
                            if ez and not 0 <= dz < nz:
                                if zper: dz = 0 if ez > 0 else nz - 1
                                else: ok = False
                            if ok: idx.append(pel(i, x, dy + ny * dz))
                        # the forward pull (fused_common.hpp pull()): unconditional, from a classified source
                        sx, sy, sz, zh, w = x - ex, y - ey, z - ez, False, False
                        if ex > 0 and x == 0:
                            if xper: sx = nx - 1
                            else: zh = True
                        if ex < 0 and x == nx - 1: sx = 0 if xper else nx - 1
                        if ey and not 0 <= sy < ny:
                            if yper: sy = ny - 1 if ey > 0 else 0
                            else: w = True
                        if ez and not 0 <= sz < nz:
                            if zper: sz = nz - 1 if ez > 0 else 0
                            else: w = True
                        plain = not zh and not w
                        srow = sy + ny * sz if plain else row
                        sxx = sx if plain else x
                        solid_src = (not zh and not w and near[z, y, x] and solid[sz, sy, sx])
                        idx.append(pel(o, x, row) if (w and not zh) or solid_src else pel(i, sxx, srow))
                    idx.append(pel(i, x, row))                          # store of mu_m
    return np.array(idx, dtype=np.int64), g


CASES = [
    (27, 24, 10, 8, (False, False, False), None),
    (27, 24, 10, 8, (False, True, True), None),
    (27, 13, 7, 6, (True, False, True), "sphere"),
    (9, 40, 12, 1, (False, False, False), "sphere"),
    (9, 37, 9, 1, (True, True, False), "sphere"),
    (27, 20, 12, 10, (False, False, False), "sphere"),
]


@pytest.mark.parametrize("Q,nx,ny,nz,per,body", CASES)
def test_every_adjoint_access_is_inside_its_buffer(Q, nx, ny, nz, per, body):
    mask = np.zeros((nz, ny, nx), bool)
    if body:
        zz, yy, xx = np.mgrid[0:nz, 0:ny, 0:nx]
        mask = (xx - nx / 3) ** 2 + (yy - ny / 2 + 0.5) ** 2 + (0 if Q == 9 else (zz - nz / 2 + 0.5) ** 2) <= 4.0
    idx, g = kernel_indices(Q, nx, ny, nz, per, mask)
    assert idx.size > 0
    assert idx.min() >= 0 and idx.max() < g["size"]


def test_an_unclassified_gather_leaves_the_buffer_at_a_wall():
    """The gather of mu at (x + e_x, y + e_y, z + e_z) issued before the wall test: for the node
    (0, 0, 0) of the 24x10x8 D3Q27 box and e = (0, -1, 0) the unsigned source row is
    2^32 - 1 and the element index about 7e12, far beyond the 8.4e4-element buffer."""
    g = geometry(27, 24, 10, 8)
    ey, ez = -1, 0
    drow = ((0 + ey) % U32 + g["ny"] * ((0 + ez) % U32)) % U32
    element = 4 * g["ds"] + drow * g["rs"] + 0
    assert element >= g["size"] and element > 1e12
