"""Thread-to-node map of the several-steps-per-launch kernel (k_tb2d, small_kernels.hip),
transcribed on the CPU, for every tile side and depth the host can select (T = 24 or 32,
K = 1..10 with a core of at least 2 x 2 nodes, and every launch length up to K):

  * threads [0, (T-2)^2) take every node within ring K - 1 of the core exactly once (the
    outermost ring is only loaded, never computed);
  * the first (W + 2m)^2 threads are exactly the nodes within Chebyshev distance m of the
    W x W core, so a step that computes the rings within m runs a prefix of the threads;
  * every LDS slot a step reads for a computed node (its 8 pull sources, before the clamp)
    lies in the tile and was written in the same step, or by the coalesced tile load (x
    fastest) before the first step; the mask-link classification reads in-tile flags;
  * the core store reads back exactly the slots the core's threads wrote;
  * padding the LDS rows to an odd length cuts the bank conflicts of the pull reads (modelled
    with 64 banks of 4 bytes, 32 lanes per cycle), and an even row of 32 doubles would put a
    row segment on one bank;
  * the busiest SIMD of a 24 x 24, K = 5 tile runs 8 wave-steps per launch, against 15 when
    every wave runs every step (waves assigned to the 4 SIMDs in turn); 14 against 28 for the
    28 x 28, K = 7 tile."""
import numpy as np
import pytest

E9 = [(0, 0), (1, 0), (0, 1), (-1, 0), (0, -1), (1, 1), (-1, 1), (-1, -1), (1, -1)]


def tile_map(T, K):
    """(tx, ty) of thread t, as k_tb2d computes it."""
    W = T - 2 * K
    out = []
    for t in range((T - 2) ** 2):
        if t < W * W:
            tx, ty = K + t // W, K + t % W
        else:
            r, side = 1, W + 2
            while t >= side * side:
                r += 1
                side += 2
            q = t - (side - 2) * (side - 2)
            lo = K - r
            hi = lo + side - 1
            if q < side:
                tx, ty = lo, lo + q
            elif q < 2 * side:
                tx, ty = hi, lo + q - side
            elif q < 3 * side - 2:
                ty, tx = lo, lo + 1 + q - 2 * side
            else:
                ty, tx = hi, lo + 1 + q - (3 * side - 2)
        out.append((tx, ty))
    return np.array(out)


def inlet_map(T, K):
    """(tx, ty) of thread t in a tile at the inlet, as k_tb2d computes it: the inlet column
    (tx = K), the core right of it, ring r = 1..K-1 right of it (right column, top row, bottom
    row), then the positions left of the inlet column."""
    W = T - 2 * K
    out = []
    for t in range((T - 2) ** 2):
        q = t
        if q < T - 2:
            out.append((K, 1 + q))
            continue
        q -= T - 2
        if q < (W - 1) * W:
            out.append((K + 1 + q // W, K + q % W))
            continue
        q -= (W - 1) * W
        r = 1
        while r < K and q >= 3 * W + 4 * r - 4:
            q -= 3 * W + 4 * r - 4
            r += 1
        if r < K:
            if q < W + 2 * r:
                out.append((K + W - 1 + r, K - r + q))
            elif q - (W + 2 * r) < W + r - 2:
                out.append((K + 1 + q - (W + 2 * r), K - r))
            else:
                out.append((K + 1 + q - (W + 2 * r) - (W + r - 2), K + W - 1 + r))
        else:
            out.append((1 + q // (T - 2), 1 + q % (T - 2)))
    return np.array(out)


def dist(m, T, K):
    W = T - 2 * K
    dx = np.maximum(np.maximum(K - m[:, 0], m[:, 0] - (K + W - 1)), 0)
    dy = np.maximum(np.maximum(K - m[:, 1], m[:, 1] - (K + W - 1)), 0)
    return np.maximum(dx, dy)


CASES = [(T, K) for T in range(24, 33) for K in range(1, 11) if T - 2 * K >= 2]
NT = {T: (T - 2) ** 2 for T in range(24, 33)}


@pytest.mark.parametrize("T,K", CASES)
def test_ring_order_is_a_bijection_with_prefix_regions(T, K):
    W = T - 2 * K
    m = tile_map(T, K)
    assert m.min() >= 1 and m.max() < T - 1
    slots = (m[:, 0] - 1) * (T - 2) + m[:, 1] - 1
    assert np.array_equal(np.sort(slots), np.arange(NT[T]))
    # Chebyshev distance of each thread's node from the core
    d = np.maximum(np.maximum(K - m[:, 0], m[:, 0] - (K + W - 1)), np.maximum(K - m[:, 1], m[:, 1] - (K + W - 1)))
    d = np.maximum(d, 0)
    for r in range(K):
        n = (W + 2 * r) ** 2
        assert (d[:n] <= r).all() and (d[n:] > r).all()


@pytest.mark.parametrize("T,K", CASES)
def test_inlet_tiles_put_the_inlet_column_in_one_wave(T, K):
    """The inlet tiles' order is a bijection onto the same positions, its first T - 2 threads
    (one wave) hold the inlet column, and the positions left of it come last."""
    W = T - 2 * K
    m = inlet_map(T, K)
    assert len(m) == NT[T] and m.min() >= 1 and m.max() < T - 1
    assert len({(int(x), int(y)) for x, y in m}) == NT[T]
    assert (m[:T - 2, 0] == K).all() and T - 2 <= 64
    assert (m[:, 0] == K).sum() == T - 2
    left = m[:, 0] < K
    assert not left[:NT[T] - (K - 1) * (T - 2)].any() and left[NT[T] - (K - 1) * (T - 2):].all()
    # the live positions within m of the core (right of the inlet column) come before the rest
    d = dist(m, T, K)
    for r in range(K):
        sel = (d <= r) & (m[:, 0] > K)
        n = sel.sum()
        if n:
            first = np.flatnonzero(sel)
            assert first.max() < T - 2 + n, (r, first.max())


@pytest.mark.parametrize("T,K", CASES)
def test_each_step_reads_only_slots_written_in_that_step(T, K):
    W = T - 2 * K
    TP = T + 1 if T % 2 == 0 else T + 2     # the odd row length (tb_row)
    m = tile_map(T, K)
    slot = m[:, 0] * TP + m[:, 1]
    assert slot.max() < T * TP and len(set(slot)) == NT[T]
    # the coalesced load: node j = t, t + NT, ... < T*T of thread t is (j % T, j // T)
    j = np.arange(T * T)
    loaded = np.zeros(T * TP, bool)
    loaded[(j % T) * TP + j // T] = True
    assert loaded.sum() == T * T and loaded[slot].all()
    # mask-link classification (every thread: rings within K - 1): neighbours' flags in the tile
    n1 = NT[T]
    for ex, ey in E9[1:]:
        sx, sy = m[:n1, 0] - ex, m[:n1, 1] - ey
        assert ((sx >= 0) & (sx < T) & (sy >= 0) & (sy < T)).all()
    for steps in range(1, K + 1):
        for s in range(steps):
            rs = steps - 1 - s
            ncomp, nsrc = (W + 2 * rs) ** 2, (W + 2 * rs + 2) ** 2
            assert nsrc <= (T * T if s == 0 else NT[T])
            written = loaded.copy() if s == 0 else np.zeros(T * TP, bool)
            if s > 0:
                written[slot[:nsrc]] = True
            for ex, ey in E9:
                sx, sy = m[:ncomp, 0] - ex, m[:ncomp, 1] - ey
                assert ((sx >= 0) & (sx < T) & (sy >= 0) & (sy < T)).all()
                assert written[sx * TP + sy].all()
        # after the launch the core (threads [0, W*W)) holds the exact values; the store reads
        # node (K + j % W, K + j // W) for j < W*W, which are exactly the slots they wrote
        assert ((m[:W * W] >= K) & (m[:W * W] < K + W)).all()
        jc = np.arange(W * W)
        assert np.array_equal(np.sort(slot[:W * W]), np.sort((K + jc % W) * TP + K + jc // W))


def bank_conflicts(addr_dwords, banks=64):
    """Extra cycles of one wave's 64-bit LDS access (2 dwords per lane), counted per group of
    32 lanes: the largest number of distinct dwords any bank receives, minus one."""
    extra = 0
    for g0 in range(0, len(addr_dwords), 32):
        per = {}
        for a in addr_dwords[g0:g0 + 32]:
            for d in (a, a + 1):
                per.setdefault(d % banks, set()).add(d)
        extra += max(len(v) for v in per.values()) - 1
    return extra


def test_padded_rows_conflict_less():
    """Over the first step's 8 waves and 8 pulled directions of the default tile: 160 extra
    cycles with rows of T + 1 doubles against 240 with rows of T."""
    T, K = 24, 5
    W = T - 2 * K
    m = tile_map(T, K)
    ncomp = (W + 2 * (K - 1)) ** 2
    cost = {}
    for TP in (T, T + 1):
        cost[TP] = 0
        for ex, ey in E9[1:]:
            for w0 in range(0, ncomp, 64):
                cost[TP] += bank_conflicts([2 * ((m[t, 0] - ex) * TP + m[t, 1] - ey) for t in range(w0, min(w0 + 64, ncomp))])
    assert (cost[T + 1], cost[T]) == (160, 240)


@pytest.mark.parametrize("T,K,expect", [(24, 5, (8, 15)), (28, 7, (14, 28))])
def test_busiest_simd_wave_steps(T, K, expect):
    """Wave-steps of the busiest SIMD per launch, ring order against every tile node's wave in
    every step: the 24 x 24 tile at depth 5 and the 28 x 28 tile at depth 7 (the default for a
    400 x 100 domain)."""
    W = T - 2 * K
    nwaves = (T * T + 63) // 64    # one thread per tile node, every wave every step

    def busiest(active_waves):
        per = [0] * 4
        for w in active_waves:
            per[w % 4] += 1
        return max(per)

    ring_order = sum(busiest(range((((W + 2 * (K - 1 - s)) ** 2) + 63) // 64)) for s in range(K))
    every_wave = K * busiest(range(nwaves))
    assert (ring_order, every_wave) == expect


def ghost_check(nx, ny, solid, T, K, steps):
    """The tile load (every live position's loader writes its values and its node's ghost
    values inside the tile, positions left of the inlet get 0) and the write phase of every
    later step (live nodes write their own position, the source position of each bounce-back
    link, and at the outlet the position beyond it) write no LDS position twice, and every pull
    of a live computed node then reads what the boundary rules take: its source's own value, the
    node's own opposite value (wall or solid source), the outlet neighbour's value (zero
    gradient), or a position left of the inlet (the inlet's unknowns, which read 0: written so
    by the load, and never written in a later step)."""
    W = T - 2 * K
    maps = {False: tile_map(T, K), True: inlet_map(T, K)}
    tiles_x, tiles_y = -(-nx // W), -(-ny // W)
    for b in range(tiles_x * tiles_y):
        X0, Y0 = (b % tiles_x) * W - K, (b // tiles_x) * W - K
        m = maps[b % tiles_x == 0]
        d = dist(m, T, K)

        def inside(x, y):
            return 0 <= x < nx and 0 <= y < ny

        def live(x, y):
            return inside(x, y) and not solid[y, x]

        for s in range(0, steps):
            rs = steps - 1 - s
            writes = {}
            if s == 0:
                # the tile load: every position's loader, positions left of the inlet to 0
                writers = [(lx, ly) for ly in range(T) for lx in range(T)]
                for lx, ly in writers:
                    if X0 + lx < 0:
                        for i in range(9):
                            writes[(i, lx, ly)] = ("zero",)
            else:
                writers = [tuple(m[t]) for t in range(len(m)) if d[t] <= rs + 1]
            for tx, ty in writers:
                X, Y = X0 + tx, Y0 + ty
                if not live(X, Y):
                    continue
                for i in range(9):
                    key = (i, tx, ty)
                    assert key not in writes
                    writes[key] = ("own", X, Y, i)
                for i, (ex, ey) in enumerate(E9):
                    if i == 0:
                        continue
                    zh = ex > 0 and X == 0
                    wall = not zh and not 0 <= Y - ey < ny
                    sx = X - ex
                    if ex < 0 and X == nx - 1:
                        sx = nx - 1
                    sol = not zh and not wall and solid[Y - ey, sx]
                    if (wall or sol) and 0 <= tx - ex < T and 0 <= ty - ey < T:
                        key = (i, tx - ex, ty - ey)
                        assert key not in writes
                        writes[key] = ("own", X, Y, E9.index((-ex, -ey)))
                    if X == nx - 1 and ex < 0 and tx + 1 < T:
                        key = (i, tx + 1, ty)
                        assert key not in writes
                        writes[key] = ("own", X, Y, i)
            for t in range(len(m)):
                if d[t] > rs:
                    continue
                tx, ty = m[t]
                X, Y = X0 + tx, Y0 + ty
                if not live(X, Y):
                    continue
                for i, (ex, ey) in enumerate(E9):
                    got = writes.get((i, tx - ex, ty - ey))
                    zh = ex > 0 and X == 0
                    if zh:
                        assert got in (None, ("zero",)) and not inside(X - ex, Y - ey)
                        continue
                    if i > 0 and not 0 <= Y - ey < ny:
                        assert got == ("own", X, Y, E9.index((-ex, -ey)))
                        continue
                    sx = X - ex
                    if ex < 0 and X == nx - 1:
                        sx = nx - 1
                    if solid[Y - ey, sx]:
                        assert got == ("own", X, Y, E9.index((-ex, -ey)))
                    else:
                        assert got == ("own", sx, Y - ey, i)


@pytest.mark.parametrize("T,K", [(24, 5), (28, 7), (24, 3)])
def test_ghost_positions_on_a_mask_touching_every_boundary(T, K):
    nx, ny = 100, 40
    solid = np.zeros((ny, nx), bool)
    solid[2:5, 0] = True            # inlet column
    solid[ny // 2, 1] = True
    solid[3:6, nx - 1] = True       # outlet column
    solid[ny - 4, nx - 2] = True
    solid[0, 40:43] = True          # against the y = 0 wall
    solid[ny // 2, 75] = True
    solid[ny // 2 + 1, 50:54] = True
    yy, xx = np.mgrid[0:ny, 0:nx]
    solid |= (xx - 25) ** 2 + (yy - 19.5) ** 2 < 25
    ghost_check(nx, ny, solid, T, K, K)


def xcd_tile(b, ntiles, tiles_x):
    """Tile of workgroup b in k_tb2d: XCD b % 8 takes a run of consecutive positions of an order
    that runs down bands of two tile rows column by column."""
    pos = (b % 8) * (ntiles // 8) + min(b % 8, ntiles % 8) + b // 8
    band, in_band = pos // (2 * tiles_x), pos % (2 * tiles_x)
    rows = min(2, ntiles // tiles_x - 2 * band)
    return (2 * band + in_band % rows) * tiles_x + in_band // rows


def test_xcd_grouping_is_a_bijection():
    """Every tile once, for every grid of up to 40 x 25 tiles; on 400 x 100 (8 rows of 29 tiles)
    XCD k holds a block of 2 rows and 14-15 columns."""
    for tiles_x in range(1, 41):
        for tiles_y in range(1, 26):
            n = tiles_x * tiles_y
            assert sorted(xcd_tile(b, n, tiles_x) for b in range(n)) == list(range(n))
    n, tiles_x = 232, 29
    for k in range(8):
        tiles = [xcd_tile(b, n, tiles_x) for b in range(k, n, 8)]
        rows = {t // tiles_x for t in tiles}
        cols = {t % tiles_x for t in tiles}
        assert len(tiles) == 29 and len(rows) == 2 and len(cols) in (14, 15)


@pytest.mark.parametrize("nx,ny", [(12, 6), (30, 9), (14, 14), (15, 3)])
def test_ghost_positions_on_domains_smaller_than_a_core(nx, ny):
    """One tile column (every tile is an inlet tile, holding the outlet too) and one or two
    tile rows (walls on both sides), with a solid node next to the inlet, the outlet and a
    wall: the write phases stay race-free and every pull reads what the rules take."""
    solid = np.zeros((ny, nx), bool)
    solid[ny // 2, 1] = True
    solid[0, nx // 2] = True
    solid[ny - 1, nx - 1] = True
    for T, K in ((28, 7), (24, 5), (24, 1)):
        ghost_check(nx, ny, solid, T, K, K)


def test_odd_rows_spread_a_row_segment_over_the_banks():
    """Lanes walking along x (phase A's writes, the core store's reads, ring rows) are one row
    apart: with the odd row length every side 24-32 spreads 16 such lanes over at least 16
    banks, where the even row of 32 doubles (T = 31 with a row of T + 1) puts them on one."""
    for T in range(24, 33):
        TP = T + 1 if T % 2 == 0 else T + 2
        assert TP % 2 == 1
        assert len({(2 * TP * k) % 64 for k in range(16)}) == 16
    assert len({(2 * 32 * k) % 64 for k in range(16)}) == 1
