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
  * padding the LDS rows to T + 1 cuts the bank conflicts of the pull reads (modelled with 64
    banks of 4 bytes, 32 lanes per cycle);
  * the busiest SIMD of the default 24 x 24, K = 5 tile runs 8 wave-steps per launch, against
    15 when every wave runs every step (waves assigned to the 4 SIMDs in turn)."""
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


CASES = [(T, K) for T in (24, 26, 28, 30, 32) for K in range(1, 11) if T - 2 * K >= 2]
NT = {T: (T - 2) ** 2 for T in (24, 26, 28, 30, 32)}


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
def test_each_step_reads_only_slots_written_in_that_step(T, K):
    W = T - 2 * K
    TP = T + 1
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


def test_busiest_simd_wave_steps_default_tile():
    T, K = 24, 5
    W = T - 2 * K
    nwaves = (T * T + 63) // 64    # one thread per tile node, every wave every step

    def busiest(active_waves):
        per = [0] * 4
        for w in active_waves:
            per[w % 4] += 1
        return max(per)

    ring_order = sum(busiest(range((((W + 2 * (K - 1 - s)) ** 2) + 63) // 64)) for s in range(K))
    every_wave = K * busiest(range(nwaves))
    assert (ring_order, every_wave) == (8, 15)
