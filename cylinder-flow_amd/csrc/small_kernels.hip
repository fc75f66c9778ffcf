// small_kernels.hip -- several D2Q9 steps per launch for small domains (plain mode, FP64).
//
// On a 400x100 domain one step moves 5.8 MB and the kernel itself takes about a microsecond,
// so a step costs what a dependent launch costs (~3.4 us on one MI355X; a captured graph of
// launches costs the same).  Temporal blocking removes all but one launch per K steps: block b
// owns a W x W core of nodes and loads its core plus a halo of K nodes on each side (T x T
// nodes, T = 24..32, the side with the least rounds of the CUs times T^2) into the LDS,
// coalesced; each of the K steps pulls from the LDS (boundary rules applied by the values the
// live nodes write, tb_links) and collides in registers (one thread per node), and the live
// nodes write their populations back to the LDS.  A halo node's result is exact as long as every node it pulled from was, so the exact
// region shrinks by one node per step and after K steps covers the core, which is written out;
// each step computes only that region (threads in ring order, the waves outside it idle).  No
// block reads another block's output within a launch: there is no grid-wide synchronisation.
// Each node's arithmetic is that of k_fused (same pull classification, inlet, moments in the
// reference's order, feq_ref, relaxation), so the populations equal the one-launch-per-step
// path bit for bit; forces are summed per tile (a different summation order than the per-row
// blocks of k_fused, the same value to rounding).
#include "fused_common.hpp"

#include <cstdlib>

namespace lbm {

constexpr int TB_KMAX = 10;   // steps per launch at most (tb_depth)

// LDS row length (doubles) of a T x T tile: the odd one of T + 1, T + 2
__host__ __device__ constexpr int tb_row(int T) { return T % 2 == 0 ? T + 1 : T + 2; }

template <int TB_T, bool ANY_FORCE>
__global__ void __launch_bounds__(1024)
    k_tb2d(const double *__restrict__ A, double *__restrict__ B, const uint8_t *__restrict__ flags,
           const uint16_t *__restrict__ lk, const double *__restrict__ uin, double *__restrict__ partials,
           const FusedArgs a, int K, int steps, long step0, int fi, int tiles_x, int ntiles, int pend) {
    using L = Lat<9>;
    // one thread per node of the rings a step can compute (within K - 1 of the core); the
    // outermost ring is only read, and is loaded with the rest of the tile
    constexpr int Q = 9, NT = (TB_T - 2) * (TB_T - 2);
    static_assert((NT + 63) / 64 <= 16, "force records hold 16 waves");
    const Geo &g = a.g;
    const int W = TB_T - 2 * K;          // core side
    // LDS: two halves of Q x TB_T x TP doubles (step s reads half s % 2; node (tx, ty) at
    // tx * TP + ty, rows padded to the odd TP = tb_row(TB_T) so that lanes TP apart hit different
    // banks: an even row of 32 doubles puts every x of a row segment on one bank),
    // then the tile's link words (TB_T x TP) and flags (TB_T x TP bytes)
    constexpr int TP = tb_row(TB_T), HALF = Q * TB_T * TP;
    extern __shared__ double lds[];
    uint16_t *llk = (uint16_t *)(lds + 2 * HALF);
    uint8_t *lfl = (uint8_t *)(llk + TB_T * TP);
    // Threads take the tile's nodes ring by ring from the core outwards: threads [0, W*W) the
    // core (y fastest), then ring r = 1..K-1 (the nodes at Chebyshev distance r from the core:
    // its left and right columns, then its top and bottom rows).  The nodes a step must compute,
    // the live nodes within the steps still to come of the core, are then the first (W + 2m)^2
    // threads, and the waves beyond them skip the step: the busiest SIMD of a 24 x 24 tile runs
    // 2+2+2+1+1 wave-steps in 5 steps instead of 3 per step.  Tiles at the inlet (their left K
    // columns lie outside the domain) take the inlet column x = 0 first, so that one wave runs
    // the inlet's long sequence, then the nodes right of it ring by ring, then those left of it
    // (never computed).
    const int t = threadIdx.x;
    // Workgroups go to the 8 XCDs in turn (b % 8), each XCD with its own L2: XCD k takes the
    // consecutive positions [k (n / 8) + min(k, n % 8), ...) of the n = ntiles in an order that
    // runs down bands of two tile rows column by column, so that an XCD holds a 2-row block of
    // neighbouring tiles, whose loads overlap, read through one L2 (on 400 x 100, 2 x 14-15
    // tiles per XCD: 10 % fewer microseconds per step than neighbouring tiles on different XCDs)
    const int xcd = (int)(blockIdx.x % 8u);
    const int pos = xcd * (ntiles / 8) + min(xcd, ntiles % 8) + (int)(blockIdx.x / 8u);
    const int band = pos / (2 * tiles_x), in_band = pos % (2 * tiles_x);
    const int band_rows = min(2, ntiles / tiles_x - 2 * band);
    const int tile = (2 * band + in_band % band_rows) * tiles_x + in_band / band_rows;
    int tx, ty;
    if (tile % tiles_x != 0) {
        if (t < W * W) {
            tx = K + t / W;
            ty = K + t % W;
        } else {
            int r = 1, side = W + 2;
            while (t >= side * side) {   // at most K - 1 rings: blockDim.x is NT
                r++;
                side += 2;
            }
            const int q = t - (side - 2) * (side - 2), lo = K - r, hi = lo + side - 1;
            if (q < side) {
                tx = lo;
                ty = lo + q;
            } else if (q < 2 * side) {
                tx = hi;
                ty = lo + q - side;
            } else if (q < 3 * side - 2) {
                ty = lo;
                tx = lo + 1 + q - 2 * side;
            } else {
                ty = hi;
                tx = lo + 1 + q - (3 * side - 2);
            }
        }
    } else {
        int q = t;
        if (q < TB_T - 2) {                      // the inlet column
            tx = K;
            ty = 1 + q;
        } else if ((q -= TB_T - 2) < (W - 1) * W) {   // the core right of it
            tx = K + 1 + q / W;
            ty = K + q % W;
        } else {
            q -= (W - 1) * W;
            int r = 1;
            while (r < K && q >= 3 * W + 4 * r - 4) {  // ring r right of the inlet column
                q -= 3 * W + 4 * r - 4;
                r++;
            }
            if (r < K) {
                if (q < W + 2 * r) {                   // its right column
                    tx = K + W - 1 + r;
                    ty = K - r + q;
                } else if ((q -= W + 2 * r) < W + r - 2) {   // its top row
                    ty = K - r;
                    tx = K + 1 + q;
                } else {                               // its bottom row
                    ty = K + W - 1 + r;
                    tx = K + 1 + q - (W + r - 2);
                }
            } else {                                   // left of the inlet column
                tx = 1 + q / (TB_T - 2);
                ty = 1 + q % (TB_T - 2);
            }
        }
    }
    // Chebyshev distance of the node from the core
    const int dx = tx < K ? K - tx : (tx >= K + W ? tx - (K + W - 1) : 0);
    const int dy = ty < K ? K - ty : (ty >= K + W ? ty - (K + W - 1) : 0);
    const int dist = dx > dy ? dx : dy;
    const int slot = tx * TP + ty;       // LDS index of the node
    const int X0 = (tile % tiles_x) * W - K, Y0 = (tile / tiles_x) * W - K;
    const int X = X0 + tx, Y = Y0 + ty;
    const bool act = X >= 0 && X < g.nx && Y >= 0 && Y < g.ny;
    const bool own = act && tx >= K && tx < K + W && ty >= K && ty < K + W;
    const double uy = uin[act ? Y : 0];
    // Every pull is a plain LDS read at the source's position (no boundary rule in the pull):
    // each live node also writes, with its own values, the values its rules pull from
    // positions no live node owns: h_opp(i) into the source position of each bounce-back link i
    // (a wall or a solid source, from its link word: only this node pulls direction i from
    // there), and at the outlet (x = nx - 1) its values of the directions with e_x < 0 into the
    // position beyond the outlet (x = nx), which the outlet node above or below pulls for the
    // zero gradient.  The inlet's unknowns read positions left of the inlet, which hold 0.
    // The tile's link words, flags and populations, x fastest (coalesced), into the LDS (into
    // half 0 with the first step's ghost values, which the loading thread writes for its node;
    // half 1's positions outside the domain to 0): node t, t + NT, ... of the tile's T*T, every
    // load issued (from node (0, 0) for positions outside the domain) before the first is used.
    {
        constexpr int NJ = (TB_T * TB_T + NT - 1) / NT;
        double v[NJ][Q];
        uint8_t f[NJ];
        uint16_t w[NJ];
#pragma unroll
        for (int k = 0; k < NJ; k++) {
            const int j = t + k * NT, lx = j % TB_T, ly = j / TB_T, AX = X0 + lx, AY = Y0 + ly;
            const bool in = j < TB_T * TB_T && AX >= 0 && AX < g.nx && AY >= 0 && AY < g.ny;
            const uint32_t ax = in ? (uint32_t)AX : 0u, ay = in ? (uint32_t)AY : 0u, node = ax + (uint32_t)g.nxp * ay;
            f[k] = flags[node];
            w[k] = lk[node];
            static_for<0, Q>([&](auto I) { v[k][I] = A[pel(g, I, ax, ay)]; });
        }
#pragma unroll
        for (int k = 0; k < NJ; k++) {
            const int j = t + k * NT, lx = j % TB_T, ly = j / TB_T, AX = X0 + lx, AY = Y0 + ly;
            const bool in = AX >= 0 && AX < g.nx && AY >= 0 && AY < g.ny;
            if (j < TB_T * TB_T) {
                const int pos = lx * TP + ly;
                const bool lv = in && f[k] != FL_SOLID;
                lfl[pos] = in ? f[k] : (uint8_t)FL_SOLID;
                llk[pos] = lv ? w[k] : (uint16_t)0;
                if (lv) {
                    static_for<0, Q>([&](auto I) { lds[I * TB_T * TP + pos] = v[k][I]; });
                    const uint32_t bbk = ((uint32_t)(w[k] & 0xFFu) | (uint32_t)(w[k] >> 8)) << 1;
                    if (bbk != 0u)
                        static_for<1, Q>([&](auto I) {
                            constexpr int i = I;
                            constexpr int ex = L::ex(i), ey = L::ey(i), o = L::opp(i);
                            // positions beyond the tile feed only its outermost ring, never computed
                            if (((bbk >> i) & 1u) && (unsigned)(lx - ex) < (unsigned)TB_T &&
                                (unsigned)(ly - ey) < (unsigned)TB_T)
                                lds[i * TB_T * TP + pos - ex * TP - ey] = v[k][o];
                        });
                    if (AX == g.nx - 1 && lx + 1 < TB_T)
                        static_for<1, Q>([&](auto I) {
                            constexpr int i = I;
                            if constexpr (L::ex(i) < 0) lds[i * TB_T * TP + pos + TP] = v[k][i];
                        });
                }
                if (AX < 0) static_for<0, Q>([&](auto I) { lds[I * TB_T * TP + pos] = 0.0; });
                if (!in) static_for<0, Q>([&](auto I) { lds[HALF + I * TB_T * TP + pos] = 0.0; });
            }
        }
    }
    __syncthreads();
    const uint8_t fl = lfl[slot];
    const bool live = fl != FL_SOLID;
    const uint32_t lkw = llk[slot];
    const uint32_t solid_src = (lkw & 0xFFu) << 1;                 // mask links (momentum exchange)
    const uint32_t bb = solid_src | ((lkw >> 8) << 1);              // every bounce-back link
    const bool outl = live && X == g.nx - 1;
    // a node's values in registers: every step pulls from the LDS, and a node writes (from the
    // second step on, and into the core store) only values a step of its own computed
    double h[Q] = {};
    // force records: the steps s = sf, sf + fi, ... of this launch (step0 + s + 1 a multiple of
    // fi); each wave's sums go to red[record][wave] and are added, waves in order, after the
    // launch's last barrier, so no force step waits for a block-wide sum
    __shared__ double red[3 * 16 * TB_KMAX];
    int sf = ANY_FORCE && fi > 0 ? (int)(fi - 1 - step0 % fi) : -1, kr = 0;
    for (int s = 0; s < steps; s++) {
        const bool force = ANY_FORCE && s == sf;
        // two halves: the barrier of step s also orders the reads of step s - 1 before the writes
        // of step s + 1 into the same half
        double *buf = lds + (size_t)(s & 1) * HALF;
        // this step computes the live nodes within rs = steps - 1 - s of the core, from the
        // values of those within rs + 1 (computed by the previous step, or loaded into half 0
        // with the ghost values)
        const int rs = steps - 1 - s;
        if (s > 0) {
            if (dist <= rs + 1 && live) {
                static_for<0, Q>([&](auto I) { buf[I * TB_T * TP + slot] = h[I]; });
                if (bb != 0u)
                    static_for<1, Q>([&](auto I) {
                        constexpr int i = I;
                        constexpr int ex = L::ex(i), ey = L::ey(i), o = L::opp(i);
                        if ((bb >> i) & 1u) buf[i * TB_T * TP + slot - ex * TP - ey] = h[o];
                    });
                if (outl)
                    static_for<1, Q>([&](auto I) {
                        constexpr int i = I;
                        if constexpr (L::ex(i) < 0) buf[i * TB_T * TP + slot + TP] = h[i];
                    });
            }
            __syncthreads();
        }
        double fx = 0.0, fy = 0.0;
        if (dist <= rs && live) {
            double gv[Q];
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                gv[i] = buf[i * TB_T * TP + slot - L::ex(i) * TP - L::ey(i)];
            });
            // momentum exchange of the mask links, in direction order, from the pulled values
            // (before the inlet rule), outside the pull so the other steps run without it
            if (ANY_FORCE && force && solid_src != 0u) {
                static_for<1, Q>([&](auto I) {
                    constexpr int i = I;
                    constexpr int ex = L::ex(i), ey = L::ey(i);
                    if ((solid_src >> i) & 1u) {
                        const double m = 2.0 * gv[i];
                        if constexpr (ex != 0) fx -= ex * m;
                        if constexpr (ey != 0) fy -= ey * m;
                    }
                });
            }
            if (X == 0 && live) velocity_inlet<Q, double>(gv, uy);
            // moments in the reference's term order (lbm.cpp:390-397), relaxation f - (f - feq)/tau
            double p = 0.0, u = 0.0, v = 0.0, w = 0.0;
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                p = p + div_cr(gv[i], 3.0, 1.0 / 3.0);
                if constexpr (L::ex(i) > 0) u = u + gv[i];
                else if constexpr (L::ex(i) < 0) u = u - gv[i];
                if constexpr (L::ey(i) > 0) v = v + gv[i];
                else if constexpr (L::ey(i) < 0) v = v - gv[i];
            });
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                const double feq = feq_ref<Q, i>(p, u, v, w);
                h[i] = gv[i] - div_cr(gv[i] - feq, a.tau, a.rtau);
            });
        }
        if (ANY_FORCE && force) {   // uniform over the grid
            double fz = 0.0;
            if (!own) fx = fy = 0.0;
            wave_sum3(fx, fy, fz);
            if ((t & 63) == 0) {
                double *R = red + 3 * (16 * kr + (t >> 6));
                R[0] = fx;
                R[1] = fy;
                R[2] = fz;
            }
            kr++;
            sf += fi;
        }
    }
    // the core through the LDS half the last step did not read (its last reader, step
    // steps - 2, finished before the barrier of step steps - 1), then stored x fastest
    double *out = lds + (size_t)(steps & 1) * HALF;
    if (dist == 0) static_for<0, Q>([&](auto I) { out[I * TB_T * TP + slot] = h[I]; });
    __syncthreads();
    if (ANY_FORCE && t < kr) {
        // record t of this launch into slot pend + t: the waves' sums in wave order
        constexpr int NW = (NT + 63) / 64;
        const double *R = red + 3 * 16 * t;
        double fx = R[0], fy = R[1], fz = R[2];
        for (int k = 1; k < NW; k++) {
            fx += R[3 * k];
            fy += R[3 * k + 1];
            fz += R[3 * k + 2];
        }
        double *P = partials + 3 * ((size_t)(pend + t) * ntiles + tile);
        P[0] = fx;
        P[1] = fy;
        P[2] = fz;
    }
    for (int j = t; j < W * W; j += NT) {
        const int lx = K + j % W, ly = K + j / W, AX = X0 + lx, AY = Y0 + ly;
        if (AX < g.nx && AY < g.ny && lfl[lx * TP + ly] != FL_SOLID)
            static_for<0, Q>([&](auto I) {
                __builtin_nontemporal_store(out[I * TB_T * TP + lx * TP + ly], &B[pel(g, I, (uint32_t)AX, (uint32_t)AY)]);
            });
    }
}

// The force records of one launch: block r sums the ntiles partials of record r in a fixed order
// into hist[rec0 + r], whose step is step1 + r*fi.
__global__ void __launch_bounds__(256) k_tb_force(const double *__restrict__ partials, int ntiles,
                                                  double *__restrict__ hist, int rec0, long step1, int fi) {
    const int r = blockIdx.x;
    const double *P = partials + 3 * (size_t)r * ntiles;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int j = threadIdx.x; j < ntiles; j += 256) {
        fx += P[3 * (size_t)j];
        fy += P[3 * (size_t)j + 1];
        fz += P[3 * (size_t)j + 2];
    }
    __shared__ double red[3 * 4];
    block_sum3<4>(fx, fy, fz, red);
    if (threadIdx.x == 0) {
        double *h = hist + 4 * (size_t)(rec0 + r);
        h[0] = (double)(step1 + (long)r * fi);
        h[1] = fx;
        h[2] = fy;
        h[3] = fz;
    }
}

// steps per launch: 7 (LBM_TB_DEPTH overrides, 1-10); with the 14 x 14 core of a 400 x 100
// domain's 28 x 28 tiles, 7 steps per launch measured fastest (2.54 us per step against 2.61 at
// 5 steps on 24 x 24 tiles and 2.60 at 8 on 30 x 30)
int tb_depth() {
    if (const char *e = std::getenv("LBM_TB_DEPTH")) {
        const int k = std::atoi(e);
        if (k >= 1 && k <= TB_KMAX) return k;
    }
    return 7;
}

// two population halves of 9 x T x tb_row(T) doubles, the tile's link words and flags (T x
// tb_row(T) of each)
size_t tb_lds_bytes(int T) { return sizeof(double) * 2 * 9 * T * tb_row(T) + (size_t)3 * T * tb_row(T); }

int tb_tiles(const Geo &g, int T, int K, int *tiles_x) {
    const int W = T - 2 * K;
    *tiles_x = (g.nx + W - 1) / W;
    return *tiles_x * ((g.ny + W - 1) / W);
}

// the tile sides instantiated, and the tile side for a domain (leaving a core of at least 2 x 2
// nodes); LBM_TB_TILE overrides with any instantiated side that leaves a 2 x 2 core
static const int kTbSides[] = {24, 25, 26, 27, 28, 29, 30, 31, 32};

int tb_tile(const Geo &g, int K, int cus) {
    if (const char *e = std::getenv("LBM_TB_TILE")) {
        const int t = std::atoi(e);
        for (int T : kTbSides)
            if (t == T && t - 2 * K >= 2) return t;
    }
    // the side with the least (rounds of the CUs) x (positions per tile); measured best on
    // 400 x 100 (27), 600 x 150 (28), 800 x 200 (29) and 1000 x 250 (30) at 7 steps per launch
    int best = 32, tx = 0;
    long best_cost = -1;
    for (int T : kTbSides) {
        if (T - 2 * K < 2) continue;
        const long rounds = (tb_tiles(g, T, K, &tx) + cus - 1) / cus, cost = rounds * T * T;
        if (best_cost < 0 || cost < best_cost) {
            best = T;
            best_cost = cost;
        }
    }
    return best;
}

template <int T>
static hipError_t tb_attr() {
    hipError_t e = hipFuncSetAttribute((const void *)k_tb2d<T, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)tb_lds_bytes(T));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_tb2d<T, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tb_lds_bytes(T));
    return e;
}

hipError_t tb_prepare() {
    hipError_t e = tb_attr<24>();
    if (e == hipSuccess) e = tb_attr<25>();
    if (e == hipSuccess) e = tb_attr<26>();
    if (e == hipSuccess) e = tb_attr<27>();
    if (e == hipSuccess) e = tb_attr<28>();
    if (e == hipSuccess) e = tb_attr<29>();
    if (e == hipSuccess) e = tb_attr<30>();
    if (e == hipSuccess) e = tb_attr<31>();
    if (e == hipSuccess) e = tb_attr<32>();
    return e;
}

std::vector<uint16_t> tb_links(const Geo &g, const std::vector<uint8_t> &flags) {
    using L = Lat<9>;
    std::vector<uint16_t> lk(flags.size(), 0);
    for (int y = 0; y < g.ny; y++)
        for (int x = 0; x < g.nx; x++) {
            const size_t node = (size_t)x + (size_t)g.nxp * y;
            if (flags[node] == FL_SOLID) continue;
            uint16_t b = 0;
            for (int i = 1; i < 9; i++) {
                const int ex = L::ex(i), ey = L::ey(i);
                if (ex > 0 && x == 0) continue;                            // Zou-He unknown
                const int sy = y - ey;
                if (sy < 0 || sy >= g.ny) {                                // wall
                    b |= (uint16_t)(1u << (7 + i));
                    continue;
                }
                const int sx = (ex < 0 && x == g.nx - 1) ? x : x - ex;     // outlet clamp
                if (flags[(size_t)sx + (size_t)g.nxp * sy] == FL_SOLID) b |= (uint16_t)(1u << (i - 1));
            }
            lk[node] = b;
        }
    return lk;
}

// record slots of the partials buffer: the force records of several launches are summed by one
// k_tb_force launch (a launch of its own costs about as much as a step)
int tb_records(int K) { return K > 128 ? K : 128; }

// `steps` (<= K) steps from A into B, the first being step0 + 1; the partials of the force records
// among them go to record slots pend, pend + 1, ... (the caller keeps pend + steps within
// tb_records(K) and sums them with tb_force_flush).  Returns the number of records.
hipError_t tb_steps(const void *A, void *B, const uint8_t *flags, const uint16_t *lk, const double *uin, double *partials, int pend,
                    int *nrec, const FusedArgs &a, int T, int K, int steps, long step0, int fi, hipStream_t s) {
    int tiles_x = 0;
    const int ntiles = tb_tiles(a.g, T, K, &tiles_x);
    const size_t lds = tb_lds_bytes(T);
    const int nforce = fi > 0 ? (int)((step0 + steps) / fi - step0 / fi) : 0;
    *nrec = nforce;
    if (pend < 0 || pend + nforce > tb_records(K)) return hipErrorInvalidValue;
    if (T - 2 * K < 1 || K > TB_KMAX || steps < 1 || steps > K) return hipErrorInvalidValue;
    auto go = [&](auto kern) {   // (T - 2)^2 threads: the kernel's thread map assumes exactly that
        hipLaunchKernelGGL(kern, dim3(ntiles), dim3((T - 2) * (T - 2)), lds, s, (const double *)A, (double *)B, flags, lk, uin,
                           partials, a, K, steps, step0, fi, tiles_x, ntiles, pend);
    };
    switch (T) {
    case 24: nforce > 0 ? go(k_tb2d<24, true>) : go(k_tb2d<24, false>); break;
    case 25: nforce > 0 ? go(k_tb2d<25, true>) : go(k_tb2d<25, false>); break;
    case 26: nforce > 0 ? go(k_tb2d<26, true>) : go(k_tb2d<26, false>); break;
    case 27: nforce > 0 ? go(k_tb2d<27, true>) : go(k_tb2d<27, false>); break;
    case 28: nforce > 0 ? go(k_tb2d<28, true>) : go(k_tb2d<28, false>); break;
    case 29: nforce > 0 ? go(k_tb2d<29, true>) : go(k_tb2d<29, false>); break;
    case 30: nforce > 0 ? go(k_tb2d<30, true>) : go(k_tb2d<30, false>); break;
    case 31: nforce > 0 ? go(k_tb2d<31, true>) : go(k_tb2d<31, false>); break;
    case 32: nforce > 0 ? go(k_tb2d<32, true>) : go(k_tb2d<32, false>); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t tb_force_flush(const double *partials, const Geo &g, int T, int K, double *hist, int rec0, int npend,
                          long step1, int fi, hipStream_t s) {
    if (npend <= 0) return hipSuccess;
    if (npend > tb_records(K) || fi <= 0) return hipErrorInvalidValue;
    int tiles_x = 0;
    const int ntiles = tb_tiles(g, T, K, &tiles_x);
    hipLaunchKernelGGL(k_tb_force, dim3(npend), dim3(256), 0, s, partials, ntiles, hist, rec0, step1, fi);
    return hipGetLastError();
}

}  // namespace lbm
