// small_kernels.hip -- several D2Q9 steps per launch for small domains (plain mode, FP64).
//
// On a 400x100 domain one step moves 5.8 MB and the kernel itself takes about a microsecond,
// so a step costs what a dependent launch costs (~3.4 us on one MI355X; a captured graph of
// launches costs the same).  Temporal blocking removes all but one launch per K steps: block b
// owns a W x W core of nodes and loads its core plus a halo of K nodes on each side (T x T
// nodes, one per thread, T = 24 or 32: the smaller tile when its blocks still fit the CUs
// once) into registers; each of the K steps writes the
// block's populations to LDS, pulls from there with the plain-mode boundary rules, and collides
// in registers.  A halo node's result is exact as long as every node it pulled from was, so the
// exact region shrinks by one node per step and after K steps covers the core, which is written
// out.  No block reads another block's output within a launch: there is no grid-wide
// synchronisation.  Each node's arithmetic is that of k_fused (same pull classification, inlet,
// moments in the reference's order, feq_ref, relaxation), so the populations equal the
// one-launch-per-step path bit for bit; forces are summed per tile (a different summation order
// than the per-row blocks of k_fused, the same value to rounding).
#include "fused_common.hpp"

#include <cstdlib>

namespace lbm {

template <int TB_T, bool ANY_FORCE>
__global__ void __launch_bounds__(1024)
    k_tb2d(const double *__restrict__ A, double *__restrict__ B, const uint8_t *__restrict__ flags,
           const double *__restrict__ uin, double *__restrict__ partials, const FusedArgs a, int K, int steps,
           long step0, int fi, int tiles_x, int ntiles, int pend) {
    using L = Lat<9>;
    constexpr int Q = 9;
    const Geo &g = a.g;
    const int W = TB_T - 2 * K;          // core side
    extern __shared__ double lds[];      // 2 x Q x TB_T*TB_T: step s writes and reads half s % 2
    // y fastest: the tile's inlet column (x = 0) lies in one or two waves, so only those run the
    // inlet's long FP64 sequence (row-major order would put one inlet lane in every wave)
    const int t = threadIdx.x, ty = t % TB_T, tx = t / TB_T;
    const int X = (int)(blockIdx.x % tiles_x) * W - K + tx, Y = (int)(blockIdx.x / tiles_x) * W - K + ty;
    const bool act = X >= 0 && X < g.nx && Y >= 0 && Y < g.ny;
    const bool own = act && tx >= K && tx < K + W && ty >= K && ty < K + W;
    const uint32_t node = act ? (uint32_t)X + (uint32_t)g.nxp * (uint32_t)Y : 0u;
    const uint8_t fl = act ? flags[node] : (uint8_t)FL_SOLID;
    const bool live = fl != FL_SOLID;
    // sources that are solid (mask bounce-back), classified once as the pull does
    uint32_t solid_src = 0;
    if (fl == FL_NEAR) {
        static_for<1, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i);
            if (ex > 0 && X == 0) return;                        // Zou-He unknown
            const int sy = Y - ey;
            if ((unsigned)sy >= (unsigned)g.ny) return;          // wall
            int sx = X - ex;
            if (ex < 0 && X == g.nx - 1) sx = g.nx - 1;          // outlet clamp
            if (flags[(uint32_t)sx + (uint32_t)g.nxp * (uint32_t)sy] == FL_SOLID) solid_src |= 1u << i;
        });
    }
    double h[Q];
    static_for<0, Q>([&](auto I) { h[I] = act ? A[pel(g, I, (uint32_t)(act ? X : 0), (uint32_t)(act ? Y : 0))] : 0.0; });
    const double uy = act ? uin[Y] : 0.0;
    for (int s = 0; s < steps; s++) {
        const long step = step0 + s + 1;
        const bool force = ANY_FORCE && fi > 0 && step % fi == 0;
        // two halves: the barrier of step s also orders the reads of step s - 1 before the writes
        // of step s + 1 into the same half
        double *buf = lds + (size_t)(s & 1) * Q * TB_T * TB_T;
        static_for<0, Q>([&](auto I) { buf[I * TB_T * TB_T + t] = h[I]; });
        __syncthreads();
        double gv[Q];
        double fx = 0.0, fy = 0.0;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), o = L::opp(i);
            const bool zh = ex > 0 && X == 0;
            const int sy = Y - ey;
            const bool wall = !zh && (unsigned)sy >= (unsigned)g.ny;
            const bool solid = (solid_src >> i) & 1u;
            int sx = X - ex;
            if (ex < 0 && X == g.nx - 1) sx = g.nx - 1;
            // tile coordinates of the source, kept inside the tile (a source beyond the tile only
            // feeds halo nodes whose results no core node uses within the launch)
            int stx = tx + (sx - X), sty = ty - ey;
            stx = stx < 0 ? 0 : (stx >= TB_T ? TB_T - 1 : stx);
            sty = sty < 0 ? 0 : (sty >= TB_T ? TB_T - 1 : sty);
            const double pulled = buf[i * TB_T * TB_T + stx * TB_T + sty];
            const double v = (wall || solid) ? h[o] : pulled;
            gv[i] = (live && !zh) ? v : 0.0;
            if (ANY_FORCE && force && solid) {
                const double m = 2.0 * gv[i];
                if constexpr (ex != 0) fx -= ex * m;
                if constexpr (ey != 0) fy -= ey * m;
            }
        });
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
            const double c = gv[i] - div_cr(gv[i] - feq, a.tau, a.rtau);
            h[i] = live ? c : 0.0;
        });
        if (ANY_FORCE && force) {   // uniform over the grid
            double fz = 0.0;
            if (!own) fx = fy = 0.0;
            __shared__ double red[3 * 16];
            block_sum3<16>(fx, fy, fz, red);
            if (t == 0) {
                // record slot: the records already pending, then the index within the launch
                const int r = pend + (int)(step / fi - (step0 + 1 + fi - 1) / fi);
                double *P = partials + 3 * ((size_t)r * ntiles + blockIdx.x);
                P[0] = fx;
                P[1] = fy;
                P[2] = fz;
            }
        }
    }
    if (own && live) static_for<0, Q>([&](auto I) { __builtin_nontemporal_store(h[I], &B[pel(g, I, (uint32_t)X, (uint32_t)Y)]); });
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

// steps per launch: 5 (LBM_TB_DEPTH overrides, 1-10)
int tb_depth() {
    if (const char *e = std::getenv("LBM_TB_DEPTH")) {
        const int k = std::atoi(e);
        if (k >= 1 && k <= 10) return k;
    }
    return 5;
}

size_t tb_lds_bytes(int T) { return sizeof(double) * 2 * 9 * T * T; }

int tb_tiles(const Geo &g, int T, int K, int *tiles_x) {
    const int W = T - 2 * K;
    *tiles_x = (g.nx + W - 1) / W;
    return *tiles_x * ((g.ny + W - 1) / W);
}

// the tile side for a domain: 24 when its tiles fit the CUs once (and leave a core of at least
// 2 x 2 nodes), else 32; LBM_TB_TILE=24 or 32 overrides where the core stays at least 2 x 2
int tb_tile(const Geo &g, int K, int cus) {
    if (const char *e = std::getenv("LBM_TB_TILE")) {
        const int t = std::atoi(e);
        if ((t == 24 || t == 32) && t - 2 * K >= 2) return t;
    }
    int tx = 0;
    return 24 - 2 * K >= 2 && tb_tiles(g, 24, K, &tx) <= cus ? 24 : 32;
}

hipError_t tb_prepare() {
    hipError_t e = hipFuncSetAttribute((const void *)k_tb2d<24, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)tb_lds_bytes(24));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_tb2d<24, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tb_lds_bytes(24));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_tb2d<32, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tb_lds_bytes(32));
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_tb2d<32, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)tb_lds_bytes(32));
    return e;
}

// record slots of the partials buffer: the force records of several launches are summed by one
// k_tb_force launch (a launch of its own costs about as much as a step)
int tb_records(int K) { return K > 128 ? K : 128; }

// `steps` (<= K) steps from A into B, the first being step0 + 1; the partials of the force records
// among them go to record slots pend, pend + 1, ... (the caller keeps pend + steps within
// tb_records(K) and sums them with tb_force_flush).  Returns the number of records.
hipError_t tb_steps(const void *A, void *B, const uint8_t *flags, const double *uin, double *partials, int pend,
                    int *nrec, const FusedArgs &a, int T, int K, int steps, long step0, int fi, hipStream_t s) {
    int tiles_x = 0;
    const int ntiles = tb_tiles(a.g, T, K, &tiles_x);
    const size_t lds = tb_lds_bytes(T);
    const int nforce = fi > 0 ? (int)((step0 + steps) / fi - step0 / fi) : 0;
    *nrec = nforce;
    if (pend < 0 || pend + nforce > tb_records(K)) return hipErrorInvalidValue;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(ntiles), dim3(T * T), lds, s, (const double *)A, (double *)B, flags, uin, partials, a,
                           K, steps, step0, fi, tiles_x, ntiles, pend);
    };
    if (T == 24) {
        if (nforce > 0) go(k_tb2d<24, true>);
        else go(k_tb2d<24, false>);
    } else {
        if (nforce > 0) go(k_tb2d<32, true>);
        else go(k_tb2d<32, false>);
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
