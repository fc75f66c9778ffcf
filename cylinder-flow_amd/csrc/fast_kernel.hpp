// fast_kernel.hpp -- fused pull step for FP32 / FP16 / BF16 storage, V nodes per lane.
//
// Same boundary rules and arithmetic as the per-node path (fused_common.hpp), organised for
// gfx950 memory throughput:
//  * every lane owns V consecutive x nodes, so each of the 27 pulls and 27 stores is one
//    V*sizeof(storage)-byte access (a dword for two FP16 nodes) instead of V narrow ones;
//  * a wave covers 64*V nodes of one x row.  A per-(row, wave) flag, computed once from the
//    mask, sends waves that touch the body (solid or solid-neighbour nodes, a solid
//    fraction), padding columns or a periodic x edge down the per-node path; every other
//    wave takes the fast path, where the source row of each direction is a scalar (walls
//    and periodic y/z are decided per block) and only three lane offsets exist (x-1, x,
//    x+1), so a pull is one load from a scalar row pointer plus a 32-bit lane offset.  The
//    inlet lane (x = 0) and the outlet lane (x = nx-1) load their row aligned and shift
//    one value;
//  * feq_i and feq_opp(i) share w_i (3p - 1.5|u|^2 + 4.5 (e.u)^2) and differ by the sign of
//    3 w_i e.u, so the equilibrium is evaluated per opposite pair.
// Both paths use collide(), so a node's result does not depend on the path it took.
#pragma once
#include "fused_common.hpp"

namespace lbm {

template <int Q>
constexpr bool pair_first(int i) {
    return i != 0 && i < Lat<Q>::opp(i);
}

// moments, BGK toward the shifted incompressible equilibrium, partial bounce-back
template <int Q, bool FORCE, bool SFRAC>
__device__ __forceinline__ void collide(const float *g, float sf, float om, float *c, float &s, float &u, float &v,
                                        float &w, double &fx, double &fy, double &fz) {
    using L = Lat<Q>;
    s = g[0];
    u = 0.f;
    v = 0.f;
    w = 0.f;
    static_for<1, Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (pair_first<Q>(i)) {
            constexpr int o = L::opp(i);
            s += g[i] + g[o];
            const float d = g[i] - g[o];
            if constexpr (L::ex(i) > 0) u += d;
            else if constexpr (L::ex(i) < 0) u -= d;
            if constexpr (L::ey(i) > 0) v += d;
            else if constexpr (L::ey(i) < 0) v -= d;
            if constexpr (L::ez(i) > 0) w += d;
            else if constexpr (L::ez(i) < 0) w -= d;
        }
    });
    if constexpr (FORCE && SFRAC) {
        fx += 2.0 * (double)sf * (double)u;
        fy += 2.0 * (double)sf * (double)v;
        fz += 2.0 * (double)sf * (double)w;
    }
    // feq_i = w_i (s - 1.5|u|^2 + 4.5 eu^2 + 3 eu), s = sum g = 3p
    const float base = __builtin_fmaf(-1.5f, __builtin_fmaf(u, u, __builtin_fmaf(v, v, w * w)), s);
    const float omb = 1.0f - om;
    c[0] = __builtin_fmaf(omb, g[0], (float)(L::w(0)) * om * base);
    static_for<1, Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (pair_first<Q>(i)) {
            constexpr int o = L::opp(i);
            float eu = 0.f;
            if constexpr (L::ex(i) > 0) eu += u;
            else if constexpr (L::ex(i) < 0) eu -= u;
            if constexpr (L::ey(i) > 0) eu += v;
            else if constexpr (L::ey(i) < 0) eu -= v;
            if constexpr (L::ez(i) > 0) eu += w;
            else if constexpr (L::ez(i) < 0) eu -= w;
            const float wom = (float)L::w(i) * om;
            const float A = wom * __builtin_fmaf(4.5f * eu, eu, base);
            const float Bq = (3.0f * wom) * eu;
            c[i] = __builtin_fmaf(omb, g[i], A + Bq);
            c[o] = __builtin_fmaf(omb, g[o], A - Bq);
        }
    });
    if constexpr (SFRAC) {
        float cc[Q];
        static_for<0, Q>([&](auto I) { cc[I] = c[I]; });
        static_for<0, Q>([&](auto I) { c[I] = __builtin_fmaf(sf, g[L::opp(I)] - cc[I], cc[I]); });
    }
}

// V storage values of one direction, loaded and stored as one access
template <class T, int V>
struct Pack {
    T e[V];
};

template <int Bytes> struct UInt;
template <> struct UInt<2> { using type = uint16_t; };
template <> struct UInt<4> { using type = uint32_t; };
template <> struct UInt<8> { using type = uint64_t; };

template <class T, int V>
__device__ __forceinline__ Pack<T, V> load_pack(const char *p) {
    // 2-byte aligned for the +-1 x shifts of 16-bit data: gfx950 serves it as one access
    using U = typename UInt<sizeof(T) * V>::type;
    U u;
    __builtin_memcpy(&u, p, sizeof(U));
    Pack<T, V> r;
    __builtin_memcpy(&r, &u, sizeof(U));
    return r;
}

template <class T, int V>
__device__ __forceinline__ void store_pack(char *p, const Pack<T, V> &r) {
    using U = typename UInt<sizeof(T) * V>::type;
    U u;
    __builtin_memcpy(&u, &r, sizeof(U));
    __builtin_nontemporal_store(u, (U *)p);
}

template <int Q, class S, int V, bool FORCE, bool MACROS, bool SFRAC>
__global__ void __launch_bounds__(256)
    k_fused_v(const typename S::T *__restrict__ A, typename S::T *__restrict__ B, const uint8_t *__restrict__ flags,
              const uint8_t *__restrict__ segslow, int nseg, const float *__restrict__ sfrac,
              const double *__restrict__ uin, const int *__restrict__ block_slot, double *__restrict__ partials,
              double *__restrict__ macros, const FusedArgs a) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // lane index within the x row
    const int y = blockIdx.y, z = blockIdx.z;
    const int x0 = t * V;
    const size_t rowoff = (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
    char *const Brow = (char *)(B + rowoff);                 // this block's row, direction 0
    const size_t pbytes = g.plane * sizeof(T);               // bytes of one direction plane
    const int seg = __builtin_amdgcn_readfirstlane(t >> 6);
    const bool slow = segslow[(size_t)(y + g.ny * z) * nseg + seg] != 0;
    double fx = 0.0, fy = 0.0, fz = 0.0;

    if (slow) {
        // per-node path: the general pull of fused_common.hpp, one node at a time
        for (int k = 0; k < V; k++) {
            const int x = x0 + k;
            const bool active = x < g.nx;
            const uint32_t node = (uint32_t)(rowoff + (active ? x : 0));   // valid address for idle lanes
            const uint8_t fl = active ? flags[node] : (uint8_t)FL_SOLID;
            const bool live = fl != FL_SOLID;
            float gv[Q], c[Q], s, u, v, w;
            pull<Q, S, FORCE>(A, flags, g, a.sp, x, y, z, node, live, fl == FL_NEAR, gv, fx, fy, fz);
            if (!g.xper && x == 0 && live) zou_he<Q, float>(gv, (float)uin[y + g.ny * z]);
            float sf = 0.f;
            if constexpr (SFRAC) sf = live ? sfrac[node] : 0.f;
            collide<Q, FORCE, SFRAC>(gv, sf, a.omega, c, s, u, v, w, fx, fy, fz);
            if (live) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    __builtin_nontemporal_store(S::template enc<i, L>(c[i], a.sp),
                                                (T *)(Brow + i * pbytes + (uint32_t)x * sizeof(T)));
                });
            }
            if constexpr (MACROS) {
                if (active) {
                    double *m = macros + rowoff + x;
                    m[0] = live ? (double)s * (1.0 / 3.0) : 0.0;
                    m[g.plane] = live ? (double)u : 0.0;
                    m[2 * g.plane] = live ? (double)v : 0.0;
                    m[3 * g.plane] = live ? (double)w : 0.0;
                }
            }
        }
    } else {
        // fast path: no solid or solid-neighbour node, no padding, x not periodic in this wave
        const bool first = x0 == 0, last = x0 + V == g.nx;
        const uint32_t b0 = (uint32_t)x0 * sizeof(T);
        const uint32_t bP = first ? b0 : b0 - (uint32_t)sizeof(T);   // e_x = +1 pulls x-1
        const uint32_t bM = last ? b0 : b0 + (uint32_t)sizeof(T);    // e_x = -1 pulls x+1
        float gv[V][Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
            int sy = y - ey, sz = z - ez;
            bool bb = false;   // wall bounce-back: uniform over the block
            if constexpr (ey != 0) {
                if ((unsigned)sy >= (unsigned)g.ny) {
                    if (g.yper) sy = ey > 0 ? g.ny - 1 : 0;
                    else bb = true;
                }
            }
            if constexpr (ez != 0) {
                if ((unsigned)sz >= (unsigned)g.nz) {
                    if (g.zper) sz = ez > 0 ? g.nz - 1 : 0;
                    else bb = true;
                }
            }
            const char *rowp = bb ? (const char *)(A + (size_t)o * g.plane + rowoff)
                                  : (const char *)(A + (size_t)i * g.plane + g.nxp * ((size_t)sy + (size_t)g.ny * sz));
            const uint32_t bo = bb ? b0 : (ex > 0 ? bP : (ex < 0 ? bM : b0));
            const Pack<T, V> r = load_pack<T, V>(rowp + bo);
            const float off = bb ? a.sp.off[o] : a.sp.off[i];
            float f[V];
            static_for<0, V>([&](auto K) { f[K] = S::template decode<i, L>(r.e[K], off); });
            if constexpr (ex > 0 && V > 1) {   // inlet lane: node k pulls element k-1 (node 0: Zou-He)
                if (!bb) static_for<1, V>([&](auto K) { f[V - K] = first ? f[V - K - 1] : f[V - K]; });
            }
            if constexpr (ex < 0 && V > 1) {   // outlet lane: node k pulls k+1, node V-1 clamps to itself
                if (!bb) static_for<0, V - 1>([&](auto K) { f[K] = last ? f[K + 1] : f[K]; });
            }
            static_for<0, V>([&](auto K) { gv[K][i] = f[K]; });
        });
        if (first) zou_he<Q, float>(gv[0], (float)uin[y + g.ny * z]);
        Pack<T, V> out[Q];
        static_for<0, V>([&](auto K) {
            float c[Q], s, u, v, w;
            collide<Q, false, false>(gv[K], 0.f, a.omega, c, s, u, v, w, fx, fy, fz);
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                out[i].e[K] = S::template enc<i, L>(c[i], a.sp);
            });
            if constexpr (MACROS) {
                double *m = macros + rowoff + x0 + K;
                m[0] = (double)s * (1.0 / 3.0);
                m[g.plane] = (double)u;
                m[2 * g.plane] = (double)v;
                m[3 * g.plane] = (double)w;
            }
        });
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            store_pack<T, V>(Brow + i * pbytes + b0, out[i]);
        });
    }
    if constexpr (FORCE) {
        const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const int slot = block_slot[bid];
        if (slot >= 0) {
            __shared__ double red[3 * 4];
            block_sum3<4>(fx, fy, fz, red);
            if (threadIdx.x == 0) {
                partials[3 * (size_t)slot] = fx;
                partials[3 * (size_t)slot + 1] = fy;
                partials[3 * (size_t)slot + 2] = fz;
            }
        }
    }
}

template <int Q, class S, int V, bool FORCE, bool MACROS>
static hipError_t fast_launch(const StepBufs &b, const FusedArgs &a, hipStream_t s) {
    using T = typename S::T;
    const Shape sh = lane_shape(a.g, V);
    const dim3 grid(sh.gx, a.g.ny, a.g.nz);
    if (b.sfrac)
        hipLaunchKernelGGL((k_fused_v<Q, S, V, FORCE, MACROS, true>), grid, dim3(sh.bx), 0, s, (const T *)b.A, (T *)b.B,
                           b.flags, b.segslow, b.nseg, (const float *)b.sfrac, b.uin, b.block_slot, b.partials, b.macros, a);
    else
        hipLaunchKernelGGL((k_fused_v<Q, S, V, FORCE, MACROS, false>), grid, dim3(sh.bx), 0, s, (const T *)b.A, (T *)b.B,
                           b.flags, b.segslow, b.nseg, (const float *)nullptr, b.uin, b.block_slot, b.partials, b.macros, a);
    return hipGetLastError();
}

template <int Q, class S, int V>
static hipError_t fast_flags(const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s) {
    if (force && mac) return fast_launch<Q, S, V, true, true>(b, a, s);
    if (force) return fast_launch<Q, S, V, true, false>(b, a, s);
    if (mac) return fast_launch<Q, S, V, false, true>(b, a, s);
    return fast_launch<Q, S, V, false, false>(b, a, s);
}

// V must be one of Vs (the instantiated nodes-per-lane widths for this storage type)
template <class S, int... Vs>
static hipError_t fast_step(int Q, int V, const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s) {
    hipError_t r = hipErrorInvalidValue;
    auto one = [&](auto vc) {
        constexpr int W = decltype(vc)::value;
        if (V == W) r = Q == 9 ? fast_flags<9, S, W>(b, force, mac, a, s) : fast_flags<27, S, W>(b, force, mac, a, s);
    };
    (one(std::integral_constant<int, Vs>{}), ...);
    return r;
}

}  // namespace lbm
