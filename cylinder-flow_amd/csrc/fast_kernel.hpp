// fast_kernel.hpp -- fused pull step for FP32 / FP16 / BF16 storage, V nodes per lane.
//
// Same boundary rules and arithmetic as the per-node path (fused_common.hpp), organised for
// gfx950 memory throughput:
//  * every lane owns V consecutive x nodes, so each of the 27 pulls and 27 stores is one
//    V*sizeof(storage)-byte access (a dword for two FP16 nodes) instead of V narrow ones;
//  * a wave covers 64*V nodes of one x row.  A per-(row, wave) class, computed once from the
//    mask, picks its path: packed (no solid node or mask link), packed with link words
//    (solid nodes or mask links: bounce-back values and momentum exchange patched in per
//    node), or per-node (padding columns, a periodic x edge, a solid fraction);
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
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
template <> struct UInt<16> { using type = u32x4; };

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

// Byte offsets, for one storage type and geometry, of the direction planes (dst) and of each
// direction's nominal source row relative to the destination row, x shift excluded (src):
//   dst[i] = i * plane * sizeof(T),   src[i] = dst[i] - (e_iy + ny e_iz) * nxp * sizeof(T)
// computed on the host per launch, so a block addresses direction i with one scalar add.
struct PlaneTab {
    int64_t dst[27];
    int64_t src[27];
};

template <int Q, class T>
inline PlaneTab plane_tab(const Geo &g) {
    using L = Lat<Q>;
    PlaneTab t{};
    for (int i = 0; i < Q; i++) {
        t.dst[i] = (int64_t)i * (int64_t)g.plane * (int64_t)sizeof(T);
        t.src[i] = t.dst[i] - ((int64_t)L::ey(i) + (int64_t)g.ny * L::ez(i)) * g.nxp * (int64_t)sizeof(T);
    }
    return t;
}

// Packed path of one lane: V consecutive nodes of a wave with no padding column and no
// periodic x edge.  The source row of each direction is a scalar (walls and periodic y/z are
// decided per block) and only three lane offsets exist (x-1, x, x+1), so a pull is one
// V-wide load from a scalar row pointer plus a 32-bit lane offset.  The inlet lane (x = 0)
// and the outlet lane (x = nx-1) load their row aligned and shift one value.
// NEAR (waves touching the mask): the lane's link words pick, per node and direction, the
// bounced value h_opp(i)(node), loaded V-wide from the node's own row when any node of the
// lane needs it; the momentum exchange of those links is accumulated node by node in
// direction order, as the per-node pull does, and solid nodes are neither updated nor stored.
template <int Q, class S, int V, bool FORCE, bool MACROS, bool NEAR>
__device__ __forceinline__ void packed_nodes(const typename S::T *__restrict__ A, char *__restrict__ Brow,
                                             const uint32_t *__restrict__ links, const double *__restrict__ uin,
                                             double *__restrict__ macros, const FusedArgs &a, const PlaneTab &tab,
                                             int y, int z, int x0, size_t rowoff, double &fx, double &fy,
                                             double &fz) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const bool first = x0 == 0, last = x0 + V == g.nx;
    const uint32_t b0 = (uint32_t)x0 * sizeof(T);
    const uint32_t bP = first ? b0 : b0 - (uint32_t)sizeof(T);   // e_x = +1 pulls x-1
    const uint32_t bM = last ? b0 : b0 + (uint32_t)sizeof(T);    // e_x = -1 pulls x+1
    uint32_t lk[V];
    uint32_t anyl = 0;
    bool all_live = true;
    if constexpr (NEAR) {
        const Pack<uint32_t, V> lp = load_pack<uint32_t, V>((const char *)(links + rowoff) + (size_t)x0 * 4);
        static_for<0, V>([&](auto K) {
            lk[K] = lp.e[K];
            anyl |= lk[K];
            all_live = all_live && !(lk[K] & LINK_SOLID);
        });
    }
    // every source load first (unconditional, valid addresses), then the bounced ones.
    // Interior rows (no wall, no periodic wrap in y/z) address direction i at the scalar
    // Arow + tab.src[i]; boundary rows pick the source row (or the wall bounce-back row) per
    // direction.
    const char *const Arow = (const char *)A + rowoff * sizeof(T);
    const bool interior = (unsigned)(y - 1) < (unsigned)(g.ny - 2) &&
                          (L::D == 2 || (unsigned)(z - 1) < (unsigned)(g.nz - 2));
    Pack<T, V> r[Q];
    bool bbw[Q];   // wall bounce-back: uniform over the block
    if (interior) {
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i);
            const uint32_t bo = ex > 0 ? bP : (ex < 0 ? bM : b0);
            r[i] = load_pack<T, V>(Arow + tab.src[i] + bo);
            bbw[i] = false;
        });
    } else {
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
            int dy = 0, dz = 0;   // source row shift beyond the nominal one (periodic wrap)
            bool bb = false;
            if constexpr (ey != 0) {
                const int sy = y - ey;
                if ((unsigned)sy >= (unsigned)g.ny) {
                    if (g.yper) dy = ey > 0 ? g.ny : -g.ny;
                    else bb = true;
                }
            }
            if constexpr (ez != 0) {
                const int sz = z - ez;
                if ((unsigned)sz >= (unsigned)g.nz) {
                    if (g.zper) dz = ez > 0 ? g.nz : -g.nz;
                    else bb = true;
                }
            }
            const int64_t shift = (int64_t)g.nxp * (dy + (int64_t)g.ny * dz) * (int64_t)sizeof(T);
            const char *rowp = bb ? Arow + tab.dst[o] : Arow + tab.src[i] + shift;
            const uint32_t bo = bb ? b0 : (ex > 0 ? bP : (ex < 0 ? bM : b0));
            r[i] = load_pack<T, V>(rowp + bo);
            bbw[i] = bb;
        });
    }
    Pack<T, V> rb[Q];
    if constexpr (NEAR) {
        static_for<1, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int o = L::opp(i);
            if (anyl & (1u << i)) rb[i] = load_pack<T, V>(Arow + tab.dst[o] + b0);
        });
    }
    float gv[V][Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), o = L::opp(i);
        const bool bb = bbw[i];
        const float off = bb ? a.sp.off[o] : a.sp.off[i];
        float f[V];
        static_for<0, V>([&](auto K) { f[K] = S::template decode<i, L>(r[i].e[K], off); });
        if constexpr (ex > 0 && V > 1) {   // inlet lane: node k pulls element k-1 (node 0: Zou-He)
            if (!bb) static_for<1, V>([&](auto K) { f[V - K] = first ? f[V - K - 1] : f[V - K]; });
        }
        if constexpr (ex < 0 && V > 1) {   // outlet lane: node k pulls k+1, node V-1 clamps to itself
            if (!bb) static_for<0, V - 1>([&](auto K) { f[K] = last ? f[K + 1] : f[K]; });
        }
        if constexpr (NEAR && i > 0) {
            static_for<0, V>([&](auto K) {
                if (lk[K] & (1u << i)) f[K] = S::template decode<i, L>(rb[i].e[K], a.sp.off[o]);
            });
        }
        static_for<0, V>([&](auto K) { gv[K][i] = f[K]; });
    });
    if constexpr (NEAR) {
        static_for<0, V>([&](auto K) {
            if (lk[K] & LINK_SOLID) static_for<0, Q>([&](auto I) { gv[K][I] = 0.f; });
            if constexpr (FORCE) {
                // momentum exchange of the bounced population h_opp(i): 2 h e_opp(i)
                static_for<1, Q>([&](auto I) {
                    constexpr int i = I;
                    constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
                    if (lk[K] & (1u << i)) {
                        const double m = 2.0 * (double)gv[K][i];
                        if constexpr (ex != 0) fx -= ex * m;
                        if constexpr (ey != 0) fy -= ey * m;
                        if constexpr (ez != 0) fz -= ez * m;
                    }
                });
            }
        });
    }
    if (first && (!NEAR || !(lk[0] & LINK_SOLID))) velocity_inlet<Q, float>(gv[0], (float)uin[y + g.ny * z]);
    Pack<T, V> out[Q];
    static_for<0, V>([&](auto K) {
        float c[Q], s, u, v, w;
        collide<Q, false, false>(gv[K], 0.f, a.omega, c, s, u, v, w, fx, fy, fz);
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            out[i].e[K] = S::template enc<i, L>(c[i], a.sp);
        });
        if constexpr (MACROS) {
            const bool live = !NEAR || !(lk[K] & LINK_SOLID);
            double *m = macros + rowoff + x0 + K;
            m[0] = live ? (double)s * (1.0 / 3.0) : 0.0;
            m[g.plane] = live ? (double)u : 0.0;
            m[2 * g.plane] = live ? (double)v : 0.0;
            m[3 * g.plane] = live ? (double)w : 0.0;
        }
    });
    if (!NEAR || all_live) {
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            store_pack<T, V>(Brow + tab.dst[i] + b0, out[i]);
        });
    } else {
        static_for<0, V>([&](auto K) {
            if (!(lk[K] & LINK_SOLID)) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    __builtin_nontemporal_store(out[i].e[K], (T *)(Brow + tab.dst[i] + b0 + K * sizeof(T)));
                });
            }
        });
    }
}

template <int Q, class S, int V, bool FORCE, bool MACROS, bool SFRAC>
__global__ void __launch_bounds__(256)
    k_fused_v(const typename S::T *__restrict__ A, typename S::T *__restrict__ B, const uint8_t *__restrict__ flags,
              const uint8_t *__restrict__ segcls, int nseg, const uint32_t *__restrict__ links,
              const float *__restrict__ sfrac, const double *__restrict__ uin, const int *__restrict__ block_slot,
              double *__restrict__ partials, double *__restrict__ macros, const FusedArgs a, const PlaneTab tab) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // lane index within the x row
    const int y = blockIdx.y, z = blockIdx.z;
    const int x0 = t * V;
    const size_t rowoff = (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
    char *const Brow = (char *)(B + rowoff);                 // this block's row, direction 0
    const int seg = __builtin_amdgcn_readfirstlane(t >> 6);
    const int cls = segcls[(size_t)(y + g.ny * z) * nseg + seg];
    double fx = 0.0, fy = 0.0, fz = 0.0;

    if (cls == WAVE_FAST) {
        packed_nodes<Q, S, V, FORCE, MACROS, false>(A, Brow, links, uin, macros, a, tab, y, z, x0, rowoff, fx, fy, fz);
    } else if (cls == WAVE_NEAR) {
        packed_nodes<Q, S, V, FORCE, MACROS, true>(A, Brow, links, uin, macros, a, tab, y, z, x0, rowoff, fx, fy, fz);
    } else {
        // per-node path: the general pull of fused_common.hpp, one node at a time
        for (int k = 0; k < V; k++) {
            const int x = x0 + k;
            const bool active = x < g.nx;
            const uint32_t node = (uint32_t)(rowoff + (active ? x : 0));   // valid address for idle lanes
            const uint8_t fl = active ? flags[node] : (uint8_t)FL_SOLID;
            const bool live = fl != FL_SOLID;
            float gv[Q], c[Q], s, u, v, w;
            pull<Q, S, FORCE>(A, flags, g, a.sp, x, y, z, node, live, fl == FL_NEAR, gv, fx, fy, fz);
            if (!g.xper && x == 0 && live) velocity_inlet<Q, float>(gv, (float)uin[y + g.ny * z]);
            float sf = 0.f;
            if constexpr (SFRAC) sf = live ? sfrac[node] : 0.f;
            collide<Q, FORCE, SFRAC>(gv, sf, a.omega, c, s, u, v, w, fx, fy, fz);
            if (live) {
                static_for<0, Q>([&](auto I) {
                    constexpr int i = I;
                    __builtin_nontemporal_store(S::template enc<i, L>(c[i], a.sp),
                                                (T *)(Brow + tab.dst[i] + (uint32_t)x * sizeof(T)));
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
    const PlaneTab tab = plane_tab<Q, T>(a.g);
    if (b.sfrac)
        hipLaunchKernelGGL((k_fused_v<Q, S, V, FORCE, MACROS, true>), grid, dim3(sh.bx), 0, s, (const T *)b.A, (T *)b.B,
                           b.flags, b.segcls, b.nseg, b.links, (const float *)b.sfrac, b.uin, b.block_slot, b.partials,
                           b.macros, a, tab);
    else
        hipLaunchKernelGGL((k_fused_v<Q, S, V, FORCE, MACROS, false>), grid, dim3(sh.bx), 0, s, (const T *)b.A, (T *)b.B,
                           b.flags, b.segcls, b.nseg, b.links, (const float *)nullptr, b.uin, b.block_slot, b.partials,
                           b.macros, a, tab);
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
