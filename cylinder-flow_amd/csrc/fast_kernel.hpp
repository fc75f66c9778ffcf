// fast_kernel.hpp -- fused pull step for FP32 / FP16 / BF16 storage, V nodes per lane.
//
// Same boundary rules and arithmetic as the per-node path (fused_common.hpp), organised for
// gfx950 memory throughput.  One step is one launch on the solver stream, plus a second when
// the fix-up list below is not empty:
//  * k_fast_v runs the packed path on every lane of every row: each lane owns V consecutive x
//    nodes, so each of the 27 pulls and 27 stores is one V*sizeof(storage)-byte access (a
//    dword for two FP16 nodes); the source row of a direction is a scalar and only three lane
//    offsets exist (x-1, x, x+1).  The populations of a row are row-interleaved (direction i
//    of row r at i*nxp + r*Q*nxp), so a row's update reads and writes Q*nxp contiguous
//    elements plus the source rows' stretches.  Optionally (aligned_pull), 16-bit interior
//    rows load every direction aligned and build the x-1 / x+1 pairs from the neighbouring
//    lane's dword, so a wave's request spans two cache lines instead of three.  It applies
//    walls, periodic y/z, the inlet and the outlet.  The mask's half-way bounce-back is
//    confined to a per-row window of lanes (built once from the mask; empty for most rows):
//    a lane inside it reads its nodes' link words (bit i: direction i pulls from a solid node)
//    and, for each linked direction, replaces the pulled value by its own node's h_opp(i),
//    loaded from the same row (a line the solid neighbour's lane reads anyway).  The step
//    reads only its input buffer and writes only its output buffer;
//  * on steps that record the force, k_link_force (one lane per linked node, before it) sums
//    each linked node's momentum exchange from the same input values into its partial slot:
//    the FP64 sums stay out of the packed kernel, which keeps one instruction schedule;
//  * after it, k_fix_v recomputes, with the per-node pull, the few nodes the packed path still
//    gets wrong (a compact list built once from the mask: nodes with a solid fraction, the
//    periodic x edges, the last node of a lane that reaches into the padding, mask neighbours
//    on the outlet column, whose clamped pull the window does not cover).
// The rare boundary work thus costs the fast kernel few registers.  Both kernels use
// collide(), so a node's result does not depend on the path it took.
//  * feq_i and feq_opp(i) share w_i (3p - 1.5|u|^2 + 4.5 (e.u)^2) and differ by the sign of
//    3 w_i e.u, so the equilibrium is evaluated per opposite pair.
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

// rank of direction i among the directions with the same sign of e_x
template <int Q>
constexpr int xrank(int i) {
    int k = 0;
    for (int j = 0; j < i; j++)
        if ((Lat<Q>::ex(j) > 0) == (Lat<Q>::ex(i) > 0) && (Lat<Q>::ex(j) < 0) == (Lat<Q>::ex(i) < 0)) k++;
    return k;
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

// Byte offsets, for one storage type and geometry, of direction i's row relative to the start
// of a row's populations (dst) and of its nominal source row, x shift excluded (src):
//   dst[i] = i * ds * sizeof(T),   src[i] = dst[i] - (e_iy + ny e_iz) * rs * sizeof(T)
// computed on the host per launch, so a block addresses direction i with one scalar add.
struct DirTab {
    int64_t dst[27];
    int64_t src[27];
};

template <int Q, class T>
inline DirTab dir_tab(const Geo &g) {
    using L = Lat<Q>;
    DirTab t{};
    for (int i = 0; i < Q; i++) {
        t.dst[i] = (int64_t)i * (int64_t)g.ds * (int64_t)sizeof(T);
        t.src[i] = t.dst[i] - ((int64_t)L::ey(i) + (int64_t)g.ny * L::ez(i)) * (int64_t)g.rs * (int64_t)sizeof(T);
    }
    return t;
}

// Aligned pull of an interior row for two 16-bit nodes per lane, the whole row in one block
// (up to 4 waves).  Every direction is loaded as the lane's own aligned dword D_t = (h[2t],
// h[2t+1]) of its source row; e_x = +1 needs (h[2t-1], h[2t]), the high half of the previous
// lane's dword and the low half of its own, e_x = -1 needs (h[2t+1], h[2t+2]).  The
// neighbour's dword comes from a wave shift (DPP wave_shr:1 / wave_shl:1); at a wave edge the
// shift returns the `old` operand, which holds the dword that the neighbouring wave's edge
// lane left in LDS.  The inlet lane (x = 0) and the lane holding the last node keep their own
// dword, which the caller's clamp then shifts as on the unaligned path.  Each request covers
// two 128-byte lines instead of three.
template <int Q, class T>
__device__ __forceinline__ void aligned_pull(const char *Arow, const DirTab &tab, uint32_t b0, bool first,
                                             bool last, Pack<T, 2> *r) {
    using L = Lat<Q>;
    constexpr int NS = Q == 27 ? 9 : 3;   // directions per sign of e_x
    __shared__ uint32_t up[4][NS], dn[4][NS];
    const int lane = (int)(threadIdx.x & 63u), wv = (int)(threadIdx.x >> 6);
    const int nw = (int)((blockDim.x + 63u) >> 6);
    uint32_t D[Q];
    static_for<0, Q>([&](auto I) { D[I] = load_pack<uint32_t, 1>(Arow + tab.src[I] + b0).e[0]; });
    if (lane == 63)
        static_for<0, Q>([&](auto I) {
            if constexpr (L::ex(I) > 0) up[wv][xrank<Q>(I)] = D[I];
        });
    if (lane == 0)
        static_for<0, Q>([&](auto I) {
            if constexpr (L::ex(I) < 0) dn[wv][xrank<Q>(I)] = D[I];
        });
    __syncthreads();
    const int wp = wv > 0 ? wv - 1 : 0, wn = wv + 1 < nw ? wv + 1 : wv;
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        uint32_t v = D[i];
        if constexpr (L::ex(i) > 0) {
            const int P = __builtin_amdgcn_update_dpp((int)up[wp][xrank<Q>(i)], (int)D[i], 0x138, 0xf, 0xf, false);
            v = first ? D[i] : __builtin_amdgcn_alignbit(D[i], (uint32_t)P, 16);
        } else if constexpr (L::ex(i) < 0) {
            const int N = __builtin_amdgcn_update_dpp((int)dn[wn][xrank<Q>(i)], (int)D[i], 0x130, 0xf, 0xf, false);
            v = last ? D[i] : __builtin_amdgcn_alignbit((uint32_t)N, D[i], 16);
        }
        __builtin_memcpy(&r[i], &v, 4);
    });
}

// Packed path of one lane: V consecutive nodes of one x row.  Interior rows (no wall, no
// periodic wrap in y/z) address direction i at the scalar Arow + tab.src[i]; boundary rows pick
// the source row (or the wall bounce-back row) per direction.  The inlet lane (x = 0) and the
// lane holding the outlet node (x = nx-1) load their row aligned and shift one value; a lane
// past the outlet (padding) clamps the same way, so every access stays inside its row.
template <int Q, class S, int V, bool MACROS, bool ALIGNED>
__device__ __forceinline__ void packed_nodes(const typename S::T *__restrict__ A, char *__restrict__ Brow,
                                             const double *__restrict__ uin, double *__restrict__ macros,
                                             const FusedArgs &a, const DirTab &tab, int y, int z, int x0,
                                             size_t rowoff, size_t poprow, int4 win, int tl,
                                             const uint32_t *__restrict__ wlk) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const bool first = x0 == 0, last = x0 + V >= g.nx;
    const uint32_t b0 = (uint32_t)x0 * sizeof(T);
    const uint32_t bP = first ? b0 : b0 - (uint32_t)sizeof(T);   // e_x = +1 pulls x-1
    const uint32_t bM = last ? b0 : b0 + (uint32_t)sizeof(T);    // e_x = -1 pulls x+1
    const char *const Arow = (const char *)A + poprow * sizeof(T);
    const bool interior = (unsigned)(y - 1) < (unsigned)(g.ny - 2) &&
                          (L::D == 2 || (unsigned)(z - 1) < (unsigned)(g.nz - 2));
    Pack<T, V> r[Q];
    bool bbw[Q];   // wall bounce-back: uniform over the block
    if (ALIGNED && interior) {
        if constexpr (ALIGNED) aligned_pull<Q, T>(Arow, tab, b0, first, last, r);
        static_for<0, Q>([&](auto I) { bbw[I] = false; });
    } else if (interior) {
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
            const int64_t shift = (int64_t)g.rs * (dy + (int64_t)g.ny * dz) * (int64_t)sizeof(T);
            const char *rowp = bb ? Arow + tab.dst[o] : Arow + tab.src[i] + shift;
            const uint32_t bo = bb ? b0 : (ex > 0 ? bP : (ex < 0 ? bM : b0));
            r[i] = load_pack<T, V>(rowp + bo);
            bbw[i] = bb;
        });
    }
    // Mask links (lanes of the row's window only; never the inlet or outlet lane, so element K
    // of a pack is node K's value): direction i of a node whose source x - e_i is solid takes the
    // node's own stored h_opp(i) (half-way bounce-back; exact as stored, because the host sends
    // link nodes to the fix-up list when opposite directions have different storage offsets).
    if (win.y > 0 && (unsigned)tl < (unsigned)win.y) {
        uint32_t lk[V], any = 0;
        static_for<0, V>([&](auto K) {
            lk[K] = wlk[(size_t)win.z + (size_t)tl * V + K];
            any |= lk[K];
        });
        Pack<T, V> own[Q];
        static_for<1, Q>([&](auto I) {   // every bounced value first, so the loads are in flight together
            constexpr int i = I;
            if ((any >> i) & 1u) own[i] = load_pack<T, V>(Arow + tab.dst[L::opp(i)] + b0);
        });
        static_for<1, Q>([&](auto I) {
            constexpr int i = I;
            static_for<0, V>([&](auto K) {
                if ((lk[K] >> i) & 1u) r[i].e[K] = own[i].e[K];
            });
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
        if constexpr (ex > 0 && V > 1) {   // inlet lane: node k pulls element k-1 (node 0: inlet)
            if (!bb) static_for<1, V>([&](auto K) { f[V - K] = first ? f[V - K - 1] : f[V - K]; });
        }
        if constexpr (ex < 0 && V > 1) {   // outlet lane: node k pulls k+1, node V-1 clamps to itself
            if (!bb) static_for<0, V - 1>([&](auto K) { f[K] = last ? f[K + 1] : f[K]; });
        }
        static_for<0, V>([&](auto K) { gv[K][i] = f[K]; });
    });
    if (first) velocity_inlet<Q, float>(gv[0], (float)uin[y + g.ny * z]);
    Pack<T, V> out[Q];
    static_for<0, V>([&](auto K) {
        float c[Q], s, u, v, w;
        double fx = 0.0, fy = 0.0, fz = 0.0;
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
        store_pack<T, V>(Brow + tab.dst[i] + b0, out[i]);
    });
}

// Every lane of every row (256-lane blocks along x, one row per (blockIdx.y, blockIdx.z)):
// lanes past the padded row return; nothing else branches on the node but the row's link
// window, and there is no barrier.  rowwin may be null (no mask links).
template <int Q, class S, int V, bool MACROS, bool ALIGNED>
__global__ void __launch_bounds__(256)
    k_fast_v(const typename S::T *__restrict__ A, typename S::T *__restrict__ B, const double *__restrict__ uin,
             double *__restrict__ macros, const int4 *__restrict__ rowwin, const uint32_t *__restrict__ wlk,
             const FusedArgs a, const DirTab tab) {
    const Geo &g = a.g;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;   // lane index within the x row
    if (t * V >= g.nxp) return;
    const int y = blockIdx.y, z = blockIdx.z;
    const size_t row = (size_t)y + (size_t)g.ny * z;
    const size_t rowoff = (size_t)g.nxp * row, poprow = g.rs * row;
    const int4 win = rowwin ? rowwin[row] : make_int4(0, 0, 0, 0);
    packed_nodes<Q, S, V, MACROS, ALIGNED>(A, (char *)(B + poprow), uin, macros, a, tab, y, z, t * V, rowoff, poprow,
                                           win, t - win.x, wlk);
}

// Momentum exchange of the linked nodes on force steps, before k_fast_v on the same stream and
// from the same input buffer (neither writes it): one lane per linked node (padded index and
// link word), each bounced value 2 h_opp(i)(x) e_opp(i) summed in direction order, as the
// per-node pull does, into the node's partial slot.
template <int Q, class S>
__global__ void __launch_bounds__(256)
    k_link_force(const typename S::T *__restrict__ A, const uint32_t *__restrict__ list,
                 const uint32_t *__restrict__ links, const int *__restrict__ slot, int nlist,
                 double *__restrict__ partials, const FusedArgs a) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nlist) return;
    const uint32_t node = list[j], lk = links[j];
    const uint32_t row = node / (uint32_t)g.nxp;
    const uint32_t x = node - row * (uint32_t)g.nxp;
    // every bounced value first, so the loads are in flight together: a direction without a
    // link loads the node's own rest value instead (a line the wave reads anyway), so no load
    // sits in a branch of its own and waits there
    T val[Q];
    const size_t self0 = pel(g, 0, x, row);
    static_for<1, Q>([&](auto I) {
        constexpr int i = I;
        val[i] = A[((lk >> i) & 1u) ? pel(g, L::opp(i), x, row) : self0];
    });
    double fx = 0.0, fy = 0.0, fz = 0.0;
    static_for<1, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
        if ((lk >> i) & 1u) {
            const double m = 2.0 * (double)S::template decode<i, L>(val[i], a.sp.off[o]);
            if constexpr (ex != 0) fx -= ex * m;
            if constexpr (ey != 0) fy -= ey * m;
            if constexpr (ez != 0) fz -= ez * m;
        }
    });
    const int k = slot[j];
    partials[3 * (size_t)k] = fx;
    partials[3 * (size_t)k + 1] = fy;
    partials[3 * (size_t)k + 2] = fz;
}

// The per-node recomputation of the listed nodes (padded-layout indices, row-major order),
// after k_fast_v on the same stream: the general pull of fused_common.hpp (walls, inlet,
// outlet, periodic axes, mask bounce-back taken from the node's link word: bit i set when
// direction i pulls from a solid node), the velocity inlet, collision with partial
// bounce-back, single-value stores.  FORCE writes each node's momentum exchange (and partial
// bounce-back force) to partials[3 * slot[j]] when slot[j] >= 0; the slots are numbered in
// node order, so the later fixed-order sum does not depend on which nodes the list holds.
template <int Q, class S, bool FORCE, bool MACROS, bool SFRAC>
__global__ void __launch_bounds__(256)
    k_fix_v(const typename S::T *__restrict__ A, typename S::T *__restrict__ B, const uint32_t *__restrict__ list,
            const uint32_t *__restrict__ links, int nlist, const int *__restrict__ slot,
            const float *__restrict__ sfrac, const double *__restrict__ uin, double *__restrict__ partials,
            double *__restrict__ macros, const FusedArgs a, const DirTab tab) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nlist) return;
    const uint32_t node = list[j];
    const uint32_t row = node / (uint32_t)g.nxp;
    const int x = (int)(node - row * (uint32_t)g.nxp);
    const int y = (int)(row % (uint32_t)g.ny), z = (int)(row / (uint32_t)g.ny);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    float gv[Q], c[Q], s, u, v, w;
    pull<Q, S, FORCE, true>(A, nullptr, g, a.sp, x, y, z, node, true, false, gv, fx, fy, fz, links[j]);
    if (!g.xper && x == 0) velocity_inlet<Q, float>(gv, (float)uin[y + g.ny * z]);
    float sf = 0.f;
    if constexpr (SFRAC) sf = sfrac[node];
    collide<Q, FORCE, SFRAC>(gv, sf, a.omega, c, s, u, v, w, fx, fy, fz);
    char *const Bn = (char *)B + pel(g, 0, (uint32_t)x, row) * sizeof(T);
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        __builtin_nontemporal_store(S::template enc<i, L>(c[i], a.sp), (T *)(Bn + tab.dst[i]));
    });
    if constexpr (MACROS) {
        double *m = macros + node;
        m[0] = (double)s * (1.0 / 3.0);
        m[g.plane] = (double)u;
        m[2 * g.plane] = (double)v;
        m[3 * g.plane] = (double)w;
    }
    if constexpr (FORCE) {
        const int k = slot[j];
        if (k >= 0) {
            partials[3 * (size_t)k] = fx;
            partials[3 * (size_t)k + 1] = fy;
            partials[3 * (size_t)k + 2] = fz;
        }
    }
}

// Dense mode: when the nodes the packed path gets wrong are a large part of the domain (a
// solid fraction over a noisy field), one pass with one node per lane replaces the copies, the
// packed kernel and the fix-up: the per-node pull classifying its sources by the node flags,
// the velocity inlet and collide() as in the other kernels, so the populations are those of
// the list path bit for bit; FORCE sums the nodes' momentum exchange and partial bounce-back
// force per block into partials[3 * block_slot[block]] (blocks without a forced node: -1).
template <int Q, class S, bool FORCE, bool MACROS, bool SFRAC>
__global__ void __launch_bounds__(256)
    k_dense_v(const typename S::T *__restrict__ A, typename S::T *__restrict__ B, const uint8_t *__restrict__ flags,
              const float *__restrict__ sfrac, const double *__restrict__ uin, const int *__restrict__ block_slot,
              double *__restrict__ partials, double *__restrict__ macros, const FusedArgs a, const DirTab tab) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    const bool active = x < g.nx;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    // lanes past the row use the row's first node as a valid, unused address
    const uint32_t node = (uint32_t)(active ? x : 0) + (uint32_t)g.nxp * row;
    const uint8_t fl = active ? flags[node] : (uint8_t)FL_SOLID;
    const bool live = fl != FL_SOLID;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    float gv[Q], c[Q], s, u, v, w;
    pull<Q, S, FORCE>(A, flags, g, a.sp, x, y, z, node, live, fl == FL_NEAR, gv, fx, fy, fz);
    if (!g.xper && x == 0 && live) velocity_inlet<Q, float>(gv, (float)uin[y + g.ny * z]);
    float sf = 0.f;
    if constexpr (SFRAC) sf = live ? sfrac[node] : 0.f;
    collide<Q, FORCE, SFRAC>(gv, sf, a.omega, c, s, u, v, w, fx, fy, fz);
    if (live) {
        char *const Bn = (char *)B + pel(g, 0, (uint32_t)x, row) * sizeof(T);
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            __builtin_nontemporal_store(S::template enc<i, L>(c[i], a.sp), (T *)(Bn + tab.dst[i]));
        });
    }
    if constexpr (MACROS) {
        if (active) {
            double *m = macros + node;
            m[0] = live ? (double)s * (1.0 / 3.0) : 0.0;
            m[g.plane] = live ? (double)u : 0.0;
            m[2 * g.plane] = live ? (double)v : 0.0;
            m[3 * g.plane] = live ? (double)w : 0.0;
        }
    }
    if constexpr (FORCE) {
        const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const int slot = block_slot[bid];
        if (slot >= 0) {   // uniform over the block
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

// One step on s: the packed path everywhere, then the listed nodes (dense mode: k_dense_v alone).
template <int Q, class S, int V, bool FORCE, bool MACROS>
static hipError_t fast_launch(const StepBufs &b, const FusedArgs &a, hipStream_t s) {
    using T = typename S::T;
    const Shape sh = lane_shape(a.g, V);
    const DirTab tab = dir_tab<Q, T>(a.g);
    hipError_t e;
    if (b.dense) {
        const Shape s1 = lane_shape(a.g, 1);
        const dim3 grid(s1.gx, a.g.ny, a.g.nz);
        if (b.ev_bulk[0] && (e = hipEventRecord(b.ev_bulk[0], s)) != hipSuccess) return e;
        if (b.sfrac)
            hipLaunchKernelGGL((k_dense_v<Q, S, FORCE, MACROS, true>), grid, dim3(s1.bx), 0, s, (const T *)b.A, (T *)b.B,
                               b.flags, (const float *)b.sfrac, b.uin, b.block_slot, b.partials, b.macros, a, tab);
        else
            hipLaunchKernelGGL((k_dense_v<Q, S, FORCE, MACROS, false>), grid, dim3(s1.bx), 0, s, (const T *)b.A, (T *)b.B,
                               b.flags, (const float *)nullptr, b.uin, b.block_slot, b.partials, b.macros, a, tab);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if (b.ev_bulk[1] && (e = hipEventRecord(b.ev_bulk[1], s)) != hipSuccess) return e;
        return hipSuccess;
    }
    if (FORCE && b.nlink > 0) {
        hipLaunchKernelGGL((k_link_force<Q, S>), dim3((unsigned)((b.nlink + 255) / 256)), dim3(256), 0, s, (const T *)b.A,
                           b.link, b.linklk, b.link_slot, b.nlink, b.partials, a);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    // the aligned pull needs the whole row in one block
    const unsigned lds = b.fast_lds;
    if (b.ev_bulk[0] && (e = hipEventRecord(b.ev_bulk[0], s)) != hipSuccess) return e;
    if (sizeof(T) == 2 && V == 2 && sh.gx == 1 && b.aligned_pull)
        hipLaunchKernelGGL((k_fast_v<Q, S, V, MACROS, (sizeof(T) == 2 && V == 2)>), dim3(sh.gx, a.g.ny, a.g.nz),
                           dim3(sh.bx), lds, s, (const T *)b.A, (T *)b.B, b.uin, b.macros, b.rowwin, b.wlk, a, tab);
    else
        hipLaunchKernelGGL((k_fast_v<Q, S, V, MACROS, false>), dim3(sh.gx, a.g.ny, a.g.nz), dim3(sh.bx), lds, s,
                           (const T *)b.A, (T *)b.B, b.uin, b.macros, b.rowwin, b.wlk, a, tab);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (b.ev_bulk[1] && (e = hipEventRecord(b.ev_bulk[1], s)) != hipSuccess) return e;
    if (b.nfix == 0) return hipSuccess;
    const dim3 grid((unsigned)((b.nfix + 255) / 256));
    if (b.sfrac)
        hipLaunchKernelGGL((k_fix_v<Q, S, FORCE, MACROS, true>), grid, dim3(256), 0, s, (const T *)b.A, (T *)b.B, b.fix,
                           b.fixlk, b.nfix, b.block_slot, (const float *)b.sfrac, b.uin, b.partials, b.macros, a, tab);
    else
        hipLaunchKernelGGL((k_fix_v<Q, S, FORCE, MACROS, false>), grid, dim3(256), 0, s, (const T *)b.A, (T *)b.B, b.fix,
                           b.fixlk, b.nfix, b.block_slot, (const float *)nullptr, b.uin, b.partials, b.macros, a, tab);
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
