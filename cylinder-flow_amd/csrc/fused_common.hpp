// fused_common.hpp -- device pieces shared by the forward (plain_kernels.hip) and the
// adjoint (adjoint_kernels.hip) kernels, so both classify every pulled entry the same way.
#pragma once
#include "internal.hpp"
#include "lattice.hpp"

namespace lbm {

// ---- counter-based perturbation RNG: SplitMix64 finaliser of (seed ^ SplitMix64(index)) ----
__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double uniform_pm1(uint64_t seed, uint64_t index) {
    return (double)(splitmix64(seed ^ splitmix64(index)) >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// a / b rounded to nearest, given y = 1/b rounded to nearest: q = RN(a y), r = a - b q (exact
// with one FMA), RN(q + r y).  Markstein's correction (P. W. Markstein, IBM J. Res. Dev. 34(1),
// 1990; the final step of IA-64 division): with a correctly rounded reciprocal and q within an
// ulp of a/b it yields the correctly rounded quotient for quotients in the normal range, the same
// bits as IEEE division in three FMA-unit operations instead of the general sequence.  Used for
// the FP64 path's divisions by 3 and by tau (checked against IEEE division on the CPU:
// tests/div_check.c).
__device__ __forceinline__ double div_cr(double a, double b, double y) {
    const double q = a * y;
    const double r = __builtin_fma(-q, b, a);
    return __builtin_fma(r, y, q);
}

// shifted incompressible equilibrium in the reference's rounding order
// (3D MPI+KOKKOS+MPIIO/lbm.cpp:263-267):  t*p*3 + t*(3 edu + 4.5 edu^2 - 1.5 udu)
template <int Q, int I>
__device__ __forceinline__ double feq_ref(double p, double u, double v, double w) {
    using L = Lat<Q>;
    double edu = (double)L::ex(I) * u + (double)L::ey(I) * v;
    double udu = u * u + v * v;
    if constexpr (L::D == 3) {
        edu = edu + (double)L::ez(I) * w;
        udu = udu + w * w;
    }
    const double eu2 = edu * edu;
    return L::w(I) * p * 3.0 + L::w(I) * (3.0 * edu + 4.5 * eu2 - 1.5 * udu);
}

// Zou-He inlet for the shifted incompressible model, inlet velocity (U, 0, 0):
//   g_i = g_opp(i) + 6 w_i U - 18 w_i (e_iy N_y + e_iz N_z),  N_t = sum_{e_kx = 0} g_k e_kt
// for every unknown i (e_ix > 0); the result carries momentum (U, 0, 0) exactly.
template <int Q, class C>
__device__ __forceinline__ void zou_he(C *g, C U) {
    using L = Lat<Q>;
    C Ny = 0, Nz = 0;
    static_for<0, Q>([&](auto K) {
        constexpr int k = K;
        if constexpr (L::ex(k) == 0) {
            if constexpr (L::ey(k) > 0) Ny += g[k];
            else if constexpr (L::ey(k) < 0) Ny -= g[k];
            if constexpr (L::ez(k) > 0) Nz += g[k];
            else if constexpr (L::ez(k) < 0) Nz -= g[k];
        }
    });
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (L::ex(i) > 0) {
            C tr = 0;
            if constexpr (L::ey(i) > 0) tr += Ny;
            else if constexpr (L::ey(i) < 0) tr -= Ny;
            if constexpr (L::ez(i) > 0) tr += Nz;
            else if constexpr (L::ez(i) < 0) tr -= Nz;
            g[i] = (g[L::opp(i)] + (C)(6.0 * L::w(i)) * U) - (C)(18.0 * L::w(i)) * tr;
        }
    });
}

// Regularised velocity inlet (Latt & Chopard's regularisation applied to the Zou-He node):
// after zou_he(), every population of the inlet node is rebuilt from S = sum g, the second
// moment P_ab = sum g e_a e_b - S/3 delta_ab and the prescribed velocity (U, 0, 0):
//   g_i = w_i (S + 3 e_ix U + 4.5 Q_i:P),   Q_i = e_i e_i - delta/3,
// i.e. the equilibrium at (S/3, U) plus the projection of the non-equilibrium stress.  S, U
// and P are reproduced exactly; the non-hydrodynamic moments that make plain Zou-He unstable
// with BGK at tau near 1/2 (tau = 0.53 at Re 100, u 0.05, D 20) are filtered out.
// Operation order is fixed (the FP64 CPU restatement follows it term by term).
template <int Q, class C>
__device__ __forceinline__ void regularize_inlet(C *g, C U) {
    using L = Lat<Q>;
    C S = 0, Mxx = 0, Myy = 0, Mzz = 0, Mxy = 0, Mxz = 0, Myz = 0;
    static_for<0, Q>([&](auto K) {
        constexpr int k = K;
        constexpr int ex = L::ex(k), ey = L::ey(k), ez = L::ez(k);
        S = S + g[k];
        if constexpr (ex != 0) Mxx = Mxx + g[k];
        if constexpr (ey != 0) Myy = Myy + g[k];
        if constexpr (ez != 0) Mzz = Mzz + g[k];
        if constexpr (ex * ey > 0) Mxy = Mxy + g[k];
        else if constexpr (ex * ey < 0) Mxy = Mxy - g[k];
        if constexpr (ex * ez > 0) Mxz = Mxz + g[k];
        else if constexpr (ex * ez < 0) Mxz = Mxz - g[k];
        if constexpr (ey * ez > 0) Myz = Myz + g[k];
        else if constexpr (ey * ez < 0) Myz = Myz - g[k];
    });
    const C S3 = S * (C)(1.0 / 3.0);
    const C Pxx = Mxx - S3, Pyy = Myy - S3, Pzz = Mzz - S3;
    C tr = Pxx + Pyy;
    if constexpr (L::D == 3) tr = tr + Pzz;
    const C d = tr * (C)(1.0 / 3.0);
    const C U3 = (C)3 * U;
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
        C q = 0, o = 0;
        if constexpr (ex != 0) q = q + Pxx;
        if constexpr (ey != 0) q = q + Pyy;
        if constexpr (ez != 0) q = q + Pzz;
        q = q - d;
        if constexpr (ex * ey > 0) o = o + Mxy;
        else if constexpr (ex * ey < 0) o = o - Mxy;
        if constexpr (ex * ez > 0) o = o + Mxz;
        else if constexpr (ex * ez < 0) o = o - Mxz;
        if constexpr (ey * ez > 0) o = o + Myz;
        else if constexpr (ey * ez < 0) o = o - Myz;
        C t = S;
        if constexpr (ex > 0) t = S + U3;
        else if constexpr (ex < 0) t = S - U3;
        g[i] = (C)L::w(i) * (t + ((C)4.5 * q + (C)9 * o));
    });
}

// The plain-mode velocity inlet: Zou-He completion of the unknowns, then regularisation.
template <int Q, class C>
__device__ __forceinline__ void velocity_inlet(C *g, C U) {
    zou_he<Q, C>(g, U);
    regularize_inlet<Q, C>(g, U);
}

// Block-wide sum of three doubles in a fixed order (wave butterfly, then the waves in order);
// the result is valid in thread 0.  `red` holds 3 doubles per wave of the block.
// wave butterfly of three doubles: every lane ends with the wave's sums
__device__ __forceinline__ void wave_sum3(double &a, double &b, double &c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
}

template <int NW>
__device__ __forceinline__ void block_sum3(double &a, double &b, double &c, double *red) {
    wave_sum3(a, b, c);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[3 * wv] = a;
        red[3 * wv + 1] = b;
        red[3 * wv + 2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        a = red[0];
        b = red[1];
        c = red[2];
        for (int k = 1; k < nw && k < NW; k++) {
            a += red[3 * k];
            b += red[3 * k + 1];
            c += red[3 * k + 2];
        }
    }
}

// Pull of the Q populations of one node from buffer A with the plain-mode boundary rules
// (first match wins):
//   x == 0 and e_ix > 0 (inlet, x not periodic)   -> Zou-He unknown (left 0, filled later)
//   source outside y/z (walls, not periodic)       -> half-way bounce-back g_i = h_opp(i)(node)
//   source x == nx (outlet, x not periodic)        -> zero gradient, source clamped to nx-1
//   source is solid (mask)                         -> half-way bounce-back, momentum exchange
//   otherwise                                      -> g_i = h_i(node - e_i) (periodic wrap)
// FORCE accumulates the momentum exchange 2 h_opp(i) e_opp(i) of the mask links.
// Which sources are solid: LINKS = false gathers the flags of the Q sources when `near`;
// LINKS = true takes them from `lk` (bit i: direction i pulls from a solid node), computed
// once on the host with the same rules, so no flag is read.
// `node` (= x + nxp*row) must be a valid node even when the lane is not live; every load below
// is issued unconditionally from a valid address (no data-dependent branches), so the Q loads
// of a lane are in flight together.
template <int Q, class S, bool FORCE, bool LINKS = false>
__device__ __forceinline__ void pull(const typename S::T *__restrict__ A, const uint8_t *__restrict__ flags,
                                     const Geo &g, const StoreParams &sp, int x, int y, int z, uint32_t node,
                                     bool live, bool near, typename S::C *gv, double &fx, double &fy, double &fz,
                                     uint32_t lk = 0) {
    using L = Lat<Q>;
    using T = typename S::T;
    using C = typename S::C;
    uint32_t src[Q], srow[Q];   // source node and its row
    bool zh[Q], wall[Q], solid[Q];
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
        int sx = x - ex, sy = y - ey, sz = z - ez;
        bool z_ = false, w_ = false;
        if constexpr (ex > 0) {
            if (x == 0) {
                if (g.xper) sx = g.nx - 1;
                else z_ = true;
            }
        }
        if constexpr (ex < 0) {
            if (x == g.nx - 1) sx = g.xper ? 0 : g.nx - 1;
        }
        if constexpr (ey != 0) {
            if ((unsigned)sy >= (unsigned)g.ny) {
                if (g.yper) sy = ey > 0 ? g.ny - 1 : 0;
                else w_ = true;
            }
        }
        if constexpr (ez != 0) {
            if ((unsigned)sz >= (unsigned)g.nz) {
                if (g.zper) sz = ez > 0 ? g.nz - 1 : 0;
                else w_ = true;
            }
        }
        zh[i] = z_;
        wall[i] = w_ && !z_;
        const bool plain = live && !z_ && !w_;
        srow[i] = plain ? (uint32_t)sy + (uint32_t)g.ny * (uint32_t)sz : row;
        src[i] = plain ? (uint32_t)sx + (uint32_t)g.nxp * srow[i] : node;
        solid[i] = false;
    });
    if constexpr (LINKS) {
        static_for<0, Q>([&](auto I) { solid[I] = !zh[I] && !wall[I] && ((lk >> (int)I) & 1u); });
    } else if (near) {   // only nodes next to the body look at their sources' flags
        uint8_t fl[Q];
        static_for<0, Q>([&](auto I) { fl[I] = flags[src[I]]; });
        static_for<0, Q>([&](auto I) { solid[I] = !zh[I] && !wall[I] && fl[I] == FL_SOLID; });
    }
    T raw[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int o = L::opp(i);
        const bool bb = wall[i] || solid[i];
        // element i*ds + row*rs + x, with node = x + nxp*row
        raw[i] = bb ? A[(size_t)o * g.ds + node + (size_t)row * (g.rs - g.nxp)]
                    : A[(size_t)i * g.ds + src[i] + (size_t)srow[i] * (g.rs - g.nxp)];
    });
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
        const bool bb = wall[i] || solid[i];
        const C v = (live && !zh[i]) ? S::template decode<i, L>(raw[i], bb ? sp.off[o] : sp.off[i]) : (C)0;
        gv[i] = v;
        if constexpr (FORCE) {
            if (solid[i]) {
                // momentum exchange of the bounced population h_opp(i): 2 h e_opp(i)
                const double m = 2.0 * (double)v;
                if constexpr (ex != 0) fx -= ex * m;
                if constexpr (ey != 0) fy -= ey * m;
                if constexpr (ez != 0) fz -= ez * m;
            }
        }
    });
}

}  // namespace lbm
