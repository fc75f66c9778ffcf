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

// Pull of the Q populations of one node from buffer A with the plain-mode boundary rules
// (first match wins):
//   x == 0 and e_ix > 0 (inlet, x not periodic)   -> Zou-He unknown (left 0, filled later)
//   source outside y/z (walls, not periodic)       -> half-way bounce-back g_i = h_opp(i)(node)
//   source x == nx (outlet, x not periodic)        -> zero gradient, source clamped to nx-1
//   source is solid (mask)                         -> half-way bounce-back, momentum exchange
//   otherwise                                      -> g_i = h_i(node - e_i) (periodic wrap)
// FORCE accumulates the momentum exchange 2 h_opp(i) e_opp(i) of the mask links.
template <int Q, class S, bool FORCE>
__device__ __forceinline__ void pull(const typename S::T *__restrict__ A, const uint8_t *__restrict__ flags,
                                     const Geo &g, const StoreParams &sp, int x, int y, int z, uint32_t node,
                                     bool live, bool near, typename S::C *gv, double &fx, double &fy, double &fz) {
    using L = Lat<Q>;
    using T = typename S::T;
    using C = typename S::C;
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
        int sx = x - ex, sy = y - ey, sz = z - ez;
        bool zh = false, bb = false, link = false;
        if constexpr (ex > 0) {
            if (x == 0) {
                if (g.xper) sx = g.nx - 1;
                else { zh = true; sx = 0; }
            }
        }
        if constexpr (ex < 0) {
            if (x == g.nx - 1) sx = g.xper ? 0 : g.nx - 1;
        }
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
        const uint32_t src = (uint32_t)sx + (uint32_t)g.nxp * ((uint32_t)sy + (uint32_t)g.ny * (uint32_t)sz);
        if (near && !zh && !bb) {
            if (flags[src] == FL_SOLID) bb = link = true;
        }
        C v = 0;
        if (live && !zh) {
            const T *p = bb ? (A + (size_t)o * g.plane + node) : (A + (size_t)i * g.plane + src);
            v = S::template decode<i, L>(*p, bb ? sp.off[o] : sp.off[i]);
        }
        gv[i] = v;
        if constexpr (FORCE) {
            if (link) {
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
