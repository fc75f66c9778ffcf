// adjoint_kernels.hip -- discrete adjoint of the plain-mode step (no reference counterpart).
//
// Forward step m (1-based) maps H_{m-1} (post-collision populations) to G_m (pulled
// populations after the boundary rules and the regularised Zou-He inlet) and H_m = (1-s) C(G_m) + s G_m[opp]
// (BGK collision C, partial bounce-back with solid fraction s).  Its drag is
//   Fx_m = sum over mask links of 2 h_opp(i) e_opp(i),x   (from H_{m-1})
//        + sum over nodes of 2 s u_x(G_m)                 (partial bounce-back).
// For J = sum_m alpha_m Fx_m the reverse sweep carries mu_m = dJ/dG_m.  One kernel per step
// and node a:
//   1. lambda(a) = dJ/dH_m(a): gather of mu_{m+1} over every pulled entry that read a
//      population of node a (transpose of the pull, evaluated as a pull along +e),
//      plus the link-force term of step m+1;
//   2. recompute G_m(a) from H_{m-1} with the forward's own pull and velocity inlet;
//   3. mu_m(a) = (collision + partial bounce-back)^T lambda + alpha_m dFx/dG, then the
//      transposes of the regularisation and of Zou-He at inlet nodes; accumulate dJ/ds(a)
//      and dJ/du_in(row).
// Traffic per node: Q reads of H, Q reads of mu_{m+1}, Q writes of mu_m.  H is a population
// buffer of the forward (read through its pull); the mu arrays use the populations' layout,
// mu[pel(i, x, row)] (row-interleaved by default).
#include "fused_common.hpp"

namespace lbm {

// a*b + c rounded once (FP32 or FP64)
template <class C>
__device__ __forceinline__ C fmac(C a, C b, C c) {
    if constexpr (sizeof(C) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// e_i . u with the zero components of e_i skipped
template <class L, int i, class C>
__device__ __forceinline__ C edot(const C *u) {
    C r = 0;
    bool first = true;
    static_for<0, 3>([&](auto K) {
        constexpr int k = K;
        constexpr int e = k == 0 ? L::ex(i) : (k == 1 ? L::ey(i) : L::ez(i));
        if constexpr (e != 0) {
            const C t = e > 0 ? u[k] : -u[k];
            r = first ? t : r + t;
            first = false;
        }
    });
    return r;
}

// Transposes of the inlet's regularisation and Zou-He completion applied to m = mu_m at an inlet
// node (x = 0, x not periodic), in place, and their contribution to dJ/du_in of its row.
template <int Q, class S>
__device__ __forceinline__ void adj_inlet_transpose(typename S::C *m, double &du) {
    using L = Lat<Q>;
    using C = typename S::C;
    // regularisation g_i = w_i (S + 3 e_ix U + 4.5 Q_i:P) is g -> A g + 3 w e_x U with
    // A_ij = w_i (1 + 4.5 Q_i:Q_j), so (A^T m)_j = M0 + 4.5 Q_j:T, M0 = sum w_i m_i,
    // T = sum w_i m_i Q_i (same structure as the forward with S -> M0, P -> T)
    {
        C M0 = 0, Rxx = 0, Ryy = 0, Rzz = 0, Rxy = 0, Rxz = 0, Ryz = 0;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
            const C wm = (C)L::w(i) * m[i];
            M0 += wm;
            if constexpr (ex != 0) Rxx += wm;
            if constexpr (ey != 0) Ryy += wm;
            if constexpr (ez != 0) Rzz += wm;
            if constexpr (ex * ey != 0) Rxy += (C)(ex * ey) * wm;
            if constexpr (ex * ez != 0) Rxz += (C)(ex * ez) * wm;
            if constexpr (ey * ez != 0) Ryz += (C)(ey * ez) * wm;
            if constexpr (ex != 0) du += (double)wm * (3.0 * ex);
        });
        const C M3 = M0 * (C)(1.0 / 3.0);
        const C Txx = Rxx - M3, Tyy = Ryy - M3, Tzz = Rzz - M3;
        const C d = (L::D == 3 ? Txx + Tyy + Tzz : Txx + Tyy) * (C)(1.0 / 3.0);
        static_for<0, Q>([&](auto J) {
            constexpr int j = J;
            constexpr int ex = L::ex(j), ey = L::ey(j), ez = L::ez(j);
            C q = -d, o = 0;
            if constexpr (ex != 0) q += Txx;
            if constexpr (ey != 0) q += Tyy;
            if constexpr (ez != 0) q += Tzz;
            if constexpr (ex * ey != 0) o += (C)(ex * ey) * Rxy;
            if constexpr (ex * ez != 0) o += (C)(ex * ez) * Rxz;
            if constexpr (ey * ez != 0) o += (C)(ey * ez) * Ryz;
            m[j] = M0 + ((C)4.5 * q + (C)9 * o);
        });
    }
    // Zou-He g_i = g_opp(i) + 6 w_i U - 18 w_i (e_iy N_y + e_iz N_z), N_t = sum_{e_kx=0} g_k e_kt
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        if constexpr (L::ex(i) > 0) {
            const C mi = m[i];
            m[L::opp(i)] += mi;
            static_for<0, Q>([&](auto K) {
                constexpr int k = K;
                constexpr int dot = L::ey(i) * L::ey(k) + L::ez(i) * L::ez(k);
                if constexpr (L::ex(k) == 0 && dot != 0) m[k] += mi * (C)(-18.0 * L::w(i) * dot);
            });
            du += (double)mi * (6.0 * L::w(i));
            m[i] = 0;
        }
    });
}

// index of the direction with components (ex, ey, ez)
template <class L>
__host__ __device__ constexpr int dir_of(int ex, int ey, int ez) {
    for (int i = 0; i < (L::D == 3 ? 27 : 9); i++)
        if (L::ex(i) == ex && L::ey(i) == ey && L::ez(i) == ez) return i;
    return -1;
}

// Sums over the lattice taken one axis at a time: a line of three values v(-1), v(0), v(+1)
// along an axis becomes v(+1) + v(-1) + c v(0) (the sum, with the weight ratio c), v(+1) - v(-1)
// (first moment) and v(+1) + v(-1) (second moment).
//
// moments01: rho = sum g, j = sum g e (c = 1).
template <class L, class C>
__device__ __forceinline__ void moments01(const C *g, C &rho, C *j) {
    constexpr int NZ = L::D == 3 ? 3 : 1;
    C a0[3], a1[3], b0[3];
    static_for<0, NZ>([&](auto ZZ) {
        constexpr int ez = L::D == 3 ? (int)ZZ - 1 : 0;
        C s0[3], s1[3];
        static_for<0, 3>([&](auto YY) {
            constexpr int ey = (int)YY - 1;
            const C fp = g[dir_of<L>(1, ey, ez)], fm = g[dir_of<L>(-1, ey, ez)];
            s0[YY] = (fp + fm) + g[dir_of<L>(0, ey, ez)];
            s1[YY] = fp - fm;
        });
        a0[ZZ] = (s0[2] + s0[0]) + s0[1];
        a1[ZZ] = s0[2] - s0[0];
        b0[ZZ] = (s1[2] + s1[0]) + s1[1];
    });
    if constexpr (L::D == 3) {
        rho = (a0[2] + a0[0]) + a0[1];
        j[0] = (b0[2] + b0[0]) + b0[1];
        j[1] = (a1[2] + a1[0]) + a1[1];
        j[2] = a0[2] - a0[0];
    } else {
        rho = a0[0];
        j[0] = b0[0];
        j[1] = a1[0];
        j[2] = 0;
    }
}

// the product form of the weights that wmoments relies on
template <class L>
constexpr bool weights_factor() {
    for (int i = 0; i < (L::D == 3 ? 27 : 9); i++) {
        double p = 1.0;
        const int e[3] = {L::ex(i), L::ey(i), L::ez(i)};
        for (int a = 0; a < L::D; a++) p *= e[a] == 0 ? 4.0 / 6.0 : 1.0 / 6.0;
        const double d = p - L::w(i);
        if (d > 1e-15 || d < -1e-15) return false;
    }
    return true;
}
static_assert(weights_factor<Lat<9>>() && weights_factor<Lat<27>>(), "lattice weights are products over the axes");

// wmoments: moments of w_i f_i to second order, M0 = sum w f, M1 = sum w f e, M2 = sum w f e e^T
// as (xx, yy, zz, xy, xz, yz).  The weights are products over the axes, w(e) = prod_a w1(e_a) with
// w1(0) = 4/6 and w1(+-1) = 1/6, so each axis pass uses c = 4 and the factor 6^-D comes last.
template <class L, class C>
__device__ __forceinline__ void wmoments(const C *f, C &M0, C *M1, C *M2) {
    constexpr int NZ = L::D == 3 ? 3 : 1;
    C a0[3], a1[3], a2[3], b0[3], b1[3], c0[3];
    static_for<0, NZ>([&](auto ZZ) {
        constexpr int ez = L::D == 3 ? (int)ZZ - 1 : 0;
        C s0[3], s1[3], s2[3];
        static_for<0, 3>([&](auto YY) {
            constexpr int ey = (int)YY - 1;
            const C fp = f[dir_of<L>(1, ey, ez)], fm = f[dir_of<L>(-1, ey, ez)];
            const C t = fp + fm;
            s0[YY] = fmac((C)4, f[dir_of<L>(0, ey, ez)], t);
            s1[YY] = fp - fm;
            s2[YY] = t;
        });
        const C t0 = s0[2] + s0[0], t1 = s1[2] + s1[0];
        a0[ZZ] = fmac((C)4, s0[1], t0);
        a1[ZZ] = s0[2] - s0[0];
        a2[ZZ] = t0;
        b0[ZZ] = fmac((C)4, s1[1], t1);
        b1[ZZ] = s1[2] - s1[0];
        c0[ZZ] = fmac((C)4, s2[1], s2[2] + s2[0]);
    });
    C r[10];
    if constexpr (L::D == 3) {
        const C t = a0[2] + a0[0];
        r[0] = fmac((C)4, a0[1], t);                    // M0
        r[1] = fmac((C)4, b0[1], b0[2] + b0[0]);        // x
        r[2] = fmac((C)4, a1[1], a1[2] + a1[0]);        // y
        r[3] = a0[2] - a0[0];                           // z
        r[4] = fmac((C)4, c0[1], c0[2] + c0[0]);        // xx
        r[5] = fmac((C)4, a2[1], a2[2] + a2[0]);        // yy
        r[6] = t;                                       // zz
        r[7] = fmac((C)4, b1[1], b1[2] + b1[0]);        // xy
        r[8] = b0[2] - b0[0];                           // xz
        r[9] = a1[2] - a1[0];                           // yz
    } else {
        r[0] = a0[0];
        r[1] = b0[0];
        r[2] = a1[0];
        r[3] = 0;
        r[4] = c0[0];
        r[5] = a2[0];
        r[6] = 0;
        r[7] = b1[0];
        r[8] = 0;
        r[9] = 0;
    }
    constexpr double sc = L::D == 3 ? 1.0 / 216.0 : 1.0 / 36.0;
    M0 = r[0] * (C)sc;
    static_for<0, 3>([&](auto K) { M1[K] = r[1 + K] * (C)sc; });
    static_for<0, 6>([&](auto K) { M2[K] = r[4 + K] * (C)sc; });
}

// Byte offsets that the straight-line path adds to its row's first element: direction i's row
// segment is at i * plane + row[(e_y + 1) + 3 (e_z + 1)] for the gather of mu_{m+1} along +e_i and
// at i * plane - row[...] for the pull of H along -e_i (H and mu elements have one size).
struct AdjTab {
    int64_t row[9];
    int64_t plane;
};

template <int Q, class C, class T>
static AdjTab adj_tab(const Geo &g) {
    static_assert(sizeof(C) == sizeof(T), "one element size for H and mu");
    AdjTab t{};
    for (int ez = -1; ez <= 1; ez++)
        for (int ey = -1; ey <= 1; ey++)
            t.row[(ey + 1) + 3 * (ez + 1)] = ((int64_t)ey + (int64_t)g.ny * ez) * (int64_t)g.rs * (int64_t)sizeof(C);
    t.plane = (int64_t)g.ds * (int64_t)sizeof(C);
    return t;
}

// Rows whose nodes are all fluid and have no solid neighbour ("clean", flagged per row on the
// host) and that are not on a periodic y/z edge take a straight-line path: every source and
// destination row of a direction is a scalar of the block (at a y/z wall, the node's own row and
// the opposite direction: half-way bounce-back), and the x edges (inlet, outlet, periodic x) are
// handled with in-row clamped indices and selects, so every load of the path is issued
// unconditionally from an address inside its row.  The loads are those of the general path, and
// the arithmetic is shared, so both give the same bits.  Other rows take the general path;
// GENERAL = false compiles the straight-line path alone (the host launches it when every row
// qualifies), with the smaller register budget of that path.
template <int Q, class S, bool SFRAC, bool GENERAL>
__global__ void __launch_bounds__(256)
    k_adj_step(const typename S::T *__restrict__ H, const typename S::C *__restrict__ mun,
               typename S::C *__restrict__ mu, const uint8_t *__restrict__ flags, const typename S::C *__restrict__ sfrac,
               const double *__restrict__ uin, double *__restrict__ dJds, double *__restrict__ dJdu,
               const uint8_t *__restrict__ rowclean, const AdjArgs a, const AdjTab tab) {
    using L = Lat<Q>;
    using C = typename S::C;
    const Geo &g = a.f.g;
    // the plane z = nz - 1 first: at a periodic z edge its rows take the general path, whose
    // dependent loads make its blocks the slowest; dispatched last they would form the tail
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    const int z = blockIdx.z == 0 ? g.nz - 1 : (int)blockIdx.z - 1;
    if (x >= g.nx) return;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    const uint32_t node = (uint32_t)x + (uint32_t)g.nxp * row;
    // uniform over the block (a block lies in one row)
    const bool ylo = y == 0, yhi = y == g.ny - 1, zlo = L::D == 3 && z == 0, zhi = L::D == 3 && z == g.nz - 1;
    const bool fast = !GENERAL || (rowclean[row] && !(g.yper && (ylo || yhi)) && !(g.zper && (zlo || zhi)));
    const uint8_t fl = fast ? (uint8_t)FL_FLUID : flags[node];
    if (fl == FL_SOLID) {
        static_for<0, Q>([&](auto I) { mu[pel(g, I, (uint32_t)x, row)] = (C)0; });
        return;
    }
    const C an = (C)a.alpha_next;
    const bool xlo = !g.xper && x == 0, xhi = !g.xper && x == g.nx - 1;
    C lam[Q];
    C gv[Q];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (fast) {
        // Every load of the path is issued before any value is used, from the scalar base of
        // its direction's row plus one of three lane offsets (x - 1, x, x + 1 after the x edge
        // rule), so the 2Q + Q/3 loads are in flight together.
        // ---- 1. lambda: entry (node, opp i) pulls from node + e_i, entry (d, i) from node;
        //         the outlet node also feeds (nx-1, y + e_y, z + e_z) along e_ix < 0 ----
        // ---- 2. recompute G_m(node): the forward pull (inlet unknowns 0, outlet clamp) ----
        const uint32_t ux = (uint32_t)x;
        const uint32_t xm = x > 0 ? ux - 1u : (g.xper ? (uint32_t)g.nx - 1u : ux);
        const uint32_t xp = x + 1 < g.nx ? ux + 1u : (g.xper ? 0u : ux);
        const bool xlive_m = x > 0 || g.xper, xlive_p = x + 1 < g.nx || g.xper;
        const char *const Mrow = (const char *)mun + (size_t)row * g.rs * sizeof(C);
        const char *const Hrow = (const char *)H + (size_t)row * g.rs * sizeof(typename S::T);
        // walls (a clean row at a periodic edge takes the general path): wm, the gather's source
        // row y + e_y, z + e_z is outside; wh, the pull's source row y - e_y, z - e_z is outside.
        // Either reads the node's own row, direction opp(i), at its own x.
        auto wall_at = [&](int ey, int ez) {
            return (ey > 0 && yhi) || (ey < 0 && ylo) || (ez > 0 && zhi) || (ez < 0 && zlo);
        };
        C rl[Q], rx[Q];
        typename S::T rh[Q];
        bool wm[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), o = L::opp(i);
            wm[i] = wall_at(L::ey(i), L::ez(i));
            const uint32_t bo = wm[i] ? ux * (uint32_t)sizeof(C)
                                      : (uint32_t)(ex > 0 ? xp : (ex < 0 ? xm : ux)) * (uint32_t)sizeof(C);
            const int64_t ro = wm[i] ? (int64_t)o * tab.plane : (int64_t)i * tab.plane + tab.row[(L::ey(i) + 1) + 3 * (L::ez(i) + 1)];
            rl[i] = *(const C *)(Mrow + ro + bo);
            // the outlet term's value, from the same row segment at the lane's own x (issued next
            // to the main load: later, the segment's lines are gone from L2 and fetched again)
            if constexpr (ex < 0) rx[i] = *(const C *)(Mrow + ro + ux * (uint32_t)sizeof(C));
        });
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), o = L::opp(i);
            const bool wh = wall_at(-L::ey(i), -L::ez(i));
            using T = typename S::T;
            const uint32_t bo = wh ? ux * (uint32_t)sizeof(T)
                                   : (uint32_t)(ex > 0 ? xm : (ex < 0 ? xp : ux)) * (uint32_t)sizeof(T);
            const int64_t ro = wh ? (int64_t)o * tab.plane : (int64_t)i * tab.plane - tab.row[(L::ey(i) + 1) + 3 * (L::ez(i) + 1)];
            rh[i] = *(const T *)(Hrow + ro + bo);
        });
        // keep the selects below after the last load: scheduled among the loads they make the
        // wave wait for the first loads before it issues the rest
        __builtin_amdgcn_sched_barrier(0);
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), o = L::opp(i);
            // a wall entry has no x edge rule (beyond the outlet it still bounces back), but an
            // inlet unknown stays zero and the outlet term needs a source row in the domain
            if constexpr (ex > 0) lam[i] = (wm[i] || xlive_p) ? rl[i] : (C)0;
            if constexpr (ex < 0) {
                const C l = xlive_m ? rl[i] : (C)0;
                lam[i] = (xhi && !wm[i]) ? l + rx[i] : l;
            }
            if constexpr (ex == 0) lam[i] = rl[i];
            const bool wh = wall_at(-L::ey(i), -L::ez(i));
            const C v = S::template decode<i, L>(rh[i], wh ? a.f.sp.off[o] : a.f.sp.off[i]);
            if constexpr (ex > 0) gv[i] = xlo ? (C)0 : v;
            else gv[i] = v;
        });
    } else if constexpr (GENERAL) {
    // ---- 1. lambda = dJ/dH_m(node) ----
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
        C l = 0;
        // entry (node, opp i) pulls from d = node + e_i; entry (d, i) pulls from node
        bool skip = false;
        if constexpr (ex < 0) skip = !g.xper && x == 0;   // (node, opp i) is a Zou-He unknown
        if (!skip) {
            int dx = x + ex, dy = y + ey, dz = z + ez;
            bool wall = false, out = false;
            if constexpr (ey != 0) {
                if ((unsigned)dy >= (unsigned)g.ny) {
                    if (g.yper) dy = ey > 0 ? 0 : g.ny - 1;
                    else wall = true;
                }
            }
            if constexpr (ez != 0) {
                if ((unsigned)dz >= (unsigned)g.nz) {
                    if (g.zper) dz = ez > 0 ? 0 : g.nz - 1;
                    else wall = true;
                }
            }
            if constexpr (ex != 0) {
                if ((unsigned)dx >= (unsigned)g.nx) {
                    if (g.xper) dx = ex > 0 ? 0 : g.nx - 1;
                    else out = true;
                }
            }
            if (wall) {
                l = mun[pel(g, o, (uint32_t)x, row)];
            } else if (!out) {
                const uint32_t drow = (uint32_t)dy + (uint32_t)g.ny * (uint32_t)dz;
                const uint32_t d = (uint32_t)dx + (uint32_t)g.nxp * drow;
                if (fl == FL_NEAR && flags[d] == FL_SOLID) l = mun[pel(g, o, (uint32_t)x, row)] + an * (C)(2 * ex);
                else l = mun[pel(g, i, (uint32_t)dx, drow)];
            }
        }
        // outlet: (nx-1, y+e_y, z+e_z) pulls h_i from this node when e_ix < 0
        if constexpr (ex < 0) {
            if (!g.xper && x == g.nx - 1) {
                int dy = y + ey, dz = z + ez;
                bool ok = true;
                if constexpr (ey != 0) {
                    if ((unsigned)dy >= (unsigned)g.ny) {
                        if (g.yper) dy = ey > 0 ? 0 : g.ny - 1;
                        else ok = false;
                    }
                }
                if constexpr (ez != 0) {
                    if ((unsigned)dz >= (unsigned)g.nz) {
                        if (g.zper) dz = ez > 0 ? 0 : g.nz - 1;
                        else ok = false;
                    }
                }
                if (ok) l += mun[pel(g, i, (uint32_t)x, (uint32_t)dy + (uint32_t)g.ny * (uint32_t)dz)];
            }
        }
        lam[i] = l;
    });

    // ---- 2. recompute G_m(node) ----
    pull<Q, S, false>(H, flags, g, a.f.sp, x, y, z, node, true, fl == FL_NEAR, gv, fx, fy, fz);
    }
    const bool inlet = xlo;
    if (inlet) velocity_inlet<Q, C>(gv, (C)uin[y + g.ny * z]);

    // ---- 3. local transpose ----
    // With c_i = g_i + om (feq_i - g_i), feq_i = w_i (3p + 3 e_i.u + 4.5 (e_i.u)^2 - 1.5 u.u):
    //   dJ/ds  += sum_i lam_i (g_opp(i) - c_i) + 2 alpha u_x
    //   mu_j    = (1-om)(1-s) lam_j + om (Lp + e_j.Lu) + s lam_opp(j) + 2 alpha s e_jx
    // where Lp = (1-s) A0 and Lu = (1-s) (3 A + 9 B u - 3 A0 u) are the transposes of the
    // moments, written with the moments of w_i lam_i: A0 = sum w lam, A = sum w lam e,
    // B = sum w lam e e^T (the per-direction sum Lu = sum (1-s) w lam (3 e + 9 (e.u) e - 3 u)
    // regrouped).  Opposite directions share |e.u| and (e.u)^2.
    C p, u[3];
    moments01<L, C>(gv, p, u);
    p = p * (C)(1.0 / 3.0);
    const C om = (C)a.f.rtau;   // 1/tau rounded once, on the host
    C sf = 0;
    if constexpr (SFRAC) sf = sfrac[node];
    const C al = (C)a.alpha;
    if constexpr (SFRAC) {
        const C usq = fmac(u[0], u[0], fmac(u[1], u[1], u[2] * u[2]));
        const C X = fmac((C)3, p, (C)-1.5 * usq);
        C acc0 = 0, acc1 = 0;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I, o = L::opp(i);
            if constexpr (i == o) {
                const C c = fmac(om, (C)L::w(i) * X - gv[i], gv[i]);
                acc0 = fmac(lam[i], gv[i] - c, acc0);
            } else if constexpr (i < o) {
                const C eu = edot<L, i>(u);
                const C wt = (C)L::w(i) * fmac((C)4.5 * eu, eu, X);
                const C w3 = (C)(3.0 * L::w(i)) * eu;
                const C ci = fmac(om, (wt + w3) - gv[i], gv[i]);
                const C co = fmac(om, (wt - w3) - gv[o], gv[o]);
                acc0 = fmac(lam[i], gv[o] - ci, acc0);
                acc1 = fmac(lam[o], gv[i] - co, acc1);
            }
        });
        dJds[node] += 2.0 * a.alpha * (double)u[0] + ((double)acc0 + (double)acc1);
    }
    // moments of w lam (B: xx, yy, zz, xy, xz, yz)
    C A0, A[3], B[6];
    wmoments<L, C>(lam, A0, A, B);
    const C ks = (C)1 - sf;
    // V = om Lu (+ 2 alpha s along x), P = om Lp
    const C t3 = (C)3 * A0;
    const C Bu0 = fmac(B[0], u[0], fmac(B[3], u[1], B[4] * u[2]));
    const C Bu1 = fmac(B[3], u[0], fmac(B[1], u[1], B[5] * u[2]));
    const C Bu2 = fmac(B[4], u[0], fmac(B[5], u[1], B[2] * u[2]));
    const C kom = ks * om;
    C V[3];
    V[0] = kom * fmac((C)9, Bu0, fmac((C)3, A[0], -t3 * u[0]));
    V[1] = kom * fmac((C)9, Bu1, fmac((C)3, A[1], -t3 * u[1]));
    V[2] = kom * fmac((C)9, Bu2, fmac((C)3, A[2], -t3 * u[2]));
    if constexpr (SFRAC) V[0] = fmac((C)2 * al, sf, V[0]);
    const C P = kom * A0;
    const C c1 = ((C)1 - om) * ks;
    C m[Q];
    static_for<0, Q>([&](auto J) {
        constexpr int j = J;
        // E = P + e_z V_z + e_y V_y + e_x V_x, in this order so that directions sharing
        // (e_z) or (e_z, e_y) share the partial sums
        C E = P;
        if constexpr (L::ez(j) > 0) E = E + V[2];
        if constexpr (L::ez(j) < 0) E = E - V[2];
        if constexpr (L::ey(j) > 0) E = E + V[1];
        if constexpr (L::ey(j) < 0) E = E - V[1];
        if constexpr (L::ex(j) > 0) E = E + V[0];
        if constexpr (L::ex(j) < 0) E = E - V[0];
        if constexpr (SFRAC) m[j] = fmac(c1, lam[j], fmac(sf, lam[L::opp(j)], E));
        else m[j] = fmac(c1, lam[j], E);
    });
    if (inlet) {
        double du = 0.0;
        adj_inlet_transpose<Q, S>(m, du);
        dJdu[y + (size_t)g.ny * z] += du;
    }
    char *const murow = (char *)mu + (size_t)row * g.rs * sizeof(C);
    const uint32_t bx = (uint32_t)x * (uint32_t)sizeof(C);
    static_for<0, Q>([&](auto I) {
        __builtin_nontemporal_store(m[I], (C *)(murow + (size_t)I * g.ds * sizeof(C) + bx));
    });
}



// initial condition H_0 = feq(0, (u_in(y,z) + perturb*xi, 0, 0)): dJ/du_in(row) +=
// sum over the row's fluid nodes of sum_i lambda_0,i dfeq_i/du_x, lambda_0 = transpose of
// the step-1 pull applied to mu_1 (plus the step-1 link-force term).  One block per row,
// fixed-order reduction.
template <int Q, class S>
__global__ void __launch_bounds__(256)
    k_adj_init(const typename S::C *__restrict__ mun, const uint8_t *__restrict__ flags, const double *__restrict__ uin,
               const double *__restrict__ xi, double perturb, uint64_t seed, double alpha1, double *__restrict__ dJdu,
               const Geo g) {
    using L = Lat<Q>;
    const int y = blockIdx.y, z = blockIdx.z;
    double acc = 0.0;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    for (int x = threadIdx.x; x < g.nx; x += blockDim.x) {
        const uint32_t node = (uint32_t)x + (uint32_t)g.nxp * row;
        const uint8_t fl = flags[node];
        if (fl == FL_SOLID) continue;
        const size_t phys = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
        const double u = uin[y + (size_t)g.ny * z] + perturb * (xi ? xi[phys] : uniform_pm1(seed, phys));
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i), o = L::opp(i);
            double l = 0.0;
            bool skip = false;
            if constexpr (ex < 0) skip = !g.xper && x == 0;
            if (!skip) {
                int dx = x + ex, dy = y + ey, dz = z + ez;
                bool wall = false, out = false;
                if constexpr (ey != 0) {
                    if ((unsigned)dy >= (unsigned)g.ny) {
                        if (g.yper) dy = ey > 0 ? 0 : g.ny - 1;
                        else wall = true;
                    }
                }
                if constexpr (ez != 0) {
                    if ((unsigned)dz >= (unsigned)g.nz) {
                        if (g.zper) dz = ez > 0 ? 0 : g.nz - 1;
                        else wall = true;
                    }
                }
                if constexpr (ex != 0) {
                    if ((unsigned)dx >= (unsigned)g.nx) {
                        if (g.xper) dx = ex > 0 ? 0 : g.nx - 1;
                        else out = true;
                    }
                }
                if (wall) {
                    l = (double)mun[pel(g, o, (uint32_t)x, row)];
                } else if (!out) {
                    const uint32_t drow = (uint32_t)dy + (uint32_t)g.ny * (uint32_t)dz;
                    const uint32_t d = (uint32_t)dx + (uint32_t)g.nxp * drow;
                    if (fl == FL_NEAR && flags[d] == FL_SOLID) l = (double)mun[pel(g, o, (uint32_t)x, row)] + alpha1 * 2.0 * ex;
                    else l = (double)mun[pel(g, i, (uint32_t)dx, drow)];
                }
            }
            if constexpr (ex < 0) {
                if (!g.xper && x == g.nx - 1) {
                    int dy = y + ey, dz = z + ez;
                    bool ok = true;
                    if constexpr (ey != 0) {
                        if ((unsigned)dy >= (unsigned)g.ny) {
                            if (g.yper) dy = ey > 0 ? 0 : g.ny - 1;
                            else ok = false;
                        }
                    }
                    if constexpr (ez != 0) {
                        if ((unsigned)dz >= (unsigned)g.nz) {
                            if (g.zper) dz = ez > 0 ? 0 : g.nz - 1;
                            else ok = false;
                        }
                    }
                    if (ok) l += (double)mun[pel(g, i, (uint32_t)x, (uint32_t)dy + (uint32_t)g.ny * (uint32_t)dz)];
                }
            }
            acc += l * L::w(i) * (3.0 * ex + 9.0 * ex * u * ex - 3.0 * u);
        });
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    __shared__ double red[4];
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) dJdu[y + (size_t)g.ny * z] += ((red[0] + red[1]) + red[2]) + red[3];
}

__global__ void __launch_bounds__(256) k_adj_strip(const double *__restrict__ in, double *__restrict__ out, Geo g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= g.nx) return;
    out[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = in[(size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z)];
}

// ------------------------------------------------------------------------------------
template <int Q, class S>
static hipError_t adj_step_qs(const void *H, const void *mun, void *mu, const uint8_t *flags, const void *sfrac,
                              const double *uin, void *dJds, double *dJdu, const uint8_t *rowclean, bool all_fast,
                              const AdjArgs &a, hipStream_t s) {
    using T = typename S::T;
    using C = typename S::C;
    // blocks of one row's lanes (at most 256), so the fast path's test is uniform per block
    const Shape sh = lane_shape(a.f.g, 1);
    const dim3 grid(sh.gx, a.f.g.ny, a.f.g.nz);
    const AdjTab tab = adj_tab<Q, C, T>(a.f.g);
    auto go = [&](auto kern, const C *sf) {
        hipLaunchKernelGGL(kern, grid, dim3(sh.bx), 0, s, (const T *)H, (const C *)mun, (C *)mu, flags, sf, uin,
                           (double *)dJds, dJdu, rowclean, a, tab);
    };
    if (sfrac) {
        if (all_fast) go(k_adj_step<Q, S, true, false>, (const C *)sfrac);
        else go(k_adj_step<Q, S, true, true>, (const C *)sfrac);
    } else {
        if (all_fast) go(k_adj_step<Q, S, false, false>, (const C *)nullptr);
        else go(k_adj_step<Q, S, false, true>, (const C *)nullptr);
    }
    return hipGetLastError();
}

hipError_t adj_step(int Q, int prec, const void *H, const void *mun, void *mu, const uint8_t *flags, const void *sfrac,
                    const double *uin, void *dJds, double *dJdu, const uint8_t *rowclean, bool all_fast, const AdjArgs &a,
                    hipStream_t s) {
    if (prec == LBM_FP64)
        return Q == 9 ? adj_step_qs<9, SF64>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, all_fast, a, s)
                      : adj_step_qs<27, SF64>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, all_fast, a, s);
    if (prec == LBM_FP32)
        return Q == 9 ? adj_step_qs<9, SF32>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, all_fast, a, s)
                      : adj_step_qs<27, SF32>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, all_fast, a, s);
    return hipErrorInvalidValue;
}

hipError_t adj_init(int Q, int prec, const void *mu1, const uint8_t *flags, const double *uin, const double *xi,
                    double perturb, uint64_t seed, double alpha1, double *dJdu, const FusedArgs &a, hipStream_t s) {
    const dim3 grid(1, a.g.ny, a.g.nz);
    if (prec == LBM_FP64) {
        if (Q == 9) hipLaunchKernelGGL((k_adj_init<9, SF64>), grid, dim3(256), 0, s, (const double *)mu1, flags, uin, xi, perturb, seed, alpha1, dJdu, a.g);
        else hipLaunchKernelGGL((k_adj_init<27, SF64>), grid, dim3(256), 0, s, (const double *)mu1, flags, uin, xi, perturb, seed, alpha1, dJdu, a.g);
    } else if (prec == LBM_FP32) {
        if (Q == 9) hipLaunchKernelGGL((k_adj_init<9, SF32>), grid, dim3(256), 0, s, (const float *)mu1, flags, uin, xi, perturb, seed, alpha1, dJdu, a.g);
        else hipLaunchKernelGGL((k_adj_init<27, SF32>), grid, dim3(256), 0, s, (const float *)mu1, flags, uin, xi, perturb, seed, alpha1, dJdu, a.g);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t adj_reduce_ds(int, const void *dJds, double *out, const Geo &g, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_strip, dim3((g.nx + 255) / 256, g.ny, g.nz), dim3(256), 0, s, (const double *)dJds, out, g);
    return hipGetLastError();
}

}  // namespace lbm
