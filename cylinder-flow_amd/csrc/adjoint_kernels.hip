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

// Interior rows (no wall or periodic wrap in y/z) whose nodes are all fluid and have no solid
// neighbour ("clean", flagged per row on the host) take a straight-line path: every source and
// destination row of a direction is a scalar of the block, and the x edges (inlet, outlet,
// periodic x) are handled with in-row clamped indices and selects, so every load of the path is
// issued unconditionally from an address inside its row.  The arithmetic is the general path's,
// term for term, so both give the same bits.  Other rows take the general path.
template <int Q, class S, bool SFRAC>
__global__ void __launch_bounds__(256)
    k_adj_step(const typename S::T *__restrict__ H, const typename S::C *__restrict__ mun,
               typename S::C *__restrict__ mu, const uint8_t *__restrict__ flags, const typename S::C *__restrict__ sfrac,
               const double *__restrict__ uin, double *__restrict__ dJds, double *__restrict__ dJdu,
               const uint8_t *__restrict__ rowclean, const AdjArgs a) {
    using L = Lat<Q>;
    using C = typename S::C;
    const Geo &g = a.f.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= g.nx) return;
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    const uint32_t node = (uint32_t)x + (uint32_t)g.nxp * row;
    const bool interior = (unsigned)(y - 1) < (unsigned)(g.ny - 2) && (L::D == 2 || (unsigned)(z - 1) < (unsigned)(g.nz - 2));
    const bool fast = interior && rowclean[row];   // uniform over the block (a block lies in one row)
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
        // ---- 1. lambda: entry (node, opp i) pulls from node + e_i, entry (d, i) from node;
        //         the outlet node also feeds (nx-1, y + e_y, z + e_z) along e_ix < 0 ----
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
            const C *mrow = mun + (size_t)i * g.ds + (size_t)(row + (uint32_t)(ey + g.ny * ez)) * g.rs;
            int dx = x + ex;
            bool live = true;
            if constexpr (ex > 0) {
                if (dx == g.nx) {
                    if (g.xper) dx = 0;
                    else live = false, dx = x;   // beyond the outlet: no entry
                }
            }
            if constexpr (ex < 0) {
                if (dx < 0) {
                    if (g.xper) dx = g.nx - 1;
                    else live = false, dx = x;   // (node, opp i) is a Zou-He unknown
                }
            }
            const C v = mrow[dx];
            lam[i] = live ? v : (C)0;
        });
        if (xhi) {   // the outlet lane only
            static_for<0, Q>([&](auto I) {
                constexpr int i = I;
                constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
                if constexpr (ex < 0)
                    lam[i] += mun[(size_t)i * g.ds + (size_t)(row + (uint32_t)(ey + g.ny * ez)) * g.rs + (size_t)x];
            });
        }
        // ---- 2. recompute G_m(node): the forward pull (inlet unknowns 0, outlet clamp) ----
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            constexpr int ex = L::ex(i), ey = L::ey(i), ez = L::ez(i);
            const typename S::T *hrow = H + (size_t)i * g.ds + (size_t)(row - (uint32_t)(ey + g.ny * ez)) * g.rs;
            int sx = x - ex;
            bool zh = false;
            if constexpr (ex > 0) {
                if (sx < 0) {
                    if (g.xper) sx = g.nx - 1;
                    else zh = true, sx = 0;
                }
            }
            if constexpr (ex < 0) {
                if (sx == g.nx) sx = g.xper ? 0 : g.nx - 1;
            }
            const C v = S::template decode<i, L>(hrow[sx], a.f.sp.off[i]);
            gv[i] = zh ? (C)0 : v;
        });
    } else {
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
    C p = 0, u[3] = {0, 0, 0};
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        p += gv[i];
        if constexpr (L::ex(i) != 0) u[0] += (C)L::ex(i) * gv[i];
        if constexpr (L::ey(i) != 0) u[1] += (C)L::ey(i) * gv[i];
        if constexpr (L::ez(i) != 0) u[2] += (C)L::ez(i) * gv[i];
    });
    p = p * (C)(1.0 / 3.0);
    const C om = (C)(1.0 / a.f.tau);
    C sf = 0;
    if constexpr (SFRAC) sf = sfrac[node];
    const C al = (C)a.alpha;
    if constexpr (SFRAC) {
        // dJ/ds = sum_i lam_i (g_opp(i) - c_i) + alpha * 2 u_x
        double ds = 2.0 * a.alpha * (double)u[0];
        const C usq = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            const C eu = (C)L::ex(i) * u[0] + (C)L::ey(i) * u[1] + (C)L::ez(i) * u[2];
            const C feq = (C)L::w(i) * ((C)3 * p + (C)3 * eu + (C)4.5 * eu * eu - (C)1.5 * usq);
            const C c = gv[i] + om * (feq - gv[i]);
            ds += (double)(lam[i] * (gv[L::opp(i)] - c));
        });
        dJds[node] += ds;
    }
    C Lp = 0, Lu[3] = {0, 0, 0};
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        const C li = ((C)1 - sf) * lam[i] * (C)L::w(i);
        const C eu = (C)L::ex(i) * u[0] + (C)L::ey(i) * u[1] + (C)L::ez(i) * u[2];
        Lp += li;
        const C k9 = (C)9 * eu;
        Lu[0] += li * ((C)3 * (C)L::ex(i) + k9 * (C)L::ex(i) - (C)3 * u[0]);
        Lu[1] += li * ((C)3 * (C)L::ey(i) + k9 * (C)L::ey(i) - (C)3 * u[1]);
        Lu[2] += li * ((C)3 * (C)L::ez(i) + k9 * (C)L::ez(i) - (C)3 * u[2]);
    });
    C m[Q];
    static_for<0, Q>([&](auto J) {
        constexpr int j = J;
        const C eL = (C)L::ex(j) * Lu[0] + (C)L::ey(j) * Lu[1] + (C)L::ez(j) * Lu[2];
        m[j] = ((C)1 - om) * ((C)1 - sf) * lam[j] + om * (Lp + eL) + sf * lam[L::opp(j)] +
               al * (C)2 * sf * (C)L::ex(j);
    });
    if (inlet) {
        double du = 0.0;
        adj_inlet_transpose<Q, S>(m, du);
        dJdu[y + (size_t)g.ny * z] += du;
    }
    static_for<0, Q>([&](auto I) { __builtin_nontemporal_store(m[I], &mu[pel(g, I, (uint32_t)x, row)]); });
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
                              const double *uin, void *dJds, double *dJdu, const uint8_t *rowclean, const AdjArgs &a,
                              hipStream_t s) {
    using T = typename S::T;
    using C = typename S::C;
    // blocks of one row's lanes (at most 256), so the fast path's test is uniform per block
    const Shape sh = lane_shape(a.f.g, 1);
    const dim3 grid(sh.gx, a.f.g.ny, a.f.g.nz);
    if (sfrac)
        hipLaunchKernelGGL((k_adj_step<Q, S, true>), grid, dim3(sh.bx), 0, s, (const T *)H, (const C *)mun, (C *)mu, flags,
                           (const C *)sfrac, uin, (double *)dJds, dJdu, rowclean, a);
    else
        hipLaunchKernelGGL((k_adj_step<Q, S, false>), grid, dim3(sh.bx), 0, s, (const T *)H, (const C *)mun, (C *)mu, flags,
                           (const C *)nullptr, uin, (double *)dJds, dJdu, rowclean, a);
    return hipGetLastError();
}

hipError_t adj_step(int Q, int prec, const void *H, const void *mun, void *mu, const uint8_t *flags, const void *sfrac,
                    const double *uin, void *dJds, double *dJdu, const uint8_t *rowclean, const AdjArgs &a,
                    hipStream_t s) {
    if (prec == LBM_FP64)
        return Q == 9 ? adj_step_qs<9, SF64>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, a, s)
                      : adj_step_qs<27, SF64>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, a, s);
    if (prec == LBM_FP32)
        return Q == 9 ? adj_step_qs<9, SF32>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, a, s)
                      : adj_step_qs<27, SF32>(H, mun, mu, flags, sfrac, uin, dJds, dJdu, rowclean, a, s);
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
