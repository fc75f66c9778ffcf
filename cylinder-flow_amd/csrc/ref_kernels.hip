// ref_kernels.hip -- the reference's own boundary rules on gfx950 (FP64 parity path).
//
// Data: the ghost-padded box of the reference (ghost 3 in 3D, 2 in 2D;
// 3D MPI+KOKKOS+MPIIO/lbm.hpp:20-22,153-163, MPI Kokkos/lbm.hpp:19,79-85), rows of px >= lx
// elements (3D: a multiple of 32, so rows start 256-byte aligned; 2D: px = lx, which keeps the
// reference's flat-index arithmetic at the row ends).  Node arrays: a[i + px*(j + ly*k)].
// Populations: f[RIDX(ii, i, j, k)] = f[ii*ds + i + rs*(j + ly*k)], in 3D row-interleaved
// (ds = px, rs = Q*px: the Q direction rows of a row side by side), in 2D one plane per
// direction (ds = px*ly, rs = px).
// One step is
//   fills (the reference's six wall/inlet/outlet kernels in their order, in place)
//   -> usrbc (in place) -> [momentum exchange] -> fused pull + Update + Collision into ft
//   -> swap f/ft.
// Fusing stream1/stream2/Update/Collision is exact: the first reference Collision acts on
// f = feq(macros) and returns f unchanged, so Collision_{n+1} can run right after
// Update_n inside the pull kernel (3D MPI+KOKKOS+MPIIO/main.cpp:29-37).  The fills write
// every ghost slot into both buffers, so ghost values persist across the swaps exactly as in
// the reference's single f array -- including the slots no fill ever refreshes.
#include "internal.hpp"
#include "lattice.hpp"

namespace lbm {

__device__ __forceinline__ uint64_t ref_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

template <int Q, int I>
__device__ __forceinline__ double ref_feq(double p, double u, double v, double w) {
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

// element of population ii at (i, j, k): ii*ds + i + rs*(j + ly*k) (one plane per direction:
// ds = plane, rs = px; row-interleaved: ds = px, rs = Q*px).  The node arrays (usr, moments)
// stay at i + px*(j + ly*k).
#define RIDX(r, ii, i, j, k) \
    ((size_t)(ii) * (r).ds + (size_t)((long)(i) + (long)(r).rs * ((long)(j) + (long)(r).ly * (long)(k))))

// u0*4*c*(n-1-c)/(n-1)^2: 3D MPI+KOKKOS+MPIIO/lbm.cpp:239,405 ; MPI Kokkos/lbm.cpp:109,446
__device__ __forceinline__ double ref_profile(double u0, int c, int n) {
    return u0 * 4.0 * (double)c * (double)(n - 1 - c) / ((double)(n - 1) * (double)(n - 1));
}

struct RefBody {
    int kind;   // LBM_BODY_*
    double cx, cy, cz, r2;
    uint64_t seed;
};

// "random", "initialize"/"initv", "initf" over every node including ghosts
// (3D MPI+KOKKOS+MPIIO/lbm.cpp:221-254 ; MPI Kokkos/lbm.cpp:92-124)
template <int Q>
__global__ void __launch_bounds__(256) k_ref_init(RefGeo r, RefBody b, double *__restrict__ f, double *__restrict__ ft,
                                                  double *__restrict__ mac, uint8_t *__restrict__ usr,
                                                  const int32_t *__restrict__ ran_in) {
    using L = Lat<Q>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y, k = blockIdx.z;
    if (i >= r.lx) return;
    const size_t n = (size_t)i + (size_t)r.px * ((size_t)j + (size_t)r.ly * k);
    const size_t nl = (size_t)i + (size_t)r.lx * ((size_t)j + (size_t)r.ly * k);   // unpadded: ran's index
    const int ran = ran_in ? ran_in[nl] : (int)(ref_splitmix64(b.seed ^ ref_splitmix64(nl)) % 100ull);
    const int x = i - r.g, y = j - r.g, z = (L::D == 3) ? k - r.g : 0;
    int inside = 0;
    if (b.kind == LBM_BODY_SPHERE || b.kind == LBM_BODY_ZCYL) {
        const double dx = (double)x - b.cx, dy = (double)y - b.cy;
        double rr = dx * dx + dy * dy;
        if (L::D == 3 && b.kind == LBM_BODY_SPHERE) {
            const double dz = (double)z - b.cz;
            rr = rr + dz * dz;
        }
        inside = rr <= b.r2;
    }
    const int us = inside ? 0 : 1;
    usr[n] = (uint8_t)us;
    double u;
    if constexpr (L::D == 3) u = ref_profile(r.u0, z, r.glz) * (double)us * (1.0 + (double)ran * 0.001);
    else if (r.variant == 1) u = ref_profile(r.u0, y, r.gly) * (double)us * (1.0 + (double)ran * 0.001);   // Kokkos/lbm.cpp:123
    else u = ref_profile(r.u0, y, r.gly) * (double)us * (0.9 + (double)ran * 0.002);
    mac[n] = 0.0;
    mac[r.plane + n] = u;
    mac[2 * r.plane + n] = 0.0;
    mac[3 * r.plane + n] = 0.0;
    static_for<0, Q>([&](auto I) {
        const double v = ref_feq<Q, I>(0.0, u, 0.0, 0.0);
        f[RIDX(r, I, i, j, k)] = v;
        ft[RIDX(r, I, i, j, k)] = v;   // the ghosts of both buffers start equal (interior: overwritten)
    });
}

// A ghost fill writes both buffers: ghost slots then hold the same values in f and ft, so the
// A/B swap after the fused kernel keeps them (the reference's f keeps its ghosts: its
// stream2 copies the interior only), and the fused kernel need not copy them.
__device__ __forceinline__ void copy2(double *__restrict__ f, double *__restrict__ ft, size_t dst, size_t src) {
    const double v = f[src];
    f[dst] = v;
    ft[dst] = v;
}


// the 9 directions a 3D face fill writes (those pointing into the domain from that face)
template <int FACE>
constexpr bool on_face(int ii) {
    using L = Lat<27>;
    return FACE == 0 ? L::ey(ii) > 0 : FACE == 1 ? L::ey(ii) < 0 : FACE == 2 ? L::ez(ii) > 0
         : FACE == 3 ? L::ez(ii) < 0 : FACE == 4 ? L::ex(ii) > 0 : L::ex(ii) < 0;
}
template <int FACE>
constexpr int face_rank(int ii) {
    int k = 0;
    for (int jj = 0; jj < ii; jj++) k += on_face<FACE>(jj);
    return k;
}

// 3D fills, one kernel per face: 3D MPI+KOKKOS+MPIIO/lbm.cpp:278-347.
// face: 0 front (y lo), 1 back (y hi), 2 bottom (z lo), 3 top (z hi), 4 inlet, 5 outlet.
// One lane per (face position, direction): blockIdx.z picks the direction among the face's
// nine.  Within one fill no copy reads a slot that another copy of the same fill writes (the
// sources lie inside the ghost layer's inner neighbour planes), so the directions need no order.
template <int FACE>
__device__ __forceinline__ void fill3(const RefGeo &r, double *__restrict__ f, double *__restrict__ ft, int slot) {
    using L = Lat<27>;
    const int a = blockIdx.x * blockDim.x + threadIdx.x + r.g - 1;   // first in-face coordinate
    const int b = blockIdx.y + r.g - 1;                              // second in-face coordinate
    const int g = r.g;
    const int le0 = g + r.glx, le1 = g + r.gly, le2 = g + r.glz;
    if (FACE <= 1) { if (a >= r.lx - g + 1 || b >= r.lz - g + 1) return; }   // (i, k)
    else if (FACE <= 3) { if (a >= r.lx - g + 1 || b >= r.ly - g + 1) return; }  // (i, j)
    else { if (a >= r.ly - g + 1 || b >= r.lz - g + 1) return; }   // (j, k)
    static_for<0, 27>([&](auto I) {
        constexpr int ii = I;
        constexpr int ex = L::ex(ii), ey = L::ey(ii), ez = L::ez(ii), bb = L::opp(ii);
        if constexpr (!on_face<FACE>(ii)) return;
        if (face_rank<FACE>(ii) != slot) return;
        if constexpr (FACE == 0 && ey > 0)
            copy2(f, ft, RIDX(r, ii, a, g - 1, b), RIDX(r, bb, a + 2 * ex, g + 1, b + 2 * ez));
        if constexpr (FACE == 1 && ey < 0)
            copy2(f, ft, RIDX(r, ii, a, le1, b), RIDX(r, bb, a + 2 * ex, le1 - 2, b + 2 * ez));
        if constexpr (FACE == 2 && ez > 0)
            copy2(f, ft, RIDX(r, ii, a, b, g - 1), RIDX(r, bb, a + 2 * ex, b + 2 * ey, g + 1));
        if constexpr (FACE == 3 && ez < 0)
            copy2(f, ft, RIDX(r, ii, a, b, le2), RIDX(r, bb, a + 2 * ex, b + 2 * ey, le2 - 2));
        if constexpr (FACE == 4 && ex > 0) copy2(f, ft, RIDX(r, ii, g - 1, a, b), RIDX(r, ii, g, a, b));
        if constexpr (FACE == 5 && ex < 0)
            copy2(f, ft, RIDX(r, ii, le0, a, b), RIDX(r, ii, le0 + ex, a + ey, b + ez));
    });
}

// The two faces of one axis in one launch (blockIdx.z < 9: face FA, else FB; the slot picks the
// direction): their copies touch disjoint slots (FA writes and reads only planes near its own
// face, FB near the opposite one, at least 3 nodes apart), so they need no order between them;
// the reference's order between the axes stays (y faces, then z, then x).
template <int FA, int FB>
__global__ void __launch_bounds__(256) k_ref_fill3(RefGeo r, double *__restrict__ f, double *__restrict__ ft) {
    const int z = (int)blockIdx.z;   // uniform over the block
    if (z < 9) fill3<FA>(r, f, ft, z);
    else fill3<FB>(r, f, ft, z - 9);
}

// 2D fills: MPI Kokkos/lbm.cpp:315-345 (stream3 inlet, stream3 outlet, stream4 bottom,
// stream4 top).  Columns i-2 = -1 and i+2 = lx wrap through the flat index, as in the
// reference's LayoutLeft views.
template <int FACE>
__device__ __forceinline__ void fill2(const RefGeo &r, double *__restrict__ f, double *__restrict__ ft) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int lx = r.lx, ly = r.ly;
    if (FACE <= 1) {
        if (t >= ly - 1) return;
        const int j = t;
        if (FACE == 0) {
            copy2(f, ft, RIDX(r, 1, 1, j, 0), RIDX(r, 1, 2, j, 0));
            copy2(f, ft, RIDX(r, 5, 1, j, 0), RIDX(r, 5, 2, j, 0));
            copy2(f, ft, RIDX(r, 8, 1, j, 0), RIDX(r, 8, 2, j, 0));
        } else {
            copy2(f, ft, RIDX(r, 3, lx - 2, j, 0), RIDX(r, 3, lx - 3, j, 0));
            copy2(f, ft, RIDX(r, 7, lx - 2, j, 0), RIDX(r, 7, lx - 3, j - 1, 0));
            copy2(f, ft, RIDX(r, 6, lx - 2, j, 0), RIDX(r, 6, lx - 3, j + 1, 0));
        }
    } else {
        if (t >= lx - 1) return;
        const int i = t;
        if (r.variant == 1 && (i < 2 || i >= lx - 2)) return;   // Kokkos/lbm.cpp:158: i in [2, lx-2)
        if (FACE == 2) {
            copy2(f, ft, RIDX(r, 2, i, 1, 0), RIDX(r, 4, i, 3, 0));
            copy2(f, ft, RIDX(r, 5, i, 1, 0), RIDX(r, 7, i + 2, 3, 0));
            copy2(f, ft, RIDX(r, 6, i, 1, 0), RIDX(r, 8, i - 2, 3, 0));
        } else {
            copy2(f, ft, RIDX(r, 4, i, ly - 2, 0), RIDX(r, 2, i, ly - 4, 0));
            copy2(f, ft, RIDX(r, 7, i, ly - 2, 0), RIDX(r, 5, i - 2, ly - 4, 0));
            copy2(f, ft, RIDX(r, 8, i, ly - 2, 0), RIDX(r, 6, i + 2, ly - 4, 0));
        }
    }
}

// Inlet and outlet (FA, FB = 0, 1), or bottom and top (2, 3), in one launch (blockIdx.y picks
// the face): the pair writes and reads columns 1 and 2 against lx-2 and lx-3, or rows 1 and
// 2..4 against ly-2 and ly-5..ly-3 (ly >= 7), so its two faces touch disjoint slots and need
// no order between them; inlet/outlet still run before bottom/top, as in the reference.
template <int FA, int FB>
__global__ void __launch_bounds__(256) k_ref_fill2(RefGeo r, double *__restrict__ f, double *__restrict__ ft) {
    if (blockIdx.y == 0) fill2<FA>(r, f, ft);
    else fill2<FB>(r, f, ft);
}

// usrbc: 3D MPI+KOKKOS+MPIIO/lbm.cpp:350-356 ; MPI Kokkos/lbm.cpp:347-353 (in place; the
// host rejects masks with a solid-fluid-solid run along a lattice link, the only case in
// which two writes and a read of this kernel touch the same slot)
template <int Q>
__global__ void __launch_bounds__(256) k_ref_usrbc(RefGeo r, double *__restrict__ f, const uint8_t *__restrict__ usr) {
    using L = Lat<Q>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + r.g, j = blockIdx.y + r.g;
    const int k = blockIdx.z + (L::D == 3 ? r.g : 0);
    if (i >= r.lx - r.g) return;
    const long n = (long)i + (long)r.px * ((long)j + (long)r.ly * k);
    if (usr[n] != 0) return;
    static_for<0, Q>([&](auto I) {
        constexpr int ii = I;
        const long nb = n + L::ex(ii) + (long)r.px * (L::ey(ii) + (long)r.ly * L::ez(ii));
        if (usr[nb] == 1) {
            f[RIDX(r, ii, i, j, k)] = f[RIDX(r, L::opp(ii), i + 2 * L::ex(ii), j + 2 * L::ey(ii), k + 2 * L::ez(ii))];
        }
    });
}

// usrbc over a list of the nodes it can change (interior solid nodes with a fluid neighbour,
// node indices built once from usr): the same per-node work as k_ref_usrbc without sweeping
// the interior
template <int Q>
__global__ void __launch_bounds__(256) k_ref_usrbc_list(RefGeo r, double *__restrict__ f, const uint8_t *__restrict__ usr,
                                                        const uint32_t *__restrict__ list, int nlist) {
    using L = Lat<Q>;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nlist) return;
    const uint32_t node = list[t], row = node / (uint32_t)r.px;
    const int i = (int)(node - row * (uint32_t)r.px), j = (int)(row % (uint32_t)r.ly), k = (int)(row / (uint32_t)r.ly);
    const long n = (long)node;
    static_for<0, Q>([&](auto I) {
        constexpr int ii = I;
        const long nb = n + L::ex(ii) + (long)r.px * (L::ey(ii) + (long)r.ly * L::ez(ii));
        if (usr[nb] == 1)
            f[RIDX(r, ii, i, j, k)] = f[RIDX(r, L::opp(ii), i + 2 * L::ex(ii), j + 2 * L::ey(ii), k + 2 * L::ez(ii))];
    });
}

// momentum exchange after usrbc: sum over fluid X, solid X+e of (f_i(X) + f_bb(i)(X+e)) e
// (MPI Kokkos/lbm.cpp:364-392), per-block partials then a fixed-order second pass.
// USRBC: the same pass also applies usrbc (solid lanes write, fluid lanes sum): for an interior
// solid X+e the sum takes f_i(X-e), the value usrbc writes into f_bb(i)(X+e) (X-e is fluid:
// the host rejects one-node fluid gaps), so no lane reads a slot another lane writes and the
// forces are those of the two separate passes, bit for bit.
template <int Q, bool USRBC>
__global__ void __launch_bounds__(256) k_ref_mem(RefGeo r, double *f, const uint8_t *__restrict__ usr,
                                                 double *__restrict__ partials) {
    using L = Lat<Q>;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + r.g, j = blockIdx.y + r.g;
    const int k = blockIdx.z + (L::D == 3 ? r.g : 0);
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (i < r.lx - r.g) {
        const long n = (long)i + (long)r.px * ((long)j + (long)r.ly * k);
        if (USRBC && usr[n] == 0) {
            static_for<0, Q>([&](auto I) {
                constexpr int ii = I;
                const long nb = n + L::ex(ii) + (long)r.px * (L::ey(ii) + (long)r.ly * L::ez(ii));
                if (usr[nb] == 1)
                    f[RIDX(r, ii, i, j, k)] =
                        f[RIDX(r, L::opp(ii), i + 2 * L::ex(ii), j + 2 * L::ey(ii), k + 2 * L::ez(ii))];
            });
        } else if (usr[n] == 1) {
            static_for<0, Q>([&](auto I) {
                constexpr int ii = I;
                constexpr int ex = L::ex(ii), ey = L::ey(ii), ez = L::ez(ii);
                const long nb = n + ex + (long)r.px * (ey + (long)r.ly * ez);
                if (usr[nb] == 0) {
                    double bb;
                    if constexpr (USRBC) {
                        const int xi = i + ex, xj = j + ey, xk = k + ez;
                        const bool inner = xi >= r.g && xi < r.lx - r.g && xj >= r.g && xj < r.ly - r.g &&
                                           (L::D == 2 || (xk >= r.g && xk < r.lz - r.g));
                        bb = inner ? f[RIDX(r, ii, i - ex, j - ey, k - ez)] : f[RIDX(r, L::opp(ii), xi, xj, xk)];
                    } else {
                        bb = f[RIDX(r, L::opp(ii), i + ex, j + ey, k + ez)];
                    }
                    const double m = f[RIDX(r, ii, i, j, k)] + bb;
                    fx += m * (double)L::ex(ii);
                    fy += m * (double)L::ey(ii);
                    fz += m * (double)L::ez(ii);
                }
            });
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        fx += __shfl_xor(fx, off);
        fy += __shfl_xor(fy, off);
        fz += __shfl_xor(fz, off);
    }
    __shared__ double red[3 * 4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[3 * wv] = fx;
        red[3 * wv + 1] = fy;
        red[3 * wv + 2] = fz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; q++) {
            red[0] += red[3 * q];
            red[1] += red[3 * q + 1];
            red[2] += red[3 * q + 2];
        }
        const size_t blk = blockIdx.x + (size_t)gridDim.x * (blockIdx.y + (size_t)gridDim.y * blockIdx.z);
        partials[3 * blk] = red[0];
        partials[3 * blk + 1] = red[1];
        partials[3 * blk + 2] = red[2];
    }
}

// Update (the moments with the reference's inlet/outlet overrides) and the next Collision of
// one interior node from its streamed populations fv: 3D MPI+KOKKOS+MPIIO/lbm.cpp:374-417 then
// :257-272 ; MPI Kokkos/lbm.cpp:420-457 then :127-139.  The collided values go to out (the other
// buffer, or f itself in place); fpre, when set, receives fv (the populations after Update).
// NT: non-temporal stores of the collided values (buffers larger than the last-level cache:
// the next step reads them from HBM anyway); small buffers keep regular stores and stay cached.
template <int Q, bool NT>
__device__ __forceinline__ void ref_update_collide(const RefGeo &r, int i, int j, int k, size_t n, size_t pe,
                                                   const double *fv,
                                                   double *__restrict__ mac, const uint8_t *__restrict__ usr,
                                                   const double *__restrict__ prof, double *__restrict__ fpre,
                                                   double *out) {
    using L = Lat<Q>;
    const int g = r.g;
    double p = 0.0, u = 0.0, v = 0.0, w = 0.0;
    static_for<0, Q>([&](auto I) {
        constexpr int ii = I;
        if (fpre) fpre[(size_t)ii * r.ds + pe] = fv[ii];
        p = p + fv[ii] / 3.0;
        if constexpr (L::ex(ii) > 0) u = u + fv[ii];
        else if constexpr (L::ex(ii) < 0) u = u - fv[ii];
        if constexpr (L::ey(ii) > 0) v = v + fv[ii];
        else if constexpr (L::ey(ii) < 0) v = v - fv[ii];
        if constexpr (L::ez(ii) > 0) w = w + fv[ii];
        else if constexpr (L::ez(ii) < 0) w = w - fv[ii];
    });
    const double us = (double)usr[n];
    u = u * us;
    v = v * us;
    w = w * us;
    p = p * us;
    if (i == g) {
        u = prof[L::D == 3 ? k - g : j - g];
        v = 0.0;
        w = 0.0;
    }
    if (i == r.lx - g - 1) p = 0.0;
    if (mac) {   // the fields are only read after a call's last step
        mac[n] = p;
        mac[r.plane + n] = u;
        mac[2 * r.plane + n] = v;
        mac[3 * r.plane + n] = w;
    }
    static_for<0, Q>([&](auto I) {
        constexpr int ii = I;
        const double feq = ref_feq<Q, ii>(p, u, v, w);
        const double h = fv[ii] - (fv[ii] - feq) / (r.tau0 + 0.5);
        if constexpr (NT) __builtin_nontemporal_store(h, out + (size_t)ii * r.ds + pe);
        else out[(size_t)ii * r.ds + pe] = h;
    });
}

// Interior rows only (the grid covers them), lanes from the row start so that every wave stays
// aligned with the row pitch; ghost slots are equal in both buffers (copy2).  false for the
// ghost lanes at either end of the row.
#define REF_INTERIOR_LANE(r, i, j, k)                                                            \
    const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y + (r).g;              \
    const int k = blockIdx.z + (Lat<Q>::D == 3 ? (r).g : 0);                                 \
    if (i < (r).g || i >= (r).lx - (r).g) return;                                             \
    const size_t n = (size_t)i + (size_t)(r).px * ((size_t)j + (size_t)(r).ly * k);          \
    const size_t pe = RIDX(r, 0, i, j, k)

// stream1 + stream2 + Update + the next Collision, one pass:
// 3D MPI+KOKKOS+MPIIO/lbm.cpp:257-272,359-417 ; MPI Kokkos/lbm.cpp:127-139,408-457
template <int Q, bool NT>
__global__ void __launch_bounds__(256) k_ref_fused(RefGeo r, const double *__restrict__ f, double *__restrict__ ft,
                                                   double *__restrict__ mac, const uint8_t *__restrict__ usr,
                                                   const double *__restrict__ prof, double *__restrict__ fpre) {
    using L = Lat<Q>;
    REF_INTERIOR_LANE(r, i, j, k);
    double fv[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int ii = I;
        fv[ii] = f[RIDX(r, ii, i - L::ex(ii), j - L::ey(ii), k - L::ez(ii))];
    });
    ref_update_collide<Q, NT>(r, i, j, k, n, pe, fv, mac, usr, prof, fpre, ft);
}

// stream1 alone (the staged step): ft = f pulled along e_i at the interior nodes;
// 3D MPI+KOKKOS+MPIIO/lbm.cpp:359-364 ; MPI Kokkos/lbm.cpp:408-412.  stream2's copy back is
// the caller's buffer swap (the ghost slots of both buffers are equal).
template <int Q, bool NT>
__global__ void __launch_bounds__(256) k_ref_stream(RefGeo r, const double *__restrict__ f, double *__restrict__ ft) {
    using L = Lat<Q>;
    REF_INTERIOR_LANE(r, i, j, k);
    (void)n;
    static_for<0, Q>([&](auto I) {
        constexpr int ii = I;
        const double h = f[RIDX(r, ii, i - L::ex(ii), j - L::ey(ii), k - L::ez(ii))];
        if constexpr (NT) __builtin_nontemporal_store(h, ft + (size_t)ii * r.ds + pe);
        else ft[(size_t)ii * r.ds + pe] = h;
    });
}

// Update and the next Collision in place (the staged step): every lane reads its node's Q
// populations before it writes them back, and no lane touches another node
template <int Q, bool NT>
__global__ void __launch_bounds__(256) k_ref_update(RefGeo r, double *f, double *__restrict__ mac,
                                                    const uint8_t *__restrict__ usr, const double *__restrict__ prof,
                                                    double *__restrict__ fpre) {
    REF_INTERIOR_LANE(r, i, j, k);
    double fv[Q];
    static_for<0, Q>([&](auto I) { fv[I] = f[(size_t)I * r.ds + pe]; });
    ref_update_collide<Q, NT>(r, i, j, k, n, pe, fv, mac, usr, prof, fpre, f);
}

__global__ void __launch_bounds__(256) k_ref_reduce(const double *__restrict__ partials, int n, double *__restrict__ hist,
                                                    int rec, long step) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) {
        a += partials[3 * (size_t)j];
        b += partials[3 * (size_t)j + 1];
        c += partials[3 * (size_t)j + 2];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    __shared__ double red[12];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane == 0) {
        red[3 * wv] = a;
        red[3 * wv + 1] = b;
        red[3 * wv + 2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; q++) {
            red[0] += red[3 * q];
            red[1] += red[3 * q + 1];
            red[2] += red[3 * q + 2];
        }
        hist[4 * (size_t)rec] = (double)step;
        hist[4 * (size_t)rec + 1] = red[0];
        hist[4 * (size_t)rec + 2] = red[1];
        hist[4 * (size_t)rec + 3] = red[2];
    }
}

// ------------------------------------------------------------------------------------
static dim3 interior_grid(const RefGeo &r) {
    return dim3((r.glx + 255) / 256, r.gly, r.D == 3 ? r.glz : 1);
}

int ref_mem_blocks(const RefGeo &r) {
    const dim3 g = interior_grid(r);
    return (int)(g.x * g.y * g.z);
}

hipError_t ref_init(const RefGeo &r, double *f, double *ft, double *mac, uint8_t *usr, const int32_t *ran,
                    int body, double cx, double cy, double cz, double r2, uint64_t seed, hipStream_t s) {
    RefBody b;
    b.kind = body;
    b.cx = cx;
    b.cy = cy;
    b.cz = cz;
    b.r2 = r2;
    b.seed = seed;
    const dim3 grid((r.lx + 255) / 256, r.ly, r.lz);
    if (r.Q == 9) hipLaunchKernelGGL(k_ref_init<9>, grid, dim3(256), 0, s, r, b, f, ft, mac, usr, ran);
    else hipLaunchKernelGGL(k_ref_init<27>, grid, dim3(256), 0, s, r, b, f, ft, mac, usr, ran);
    return hipGetLastError();
}

hipError_t ref_fills(const RefGeo &r, double *f, double *ft, hipStream_t s) {
    if (r.D == 3) {
        const int nI = r.lx - 2 * (r.g - 1), nJ = r.ly - 2 * (r.g - 1), nK = r.lz - 2 * (r.g - 1);
        hipLaunchKernelGGL((k_ref_fill3<0, 1>), dim3((nI + 255) / 256, nK, 18), dim3(256), 0, s, r, f, ft);
        hipLaunchKernelGGL((k_ref_fill3<2, 3>), dim3((nI + 255) / 256, nJ, 18), dim3(256), 0, s, r, f, ft);
        hipLaunchKernelGGL((k_ref_fill3<4, 5>), dim3((nJ + 255) / 256, nK, 18), dim3(256), 0, s, r, f, ft);
    } else {
        hipLaunchKernelGGL((k_ref_fill2<0, 1>), dim3((r.ly + 255) / 256, 2), dim3(256), 0, s, r, f, ft);
        hipLaunchKernelGGL((k_ref_fill2<2, 3>), dim3((r.lx + 255) / 256, 2), dim3(256), 0, s, r, f, ft);
    }
    return hipGetLastError();
}

hipError_t ref_usrbc(const RefGeo &r, double *f, const uint8_t *usr, const uint32_t *list, int nlist, hipStream_t s) {
    if (list) {   // the listed nodes only (none: nothing to do)
        if (nlist == 0) return hipSuccess;
        const dim3 grid((unsigned)((nlist + 255) / 256));
        if (r.Q == 9) hipLaunchKernelGGL(k_ref_usrbc_list<9>, grid, dim3(256), 0, s, r, f, usr, list, nlist);
        else hipLaunchKernelGGL(k_ref_usrbc_list<27>, grid, dim3(256), 0, s, r, f, usr, list, nlist);
        return hipGetLastError();
    }
    if (r.Q == 9) hipLaunchKernelGGL(k_ref_usrbc<9>, interior_grid(r), dim3(256), 0, s, r, f, usr);
    else hipLaunchKernelGGL(k_ref_usrbc<27>, interior_grid(r), dim3(256), 0, s, r, f, usr);
    return hipGetLastError();
}

hipError_t ref_mem(const RefGeo &r, double *f, const uint8_t *usr, double *partials, int nblk, double *hist,
                   int rec, long step, bool with_usrbc, hipStream_t s) {
    if (with_usrbc) {
        if (r.Q == 9) hipLaunchKernelGGL((k_ref_mem<9, true>), interior_grid(r), dim3(256), 0, s, r, f, usr, partials);
        else hipLaunchKernelGGL((k_ref_mem<27, true>), interior_grid(r), dim3(256), 0, s, r, f, usr, partials);
    } else {
        if (r.Q == 9) hipLaunchKernelGGL((k_ref_mem<9, false>), interior_grid(r), dim3(256), 0, s, r, f, usr, partials);
        else hipLaunchKernelGGL((k_ref_mem<27, false>), interior_grid(r), dim3(256), 0, s, r, f, usr, partials);
    }
    hipLaunchKernelGGL(k_ref_reduce, dim3(1), dim3(256), 0, s, partials, nblk, hist, rec, step);
    return hipGetLastError();
}

// grid of the interior-row kernels (lanes from the row start) and their store kind: measured,
// 3D 512x128x128 (2.1 GB per buffer) 9 % faster with non-temporal stores, 2D 512x128 (4.9 MB)
// 12 % slower; 256 MB is the size of the last-level (Infinity) cache
static dim3 row_grid(const RefGeo &r) { return dim3((r.lx + 255) / 256, r.gly, r.D == 3 ? r.glz : 1); }
static bool nt_stores(const RefGeo &r) { return (double)r.Q * (double)r.plane * sizeof(double) > 256.0 * (1 << 20); }

// K<Q, NT> launched over the interior rows with the store kind for r's buffers
template <template <int, bool> class K, class... Args>
static hipError_t launch_rows(const RefGeo &r, hipStream_t s, Args... args) {
    const dim3 grid = row_grid(r);
    const bool nt = nt_stores(r);
    if (r.Q == 9 && nt) K<9, true>::launch(grid, s, r, args...);
    else if (r.Q == 9) K<9, false>::launch(grid, s, r, args...);
    else if (nt) K<27, true>::launch(grid, s, r, args...);
    else K<27, false>::launch(grid, s, r, args...);
    return hipGetLastError();
}

// dynamic LDS of the fused kernel's blocks: the resident-block cap (RefGeo::lds, unused memory)
template <class... A> static unsigned ref_lds(const RefGeo &r, A...) { return r.lds; }

template <int Q, bool NT> struct FusedK {
    template <class... A> static void launch(dim3 g, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_ref_fused<Q, NT>), g, dim3(256), ref_lds(a...), s, a...);
    }
};
template <int Q, bool NT> struct StreamK {
    template <class... A> static void launch(dim3 g, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_ref_stream<Q, NT>), g, dim3(256), 0, s, a...);
    }
};
template <int Q, bool NT> struct UpdateK {
    template <class... A> static void launch(dim3 g, hipStream_t s, A... a) {
        hipLaunchKernelGGL((k_ref_update<Q, NT>), g, dim3(256), 0, s, a...);
    }
};

hipError_t ref_fused(const RefGeo &r, const double *f, double *ft, double *mac, const uint8_t *usr,
                     const double *prof, double *fpre, hipStream_t s) {
    return launch_rows<FusedK>(r, s, f, ft, mac, usr, prof, fpre);
}

hipError_t ref_stream(const RefGeo &r, const double *f, double *ft, hipStream_t s) {
    return launch_rows<StreamK>(r, s, f, ft);
}

hipError_t ref_update(const RefGeo &r, double *f, double *mac, const uint8_t *usr, const double *prof, double *fpre,
                      hipStream_t s) {
    return launch_rows<UpdateK>(r, s, f, mac, usr, prof, fpre);
}

}  // namespace lbm
