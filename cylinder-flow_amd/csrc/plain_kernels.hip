// plain_kernels.hip -- fused pull-scheme lattice update for gfx950 (plain boundary mode).
//
// One kernel per time step replaces the reference's Collision -> boundary fills -> usrbc
// -> stream1 -> stream2 -> Update sequence (3D MPI+KOKKOS+MPIIO/lbm.cpp:257-417;
// MPI Kokkos/lbm.cpp:127-139,313-457).  Each lane owns one node and
//   1. pulls Q populations from buffer A (neighbour x - e_i, or a boundary rule),
//   2. applies the velocity inlet at x = 0 (Zou-He, regularised),
//   3. takes the moments p = sum g / 3, u = sum g e (the reference's Update),
//   4. relaxes toward the shifted incompressible equilibrium (the reference's Collision,
//      for the next step) and applies partial bounce-back where a solid fraction is set,
//   5. writes Q populations to buffer B.
// A and B swap every step, so the kernel never reads what it writes (race-free by
// construction; the reference's in-place usrbc is not).  Global memory traffic per node is
// Q reads + Q writes -- the 2*Q*sizeof(storage) bytes the roofline counts -- plus one byte of
// node flags.
//
// Boundary rules of a pulled entry (node, i), first match wins:
//   x == 0 and e_ix > 0 (inlet, x not periodic)   -> Zou-He unknown, computed in step 2
//   source outside y/z (walls, not periodic)       -> half-way bounce-back g_i = h_opp(i)(node)
//   source x == nx (outlet, x not periodic)        -> zero gradient, source clamped to nx-1
//   source is solid (mask)                         -> half-way bounce-back, momentum exchange
//   otherwise                                      -> g_i = h_i(node - e_i) (periodic wrap)
#include <type_traits>

#include "fused_common.hpp"

namespace lbm {

Shape node_shape(const Geo &g) {
    Shape s;
    s.bx = g.nxp <= 512 ? g.nxp : 256;
    s.gx = (g.nxp + s.bx - 1) / s.bx;
    return s;
}

template <int Q, class S>
__global__ void __launch_bounds__(512) k_init(typename S::T *__restrict__ A, typename S::T *__restrict__ B,
                                              double *__restrict__ macros, const uint8_t *__restrict__ flags, const double *__restrict__ uin,
                                              const double *__restrict__ xi, double perturb, uint64_t seed,
                                              const FusedArgs a) {
    using L = Lat<Q>;
    using T = typename S::T;
    const Geo &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= g.nxp) return;
    const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
    const bool live = x < g.nx && flags[node] != FL_SOLID;
    double u = 0.0;
    if (live) {
        const size_t phys = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
        const double r = xi ? xi[phys] : uniform_pm1(seed, phys);
        u = uin[y + (size_t)g.ny * z] + perturb * r;
    }
    macros[node] = 0.0;
    macros[g.plane + node] = u;
    macros[2 * g.plane + node] = 0.0;
    macros[3 * g.plane + node] = 0.0;
    static_for<0, Q>([&](auto I) {
        constexpr int i = I;
        const double f = live ? feq_ref<Q, i>(0.0, u, 0.0, 0.0) : 0.0;
        const T h = S::encode(f, (float)(1.0 / L::w(i)), a.sp.offw[i]);
        const size_t e = pel(g, i, (uint32_t)x, (uint32_t)y + (uint32_t)g.ny * (uint32_t)z);
        A[e] = h;
        B[e] = h;
    });
}

template <int Q, class S, bool FORCE, bool MACROS, bool SFRAC>
__global__ void __launch_bounds__(512)
    k_fused(const typename S::T *__restrict__ A, typename S::T *__restrict__ B, const uint8_t *__restrict__ flags,
            const typename S::C *__restrict__ sfrac, const double *__restrict__ uin,
            const int *__restrict__ block_slot, double *__restrict__ partials, double *__restrict__ macros,
            const FusedArgs a) {
    using L = Lat<Q>;
    using C = typename S::C;
    const Geo &g = a.g;
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y, z = blockIdx.z;
    const bool active = x < g.nx;
    // node offset (a plane is < 2^32 nodes); lanes past the row (x >= nx) use the row's first
    // node as a valid, unused address.  Its populations start at element pnode (+ i*ds).
    const uint32_t row = (uint32_t)y + (uint32_t)g.ny * (uint32_t)z;
    const uint32_t node = (uint32_t)(active ? x : 0) + (uint32_t)g.nxp * row;
    const size_t pnode = node + (size_t)row * (g.rs - g.nxp);
    const uint8_t fl = active ? flags[node] : (uint8_t)FL_SOLID;
    const bool live = fl != FL_SOLID;
    const bool near = fl == FL_NEAR;

    C gv[Q];
    double fx = 0.0, fy = 0.0, fz = 0.0;

    // ---- 1. pull ----
    pull<Q, S, FORCE>(A, flags, g, a.sp, x, y, z, node, live, near, gv, fx, fy, fz);

    // ---- 2. velocity inlet ----
    if (!g.xper && x == 0 && live) velocity_inlet<Q, C>(gv, (C)uin[y + g.ny * z]);

    // ---- 3./4. moments, collision, partial bounce-back, store ----
    C sf = 0;
    if constexpr (SFRAC) sf = live ? sfrac[node] : (C)0;
    double mp, mu, mv, mw;
    if constexpr (sizeof(C) == 8) {
        // FP64: the reference's term order -- p += f/3.0 per direction (lbm.cpp:390-397)
        double p = 0.0, u = 0.0, v = 0.0, w = 0.0;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            p = p + div_cr(gv[i], 3.0, 1.0 / 3.0);
            if constexpr (L::ex(i) > 0) u = u + gv[i];
            else if constexpr (L::ex(i) < 0) u = u - gv[i];
            if constexpr (L::ey(i) > 0) v = v + gv[i];
            else if constexpr (L::ey(i) < 0) v = v - gv[i];
            if constexpr (L::ez(i) > 0) w = w + gv[i];
            else if constexpr (L::ez(i) < 0) w = w - gv[i];
        });
        if constexpr (FORCE && SFRAC) {
            fx += 2.0 * sf * u;
            fy += 2.0 * sf * v;
            fz += 2.0 * sf * w;
        }
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            const double feq = feq_ref<Q, i>(p, u, v, w);
            double c = gv[i] - div_cr(gv[i] - feq, a.tau, a.rtau);
            if constexpr (SFRAC) c = (1.0 - sf) * c + sf * gv[L::opp(i)];
            if (live) S::template store<i, L>(B + (size_t)i * g.ds + pnode, c, a.sp);
        });
        mp = p; mu = u; mv = v; mw = w;
    } else {
        float s = 0.f, u = 0.f, v = 0.f, w = 0.f;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            s += gv[i];
            if constexpr (L::ex(i) > 0) u += gv[i];
            else if constexpr (L::ex(i) < 0) u -= gv[i];
            if constexpr (L::ey(i) > 0) v += gv[i];
            else if constexpr (L::ey(i) < 0) v -= gv[i];
            if constexpr (L::ez(i) > 0) w += gv[i];
            else if constexpr (L::ez(i) < 0) w -= gv[i];
        });
        if constexpr (FORCE && SFRAC) {
            fx += 2.0 * (double)sf * (double)u;
            fy += 2.0 * (double)sf * (double)v;
            fz += 2.0 * (double)sf * (double)w;
        }
        // feq_i = w_i (3p - 1.5|u|^2 + e.u (3 + 4.5 e.u)), with 3p = sum g
        const float base = __builtin_fmaf(-1.5f, __builtin_fmaf(u, u, __builtin_fmaf(v, v, w * w)), s);
        const float om = a.omega;
        static_for<0, Q>([&](auto I) {
            constexpr int i = I;
            float eu = 0.f;
            if constexpr (L::ex(i) > 0) eu += u;
            else if constexpr (L::ex(i) < 0) eu -= u;
            if constexpr (L::ey(i) > 0) eu += v;
            else if constexpr (L::ey(i) < 0) eu -= v;
            if constexpr (L::ez(i) > 0) eu += w;
            else if constexpr (L::ez(i) < 0) eu -= w;
            const float feq = (float)L::w(i) * __builtin_fmaf(eu, __builtin_fmaf(4.5f, eu, 3.0f), base);
            float c = __builtin_fmaf(om, feq - gv[i], gv[i]);
            if constexpr (SFRAC) c = __builtin_fmaf(sf, gv[L::opp(i)] - c, c);
            if (live) S::template store<i, L>(B + (size_t)i * g.ds + pnode, c, a.sp);
        });
        mp = (double)s * (1.0 / 3.0); mu = u; mv = v; mw = w;
    }
    if constexpr (MACROS) {
        if (active) {
            macros[node] = live ? mp : 0.0;
            macros[g.plane + node] = live ? mu : 0.0;
            macros[2 * g.plane + node] = live ? mv : 0.0;
            macros[3 * g.plane + node] = live ? mw : 0.0;
        }
    }
    if constexpr (FORCE) {
        const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
        const int slot = block_slot[bid];
        if (slot >= 0) {
            __shared__ double red[3 * 8];
            block_sum3<8>(fx, fy, fz, red);
            if (threadIdx.x == 0) {
                partials[3 * (size_t)slot] = fx;
                partials[3 * (size_t)slot + 1] = fy;
                partials[3 * (size_t)slot + 2] = fz;
            }
        }
    }
}

// Force reduction after the per-block / per-node partials.  Above FORCE_CHUNK slots a first
// pass sums fixed chunks of FORCE_CHUNK slots in parallel (block b: slots [b*CHUNK, (b+1)*CHUNK)),
// then one block sums the chunk sums; both passes use a fixed order, so the result is
// deterministic and depends only on the slot values.
constexpr int FORCE_CHUNK = 4096;

__global__ void __launch_bounds__(256) k_force_chunks(const double *__restrict__ partials, int nslots,
                                                      double *__restrict__ out) {
    const int lo = blockIdx.x * FORCE_CHUNK, hi = min(nslots, lo + FORCE_CHUNK);
    double a = 0.0, b = 0.0, c = 0.0;
    // the FORCE_CHUNK / 256 = 16 slots of this thread, all loads in flight before the adds
    // (which keep the order j = lo + threadIdx.x, + 256, ...)
    constexpr int PER = FORCE_CHUNK / 256;
    double va[PER], vb[PER], vc[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int j = lo + (int)threadIdx.x + 256 * k;
        const size_t m = j < hi ? (size_t)j : (size_t)lo;
        va[k] = partials[3 * m];
        vb[k] = partials[3 * m + 1];
        vc[k] = partials[3 * m + 2];
    }
#pragma unroll
    for (int k = 0; k < PER; k++)
        if (lo + (int)threadIdx.x + 256 * k < hi) {
            a += va[k];
            b += vb[k];
            c += vc[k];
        }
    __shared__ double red[3 * 4];
    block_sum3<4>(a, b, c, red);
    if (threadIdx.x == 0) {
        out[3 * (size_t)blockIdx.x] = a;
        out[3 * (size_t)blockIdx.x + 1] = b;
        out[3 * (size_t)blockIdx.x + 2] = c;
    }
}

// final pass: fixed order, one block of 256 lanes
__global__ void __launch_bounds__(256) k_force_reduce(const double *__restrict__ partials, int nslots,
                                                      double *__restrict__ hist, int rec, long step) {
    double a = 0.0, b = 0.0, c = 0.0;
    for (int j = threadIdx.x; j < nslots; j += 256) {
        a += partials[3 * (size_t)j];
        b += partials[3 * (size_t)j + 1];
        c += partials[3 * (size_t)j + 2];
    }
    __shared__ double red[3 * 4];
    block_sum3<4>(a, b, c, red);
    if (threadIdx.x == 0) {
        hist[4 * (size_t)rec] = (double)step;
        hist[4 * (size_t)rec + 1] = a;
        hist[4 * (size_t)rec + 2] = b;
        hist[4 * (size_t)rec + 3] = c;
    }
}

template <class S>
__global__ void __launch_bounds__(512) k_decode(const typename S::T *__restrict__ A, double *__restrict__ out,
                                                const uint8_t *__restrict__ flags, float w, float off, Geo g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= g.nx) return;
    const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
    const size_t phys = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    const size_t e = (size_t)x + g.rs * ((size_t)y + (size_t)g.ny * z);   // A: this direction's element 0
    out[phys] = flags[node] == FL_SOLID ? 0.0 : (double)S::decode(A[e], w, off);
}

template <class S>
__global__ void __launch_bounds__(512) k_encode(typename S::T *__restrict__ A, const double *__restrict__ in,
                                                const uint8_t *__restrict__ flags, float winv, float offw, Geo g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= g.nx) return;
    const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
    const size_t phys = (size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z);
    const size_t e = (size_t)x + g.rs * ((size_t)y + (size_t)g.ny * z);
    A[e] = S::encode(flags[node] == FL_SOLID ? 0.0 : in[phys], winv, offw);
}

__global__ void __launch_bounds__(512) k_macros_extract(const double *__restrict__ mac, const uint8_t *__restrict__ flags,
                                                        double *__restrict__ out, Geo g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= g.nx) return;
    const size_t node = (size_t)x + (size_t)g.nxp * ((size_t)y + (size_t)g.ny * z);
    out[(size_t)x + (size_t)g.nx * ((size_t)y + (size_t)g.ny * z)] = flags[node] == FL_SOLID ? 0.0 : mac[node];
}

// Health check of a population buffer (lbm_sync): counts the non-finite values of one direction
// over a box of nodes, element base + x + sy*y + sz*z (x < nx; y, z the grid), skipping the nodes
// flagged solid when flags is set (flags at x + fpitch*(y + ny*z)).  A non-finite population
// anywhere reaches the rest population of the nodes that pull it within one step, through the
// moments, so one direction of the live nodes suffices.  16-bit formats are tested on their bits
// (exponent field all ones), FP32/FP64 with isfinite.
template <class T>
__device__ __forceinline__ bool finite_value(T v) {
    if constexpr (sizeof(T) == 8) return __builtin_isfinite((double)v);
    else if constexpr (sizeof(T) == 4) return __builtin_isfinite((float)v);
    else {
        uint16_t b;
        __builtin_memcpy(&b, &v, 2);
        return std::is_same<T, _Float16>::value ? (b & 0x7c00u) != 0x7c00u : (b & 0x7f80u) != 0x7f80u;
    }
}

template <class T>
__global__ void __launch_bounds__(256) k_count_nonfinite(const T *__restrict__ base, size_t sy, size_t sz, int nx,
                                                         const uint8_t *__restrict__ flags, int fpitch, int ny,
                                                         unsigned *__restrict__ bad) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    if (x >= nx) return;
    if (flags && flags[(size_t)x + (size_t)fpitch * ((size_t)y + (size_t)ny * z)] == FL_SOLID) return;
    if (!finite_value(base[(size_t)x + sy * y + sz * z])) atomicAdd(bad, 1u);
}

// ------------------------------------------------------------------------------------
// host launchers
// ------------------------------------------------------------------------------------
static dim3 grid_of(const Geo &g, const Shape &s) { return dim3(s.gx, g.ny, g.nz); }

template <class F>
static hipError_t dispatch_store(int prec, F &&f) {
    switch (prec) {
    case LBM_FP64: return f(SF64{});
    case LBM_FP32: return f(SF32{});
    case LBM_FP16S: return f(SF16{});
    case LBM_BF16S: return f(SBF16{});
    default: return hipErrorInvalidValue;
    }
}

template <int Q>
static hipError_t init_q(int prec, void *A, void *B, double *macros, const uint8_t *flags, const double *uin, const double *xi,
                         double perturb, uint64_t seed, const FusedArgs &a, hipStream_t s) {
    const Shape sh = node_shape(a.g);
    return dispatch_store(prec, [&](auto st) {
        using S = decltype(st);
        using T = typename S::T;
        hipLaunchKernelGGL((k_init<Q, S>), grid_of(a.g, sh), dim3(sh.bx), 0, s, (T *)A, (T *)B, macros, flags, uin,
                           xi, perturb, seed, a);
        return hipGetLastError();
    });
}

hipError_t plain_init(int Q, int prec, void *A, void *B, double *macros, const uint8_t *flags, const double *uin,
                      const double *xi, double perturb, uint64_t seed, const FusedArgs &a, hipStream_t s) {
    return Q == 9 ? init_q<9>(prec, A, B, macros, flags, uin, xi, perturb, seed, a, s)
                  : init_q<27>(prec, A, B, macros, flags, uin, xi, perturb, seed, a, s);
}

Shape fp64_shape(const Geo &g, int Q) {
    if ((size_t)8 * Q * g.plane <= ((size_t)16 << 20)) {
        Shape s;
        s.bx = 64;
        s.gx = g.nxp / 64;
        return s;
    }
    return lane_shape(g, 1);
}

Shape lane_shape(const Geo &g, int V) {
    Shape s;
    const int lanes = g.nxp / V;
    s.bx = lanes <= 256 ? lanes : 256;
    s.gx = (lanes + s.bx - 1) / s.bx;
    return s;
}

// FP64 storage: one node per lane, the reference's rounding order
template <int Q, bool FORCE, bool MACROS>
static hipError_t step_f64(const StepBufs &b, const FusedArgs &a, hipStream_t s) {
    using S = SF64;
    const Shape sh = fp64_shape(a.g, Q);
    if (b.ev_bulk[0]) {
        const hipError_t e = hipEventRecord(b.ev_bulk[0], s);
        if (e != hipSuccess) return e;
    }
    if (b.sfrac)
        hipLaunchKernelGGL((k_fused<Q, S, FORCE, MACROS, true>), grid_of(a.g, sh), dim3(sh.bx), b.fast_lds, s,
                           (const double *)b.A, (double *)b.B, b.flags, (const double *)b.sfrac, b.uin, b.block_slot,
                           b.partials, b.macros, a);
    else
        hipLaunchKernelGGL((k_fused<Q, S, FORCE, MACROS, false>), grid_of(a.g, sh), dim3(sh.bx), b.fast_lds, s,
                           (const double *)b.A, (double *)b.B, b.flags, (const double *)nullptr, b.uin, b.block_slot,
                           b.partials, b.macros, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && b.ev_bulk[1]) e = hipEventRecord(b.ev_bulk[1], s);
    return e;
}

template <int Q>
static hipError_t step_f64_q(const StepBufs &b, bool force, bool mac, const FusedArgs &a, hipStream_t s) {
    if (force && mac) return step_f64<Q, true, true>(b, a, s);
    if (force) return step_f64<Q, true, false>(b, a, s);
    if (mac) return step_f64<Q, false, true>(b, a, s);
    return step_f64<Q, false, false>(b, a, s);
}

hipError_t plain_step(int Q, int prec, int V, const StepBufs &b, bool force, bool macros_on, const FusedArgs &a,
                      hipStream_t s) {
    switch (prec) {
    case LBM_FP64: return Q == 9 ? step_f64_q<9>(b, force, macros_on, a, s) : step_f64_q<27>(b, force, macros_on, a, s);
    case LBM_FP32: return fast_step_f32(Q, V, b, force, macros_on, a, s);
    case LBM_FP16S: return fast_step_f16(Q, V, b, force, macros_on, a, s);
    case LBM_BF16S: return fast_step_bf16(Q, V, b, force, macros_on, a, s);
    default: return hipErrorInvalidValue;
    }
}

int force_chunks(int nslots) { return nslots > FORCE_CHUNK ? (nslots + FORCE_CHUNK - 1) / FORCE_CHUNK : 0; }

hipError_t force_reduce(const double *partials, int nslots, double *chunks, double *hist, int rec, long step,
                        hipStream_t s) {
    const int nc = force_chunks(nslots);
    if (nc > 0) {
        hipLaunchKernelGGL(k_force_chunks, dim3(nc), dim3(256), 0, s, partials, nslots, chunks);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
        partials = chunks;
        nslots = nc;
    }
    hipLaunchKernelGGL(k_force_reduce, dim3(1), dim3(256), 0, s, partials, nslots, hist, rec, step);
    return hipGetLastError();
}

template <int Q>
static double lat_w(int i) {
    double w = 0;
    static_for<0, Q>([&](auto I) {
        if (I == i) w = Lat<Q>::w(I);
    });
    return w;
}

hipError_t plain_decode(int Q, int prec, const void *A, double *out, const uint8_t *flags, int dir,
                        const FusedArgs &a, hipStream_t s) {
    const Shape sh = node_shape(a.g);
    const double w = Q == 9 ? lat_w<9>(dir) : lat_w<27>(dir);
    return dispatch_store(prec, [&](auto st) {
        using S = decltype(st);
        using T = typename S::T;
        hipLaunchKernelGGL(k_decode<S>, grid_of(a.g, sh), dim3(sh.bx), 0, s,
                           (const T *)A + (size_t)dir * a.g.ds, out, flags, (float)w, a.sp.off[dir], a.g);
        return hipGetLastError();
    });
}

hipError_t plain_encode(int Q, int prec, void *A, const double *in, const uint8_t *flags, int dir,
                        const FusedArgs &a, hipStream_t s) {
    const Shape sh = node_shape(a.g);
    const double w = Q == 9 ? lat_w<9>(dir) : lat_w<27>(dir);
    return dispatch_store(prec, [&](auto st) {
        using S = decltype(st);
        using T = typename S::T;
        hipLaunchKernelGGL(k_encode<S>, grid_of(a.g, sh), dim3(sh.bx), 0, s, (T *)A + (size_t)dir * a.g.ds,
                           in, flags, (float)(1.0 / w), a.sp.offw[dir], a.g);
        return hipGetLastError();
    });
}

hipError_t count_nonfinite(int prec, const void *base, size_t sy, size_t sz, int nx, int ny, int nz,
                           const uint8_t *flags, int fpitch, unsigned *bad, hipStream_t s) {
    hipError_t e = hipMemsetAsync(bad, 0, sizeof(unsigned), s);
    if (e != hipSuccess) return e;
    const dim3 grid((unsigned)((nx + 255) / 256), (unsigned)ny, (unsigned)nz);
    return dispatch_store(prec, [&](auto st) {
        using T = typename decltype(st)::T;
        hipLaunchKernelGGL(k_count_nonfinite<T>, grid, dim3(256), 0, s, (const T *)base, sy, sz, nx, flags, fpitch, ny,
                           bad);
        return hipGetLastError();
    });
}

hipError_t macros_extract(const double *macros, const uint8_t *flags, double *out, int k, const Geo &g, hipStream_t s) {
    const Shape sh = node_shape(g);
    hipLaunchKernelGGL(k_macros_extract, grid_of(g, sh), dim3(sh.bx), 0, s, macros + (size_t)k * g.plane, flags, out, g);
    return hipGetLastError();
}

}  // namespace lbm
