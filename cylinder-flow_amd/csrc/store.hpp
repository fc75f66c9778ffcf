// store.hpp -- population storage formats.
//
// All formats hold the reference's shifted populations f_i = f_i^std - w_i*rho0
// (3D MPI+KOKKOS+MPIIO/lbm.cpp:263-267).  The 16-bit formats additionally subtract a
// per-direction offset o_i = feq_i(0, (u_ref,0,0)) (0 when u_ref = 0) and divide by w_i,
// so every direction spans the same range and the free stream is stored near zero:
//     h_i = round16((f_i - o_i) / w_i),   f_i = w_i * h_i + o_i   (FP32 arithmetic)
// Opposite directions share w_i, so a bounced-back value is decoded with the weight of
// the pulled direction and the offset of the direction it was stored under.
// Collision runs in FP32 for the FP32 and 16-bit formats and in FP64 for FP64.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

struct StoreParams {
    float off[27];    // o_i
    float offw[27];   // o_i / w_i
};

struct SF64 {
    using T = double;
    using C = double;
    static constexpr int bytes = 8;
    template <int I, class L>
    __device__ __forceinline__ static C decode(T h, float) { return h; }
    __device__ __forceinline__ static C decode(T h, float, float) { return h; }
    template <int I, class L>
    __device__ __forceinline__ static T enc(C v, const StoreParams &) { return v; }
    template <int I, class L>
    __device__ __forceinline__ static void store(T *p, C v, const StoreParams &) { __builtin_nontemporal_store(v, p); }
    __device__ __forceinline__ static T encode(double v, float, float) { return v; }
};

struct SF32 {
    using T = float;
    using C = float;
    static constexpr int bytes = 4;
    template <int I, class L>
    __device__ __forceinline__ static C decode(T h, float) { return h; }
    __device__ __forceinline__ static C decode(T h, float, float) { return h; }
    template <int I, class L>
    __device__ __forceinline__ static T enc(C v, const StoreParams &) { return v; }
    template <int I, class L>
    __device__ __forceinline__ static void store(T *p, C v, const StoreParams &) { __builtin_nontemporal_store(v, p); }
    __device__ __forceinline__ static T encode(double v, float, float) { return (float)v; }
};

// IEEE binary16, round-to-nearest-even on store
struct SF16 {
    using T = _Float16;
    using C = float;
    static constexpr int bytes = 2;
    template <int I, class L>
    __device__ __forceinline__ static C decode(T h, float off) {
        return __builtin_fmaf((float)h, (float)L::w(I), off);
    }
    __device__ __forceinline__ static C decode(T h, float w, float off) { return __builtin_fmaf((float)h, w, off); }
    // The empty asm keeps the FP32 result in a register, so the compiler cannot fuse the FMA
    // and the f16 conversion into one mixed-precision instruction: every value is rounded to
    // FP32 and then to FP16 exactly as written, whichever code path encodes it.
    __device__ __forceinline__ static T round16(float t) {
        asm volatile("" : "+v"(t));
        return (T)t;
    }
    template <int I, class L>
    __device__ __forceinline__ static T enc(C v, const StoreParams &sp) {
        return round16(__builtin_fmaf(v, (float)(1.0 / L::w(I)), -sp.offw[I]));
    }
    template <int I, class L>
    __device__ __forceinline__ static void store(T *p, C v, const StoreParams &sp) {
        __builtin_nontemporal_store(enc<I, L>(v, sp), p);
    }
    __device__ __forceinline__ static T encode(double v, float winv, float offw) {
        return round16(__builtin_fmaf((float)v, winv, -offw));
    }
};

// bfloat16 (upper half of an FP32), round-to-nearest-even on store
struct SBF16 {
    using T = uint16_t;
    using C = float;
    static constexpr int bytes = 2;
    __device__ __forceinline__ static float widen(T h) { return __uint_as_float((uint32_t)h << 16); }
    __device__ __forceinline__ static T narrow(float f) {
        uint32_t u = __float_as_uint(f);
        u += 0x7fffu + ((u >> 16) & 1u);
        return (T)(u >> 16);
    }
    template <int I, class L>
    __device__ __forceinline__ static C decode(T h, float off) {
        return __builtin_fmaf(widen(h), (float)L::w(I), off);
    }
    __device__ __forceinline__ static C decode(T h, float w, float off) { return __builtin_fmaf(widen(h), w, off); }
    template <int I, class L>
    __device__ __forceinline__ static T enc(C v, const StoreParams &sp) {
        return narrow(__builtin_fmaf(v, (float)(1.0 / L::w(I)), -sp.offw[I]));
    }
    template <int I, class L>
    __device__ __forceinline__ static void store(T *p, C v, const StoreParams &sp) {
        __builtin_nontemporal_store(enc<I, L>(v, sp), p);
    }
    __device__ __forceinline__ static T encode(double v, float winv, float offw) {
        return narrow(__builtin_fmaf((float)v, winv, -offw));
    }
};

}  // namespace lbm
