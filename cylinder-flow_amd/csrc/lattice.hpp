// lattice.hpp -- compile-time D2Q9 / D3Q27 velocity sets in the reference's ordering.
//
// D2Q9 : MPI Kokkos/lbm.cpp:52-90   (0; +x,+y,-x,-y; (1,1),(-1,1),(-1,-1),(1,-1))
// D3Q27: 3D MPI+KOKKOS+MPIIO/lbm.cpp:55-218 (0; axis pairs; edge pairs; corner pairs,
//        each direction followed by its opposite)
// The ordering is kept so populations, fixtures and the reference-mode fills line up
// index for index.  Every lookup is a constant expression: kernels iterate directions
// with static_for, so e_i, w_i and opp(i) become immediates.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>

namespace lbm {

template <int Q> struct Lat;

template <> struct Lat<9> {
    static constexpr int D = 2;
    __host__ __device__ static constexpr int ex(int i) {
        return i == 1 || i == 5 || i == 8 ? 1 : (i == 3 || i == 6 || i == 7 ? -1 : 0);
    }
    __host__ __device__ static constexpr int ey(int i) {
        return i == 2 || i == 5 || i == 6 ? 1 : (i == 4 || i == 7 || i == 8 ? -1 : 0);
    }
    __host__ __device__ static constexpr int ez(int) { return 0; }
    // 1<->3, 2<->4, 5<->7, 6<->8 (MPI Kokkos/lbm.cpp:61-69)
    __host__ __device__ static constexpr int opp(int i) {
        return i == 0 ? 0 : (i <= 4 ? (i + 1) % 4 + 1 : (i - 3) % 4 + 5);
    }
    __host__ __device__ static constexpr double w(int i) {
        return i == 0 ? 4.0 / 9.0 : (i <= 4 ? 1.0 / 9.0 : 1.0 / 36.0);
    }
};

template <> struct Lat<27> {
    static constexpr int D = 3;
    // components packed as (ex+1) + 3*(ey+1) + 9*(ez+1) for each index of the reference
    __host__ __device__ static constexpr int code(int i) {
        constexpr int c[27] = {13, 14, 12, 16, 10, 22, 4,  17, 9,  11, 15, 23, 3, 5,
                               21, 25, 1,  7,  19, 26, 0,  20, 6,  8,  18, 2,  24};
        return c[i];
    }
    __host__ __device__ static constexpr int ex(int i) { return code(i) % 3 - 1; }
    __host__ __device__ static constexpr int ey(int i) { return (code(i) / 3) % 3 - 1; }
    __host__ __device__ static constexpr int ez(int i) { return code(i) / 9 - 1; }
    __host__ __device__ static constexpr int opp(int i) { return i == 0 ? 0 : (i % 2 ? i + 1 : i - 1); }
    __host__ __device__ static constexpr double w(int i) {
        return i == 0 ? 8.0 / 27.0 : (i <= 6 ? 2.0 / 27.0 : (i <= 18 ? 1.0 / 54.0 : 1.0 / 216.0));
    }
};

// compile-time loop: f(std::integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__host__ __device__ __forceinline__ void static_for(F &&f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(static_cast<F &&>(f));
    }
}

// self-checks of the tables (isotropy, opposites) evaluated by the compiler
template <int Q> constexpr bool lattice_ok() {
    using L = Lat<Q>;
    double s = 0, sxx = 0, syy = 0, sxy = 0;
    for (int i = 0; i < Q; i++) {
        if (L::ex(L::opp(i)) != -L::ex(i) || L::ey(L::opp(i)) != -L::ey(i) || L::ez(L::opp(i)) != -L::ez(i))
            return false;
        if (L::opp(L::opp(i)) != i) return false;
        s += L::w(i);
        sxx += L::w(i) * L::ex(i) * L::ex(i);
        syy += L::w(i) * L::ey(i) * L::ey(i);
        sxy += L::w(i) * L::ex(i) * L::ey(i);
    }
    auto close = [](double a, double b) { return a - b < 1e-15 && b - a < 1e-15; };
    return close(s, 1.0) && close(sxx, 1.0 / 3.0) && close(syy, 1.0 / 3.0) && close(sxy, 0.0);
}
static_assert(lattice_ok<9>(), "D2Q9 table");
static_assert(lattice_ok<27>(), "D3Q27 table");
static_assert(Lat<9>::opp(1) == 3 && Lat<9>::opp(2) == 4 && Lat<9>::opp(5) == 7 && Lat<9>::opp(6) == 8,
              "D2Q9 opposites follow MPI Kokkos/lbm.cpp:61-69");
static_assert(Lat<27>::ex(19) == 1 && Lat<27>::ey(19) == 1 && Lat<27>::ez(19) == 1 &&
                  Lat<27>::ex(26) == -1 && Lat<27>::ey(26) == 1 && Lat<27>::ez(26) == 1,
              "D3Q27 corners follow 3D MPI+KOKKOS+MPIIO/lbm.cpp:191-217");

}  // namespace lbm
