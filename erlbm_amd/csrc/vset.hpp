// vset.hpp -- D2Q9 / D3Q19 velocity sets as compile-time tables.
//
// Velocity order and weights follow the reference's lattice definitions
// (SimBoltz_Functions/LBM_functions.cpp:1437-1501 for D2Q9, :577-708 for
// D3Q19): opposite directions are consecutive pairs (1,2), (3,4), ...
// Weights are the IEEE-rounded quotients 4./9., 1./9., 1./36. (D2Q9) and
// 1./3., 1./18., 1./36. (D3Q19), exactly as the reference computes them.
#pragma once

#include <cstdint>
#include <type_traits>
#include <utility>

namespace erlbm {

template <int Q>
struct VSet;

template <>
struct VSet<9> {
    static constexpr int D = 2;
    static constexpr int cx[9] = {0, 1, -1, 0, 0, 1, -1, 1, -1};
    static constexpr int cy[9] = {0, 0, 0, 1, -1, 1, -1, -1, 1};
    static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int opp[9] = {0, 2, 1, 4, 3, 6, 5, 8, 7};
    static constexpr double w[9] = {4. / 9.,  1. / 9.,  1. / 9.,  1. / 9., 1. / 9.,
                                    1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

template <>
struct VSet<19> {
    static constexpr int D = 3;
    static constexpr int cx[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int cy[19] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, -1, 1, -1, 1};
    static constexpr int cz[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, 1, -1};
    static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
    static constexpr double w[19] = {1. / 3.,  1. / 18., 1. / 18., 1. / 18., 1. / 18.,
                                     1. / 18., 1. / 18., 1. / 36., 1. / 36., 1. / 36.,
                                     1. / 36., 1. / 36., 1. / 36., 1. / 36., 1. / 36.,
                                     1. / 36., 1. / 36., 1. / 36., 1. / 36.};
};

// Compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, in
// order (the comma fold evaluates left to right), so that sums keep the
// reference's accumulation order and `if constexpr` can drop zero terms.
template <class F, int... I>
__host__ __device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, I>{}), ...);
}

template <int N, class F>
__host__ __device__ __forceinline__ void static_for(F &&f)
{
    static_for_impl(f, std::make_integer_sequence<int, N>{});
}

}  // namespace erlbm
