// collide_f32.hpp -- FP32 collision on shifted populations.
//
// Storage holds d_i = f_i - w_i rho0 (rho0 the reference density), so FP32
// keeps its 24 bits for the deviations instead of spending them on the
// w_i rho0 bulk.  The operator is the same E-RLBM / RLBM / BGK step as
// collide.hpp (SimBoltz_Functions/Collision_operators.cpp:1181-1307,
// :859-967, :11-49), rewritten in deviations:
//   rho = rho0 + drho,  drho = sum_i d_i,  j = sum_i c_i d_i,  u = j / rho
//   d_eq_i = w_i (drho + rho (3 c.u + 4.5 (c.u)^2 - 1.5 u.u))
//   Pi = sum_i (d_i - d_eq_i) c_i c_i,   f1_i = 4.5 w_i Q_i : Pi
//   d*_i = d_eq_i + (1 - 1/tau) f1_i + (1 - 1/tau_nh) (d_i - d_eq_i - f1_i)
// The moment sums use Neumaier compensation; FMA is used explicitly.
//
// Momentum of the equilibrium.  Its odd part 3 w_i rho (c.u) carries the
// momentum j; formed in FP32 as written it does not reproduce j: the weights
// round (sum_i 3 w_i c_ix^2 = 1 + O(1e-8)) and so does rho = rho0 + drho near
// 1, whose rounding u = j / rho does not undo when multiplied back.  Both
// errors repeat every step with the same sign, so they grow linearly with
// the step count: +18% on the Taylor-Green L = 64 velocity error after its
// 2,048 steps.  So the odd part is formed as b_i (c.j) -- rho (c.u) is c.j
// exactly -- with coefficients b_axis / b_diag (3 w_i up to one FP32 unit)
// chosen so that sum_i b_i c_ix^2 = 2 b_axis + K b_diag is exactly 1 (K = 4
// for D2Q9, 8 for D3Q19): the momentum of d_eq is j up to rounding that does
// not repeat (Taylor-Green L = 64: +0.43%).
#pragma once

#include "collide.hpp"
#include "vset.hpp"

namespace erlbm {

struct CollideConstsF32 {
    float omega1, omega2, oot, rho0;
};

// Odd-part coefficients of the FP32 equilibrium (the 3 w_i of c.u terms):
// axis directions (|c|^2 = 1) and diagonals (|c|^2 = 2).
template <int Q>
struct OddCoefF32;
template <>
struct OddCoefF32<9> {   // 1/3 and 1/12, 2 b_axis + 4 b_diag == 1 exactly
    static constexpr float axis = 0x1.55555cp-2f, diag = 0x1.555548p-4f;
};
template <>
struct OddCoefF32<19> {   // 1/6 and 1/12, 2 b_axis + 8 b_diag == 1 exactly
    static constexpr float axis = 0x1.555554p-3f, diag = 0x1.555556p-4f;
};
static_assert(2.0 * (double)OddCoefF32<9>::axis + 4.0 * (double)OddCoefF32<9>::diag == 1.0,
              "D2Q9 odd coefficients must carry the momentum exactly");
static_assert(2.0 * (double)OddCoefF32<19>::axis + 8.0 * (double)OddCoefF32<19>::diag == 1.0,
              "D3Q19 odd coefficients must carry the momentum exactly");

// Neumaier-compensated accumulation s += v (c holds the lost low part).
__device__ __forceinline__ void kadd(float &s, float &c, float v)
{
    const float t = s + v;
    if (fabsf(s) >= fabsf(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
}

// drho, rho and u; jx, jy, jz receive the momentum (compensated sums).
template <int Q>
__device__ __forceinline__ void moments_shifted(const float (&d)[Q], float rho0, float &drho, float &rho, float &vx,
                                                float &vy, float &vz, float &jx, float &jy, float &jz)
{
    using V = VSet<Q>;
    float r = 0.f, rc = 0.f, mx = 0.f, mxc = 0.f, my = 0.f, myc = 0.f, mz = 0.f, mzc = 0.f;
    static_for<Q>([&](auto I) {
        constexpr int i = I;
        kadd(r, rc, d[i]);
        if constexpr (V::cx[i] != 0) kadd(mx, mxc, V::cx[i] > 0 ? d[i] : -d[i]);
        if constexpr (V::cy[i] != 0) kadd(my, myc, V::cy[i] > 0 ? d[i] : -d[i]);
        if constexpr (V::cz[i] != 0) kadd(mz, mzc, V::cz[i] > 0 ? d[i] : -d[i]);
    });
    drho = r + rc;
    rho = rho0 + drho;
    jx = mx + mxc;
    jy = my + myc;
    jz = (V::D == 3) ? mz + mzc : 0.f;
    const float oor = 1.0f / rho;
    vx = jx * oor;
    vy = jy * oor;
    vz = jz * oor;
}

template <int Q, int I>
__device__ __forceinline__ float cdotu_f(float vx, float vy, float vz)
{
    using V = VSet<Q>;
    float cv = 0.f;
    if constexpr (V::cx[I] != 0) cv += V::cx[I] > 0 ? vx : -vx;
    if constexpr (V::cy[I] != 0) cv += V::cy[I] > 0 ? vy : -vy;
    if constexpr (V::cz[I] != 0) cv += V::cz[I] > 0 ? vz : -vz;
    return cv;
}

template <int Q, int COL>
__device__ __forceinline__ void collide_shifted_f32(float (&d)[Q], const CollideConstsF32 &k, float &drho_out,
                                                    float &vx, float &vy, float &vz)
{
    using V = VSet<Q>;
    float drho, rho, jx, jy, jz;
    moments_shifted<Q>(d, k.rho0, drho, rho, vx, vy, vz, jx, jy, jz);
    drho_out = drho;
    const float s = 1.5f * (vx * vx + vy * vy + vz * vz);
    float deq[Q];
    deq[0] = (float)V::w[0] * fmaf(-rho, s, drho);
    static_for<(Q - 1) / 2>([&](auto P) {
        constexpr int i = 2 * P + 1;
        constexpr int c2 = V::cx[i] * V::cx[i] + V::cy[i] * V::cy[i] + V::cz[i] * V::cz[i];
        constexpr float b = c2 == 1 ? OddCoefF32<Q>::axis : OddCoefF32<Q>::diag;
        const float cv = cdotu_f<Q, i>(vx, vy, vz);
        const float qq = fmaf(4.5f * cv, cv, -s);   // 4.5 cv^2 - 1.5 u.u
        const float even = (float)V::w[i] * fmaf(rho, qq, drho);
        const float odd = b * cdotu_f<Q, i>(jx, jy, jz);   // 3 w_i rho (c.u) = 3 w_i (c.j)
        deq[i] = even + odd;
        deq[i + 1] = even - odd;
    });
    if constexpr (COL == COL_BGK) {
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            d[i] = fmaf(deq[i] - d[i], k.oot, d[i]);
        });
        return;
    } else {
        float n[Q];
        float Pxx = 0.f, Pyy = 0.f, Pzz = 0.f, Pxy = 0.f, Pxz = 0.f, Pyz = 0.f;
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            n[i] = d[i] - deq[i];
            if constexpr (V::cx[i] != 0) Pxx += n[i];
            if constexpr (V::cy[i] != 0) Pyy += n[i];
            if constexpr (V::D == 3 && V::cz[i] != 0) Pzz += n[i];
            if constexpr (V::cx[i] * V::cy[i] > 0) Pxy += n[i];
            if constexpr (V::cx[i] * V::cy[i] < 0) Pxy -= n[i];
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] > 0) Pxz += n[i];
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] < 0) Pxz -= n[i];
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] > 0) Pyz += n[i];
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] < 0) Pyz -= n[i];
        });
        // Q_aa : Pi = (c_a^2 - 1/3) Pi_aa; trace part shared
        const float tr = (Pxx + Pyy + Pzz) * (1.f / 3.f);
        static_for<Q>([&](auto I) {
            constexpr int i = I;
            float t = -tr;
            if constexpr (V::cx[i] != 0) t += Pxx;
            if constexpr (V::cy[i] != 0) t += Pyy;
            if constexpr (V::D == 3 && V::cz[i] != 0) t += Pzz;
            if constexpr (V::cx[i] * V::cy[i] > 0) t = fmaf(2.f, Pxy, t);
            if constexpr (V::cx[i] * V::cy[i] < 0) t = fmaf(-2.f, Pxy, t);
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] > 0) t = fmaf(2.f, Pxz, t);
            if constexpr (V::D == 3 && V::cx[i] * V::cz[i] < 0) t = fmaf(-2.f, Pxz, t);
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] > 0) t = fmaf(2.f, Pyz, t);
            if constexpr (V::D == 3 && V::cy[i] * V::cz[i] < 0) t = fmaf(-2.f, Pyz, t);
            const float f1 = (float)(4.5 * V::w[i]) * t;
            if constexpr (COL == COL_REG) {
                d[i] = fmaf(k.omega1, f1, deq[i]);
            } else {
                d[i] = fmaf(k.omega2, n[i] - f1, fmaf(k.omega1, f1, deq[i]));
            }
        });
    }
}

// Density from shifted populations in FP64 (the stability tests' reduction).
template <int Q>
__device__ __forceinline__ double density_shifted_f64(const float (&d)[Q], double rho0)
{
    double r = 0.0;
    static_for<Q>([&](auto I) { r += (double)d[(int)I]; });
    return rho0 + r;
}

}  // namespace erlbm
