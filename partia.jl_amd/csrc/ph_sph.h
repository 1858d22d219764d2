// ph_sph.h -- SPH kernel functions on the device.
//
// Evaluation order follows the reference expression by expression so that the
// float64 path reproduces the reference's Float64 arithmetic (all HIP sources
// are compiled with -ffp-contract=off; explicit fma() is used only where a
// fused product is intended, in the FP32-compensated path):
//   shapes   KI/kernel_function/kernels/{M4,M5,M6}_spline.jl:13-37, {C2,C4,C6}_Wendland.jl:13-36
//   sigma_3  same files, KernelFunctionnorm(K, Val(3), T)
//   W        KI/kernel_function/kernel.jl:141-146 (with :86-88)
//   grad W   KI/kernel_function/kernel.jl:273-285, 360-365
//   line LUT KI/kernel_function/line_integrated_kernel.jl:13-19, 90-94, 132-137
// Integer powers: ^2 and ^3 are x*x and x*x*x (Julia's literal_pow); higher
// powers are repeated multiplication (within 1 ulp of Julia's compensated `^`).
#pragma once

#include <type_traits>

#include "ph_common.h"

namespace ph {

template <class T> __host__ __device__ inline T q2(T x) { return x * x; }
template <class T> __host__ __device__ inline T q3(T x) { return x * x * x; }
template <class T> __host__ __device__ inline T q4(T x) { T x2 = x * x; return x2 * x2; }
template <class T> __host__ __device__ inline T q5(T x) { return q4(x) * x; }
template <class T> __host__ __device__ inline T q6(T x) { T x2 = x * x; return (x2 * x2) * x2; }
template <class T> __host__ __device__ inline T q7(T x) { return q6(x) * x; }
template <class T> __host__ __device__ inline T q8(T x) { T x4 = q4(x); return x4 * x4; }

template <class T> __host__ __device__ inline T sph_nan();
template <> __host__ __device__ inline double sph_nan<double>() { return __builtin_nan(""); }
template <> __host__ __device__ inline float sph_nan<float>() { return __builtin_nanf(""); }

// Dimensionless shape f(q).
template <class T, int K> __host__ __device__ inline T shape_f(T q) {
    if (q < T(0)) return sph_nan<T>();
    if constexpr (K == PH_KERNEL_M4) {
        if (q < T(1)) return T(0.25) * q3(T(2) - q) - q3(T(1) - q);
        if (q < T(2)) return T(0.25) * q3(T(2) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_M5) {
        if (q < T(0.5)) return q4(T(2.5) - q) - T(5) * q4(T(1.5) - q) + T(10) * q4(T(0.5) - q);
        if (q < T(1.5)) return q4(T(2.5) - q) - T(5) * q4(T(1.5) - q);
        if (q < T(2.5)) return q4(T(2.5) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_M6) {
        if (q < T(1)) return q5(T(3) - q) - T(6) * q5(T(2) - q) + T(15) * q5(T(1) - q);
        if (q < T(2)) return q5(T(3) - q) - T(6) * q5(T(2) - q);
        if (q < T(3)) return q5(T(3) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_C2) {
        if (q < T(2)) return q4(T(1) - T(0.5) * q) * (T(2) * q + T(1));
        return T(0);
    } else if constexpr (K == PH_KERNEL_C4) {
        if (q < T(2)) return q6(T(1) - T(0.5) * q) * ((T(35) / T(12)) * q2(q) + T(3) * q + T(1));
        return T(0);
    } else {
        if (q < T(2)) return q8(T(1) - T(0.5) * q) * (T(4) * q3(q) + T(6.25) * q2(q) + T(4) * q + T(1));
        return T(0);
    }
}

// Derivative f'(q).
template <class T, int K> __host__ __device__ inline T shape_df(T q) {
    if (q < T(0)) return sph_nan<T>();
    if constexpr (K == PH_KERNEL_M4) {
        if (q < T(1)) return T(-0.75) * q2(T(2) - q) + T(3) * q2(T(1) - q);
        if (q < T(2)) return T(-0.75) * q2(T(2) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_M5) {
        if (q < T(0.5)) return T(-4) * q3(T(2.5) - q) + T(20) * q3(T(1.5) - q) - T(40) * q3(T(0.5) - q);
        if (q < T(1.5)) return T(-4) * q3(T(2.5) - q) + T(20) * q3(T(1.5) - q);
        if (q < T(2.5)) return T(-4) * q3(T(2.5) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_M6) {
        if (q < T(1)) return T(-5) * q4(T(3) - q) + T(30) * q4(T(2) - q) - T(75) * q4(T(1) - q);
        if (q < T(2)) return T(-5) * q4(T(3) - q) + T(30) * q4(T(2) - q);
        if (q < T(3)) return T(-5) * q4(T(3) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_C2) {
        if (q < T(2)) return q4(T(1) - T(0.5) * q) * T(2) - q3(T(1) - T(0.5) * q) * (T(4) * q + T(2));
        return T(0);
    } else if constexpr (K == PH_KERNEL_C4) {
        if (q < T(2))
            return q6(T(1) - T(0.5) * q) * ((T(35) / T(6)) * q + T(3)) -
                   q5(T(1) - T(0.5) * q) * ((T(35) / T(4)) * q2(q) + T(9) * q + T(3));
        return T(0);
    } else {
        if (q < T(2))
            return q8(T(1) - T(0.5) * q) * (T(12) * q2(q) + T(12.5) * q + T(4)) -
                   q7(T(1) - T(0.5) * q) * (T(16) * q3(q) + T(25) * q2(q) + T(16) * q + T(4));
        return T(0);
    }
}

// Second derivative f''(q), the derivative of shape_df (used by the adjoints of
// gradient, divergence and curl outputs).  Wendland shapes t^k g(q), t = 1 - q/2:
// f'' = t^(k-2) (t^2 g'' - k t g' + k (k - 1) / 4 g).
template <class T, int K> __host__ __device__ inline T shape_d2f(T q) {
    if (q < T(0)) return sph_nan<T>();
    const T t = T(1) - T(0.5) * q;
    if constexpr (K == PH_KERNEL_M4) {
        if (q < T(1)) return T(1.5) * (T(2) - q) - T(6) * (T(1) - q);
        if (q < T(2)) return T(1.5) * (T(2) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_M5) {
        if (q < T(0.5)) return T(12) * q2(T(2.5) - q) - T(60) * q2(T(1.5) - q) + T(120) * q2(T(0.5) - q);
        if (q < T(1.5)) return T(12) * q2(T(2.5) - q) - T(60) * q2(T(1.5) - q);
        if (q < T(2.5)) return T(12) * q2(T(2.5) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_M6) {
        if (q < T(1)) return T(20) * q3(T(3) - q) - T(120) * q3(T(2) - q) + T(300) * q3(T(1) - q);
        if (q < T(2)) return T(20) * q3(T(3) - q) - T(120) * q3(T(2) - q);
        if (q < T(3)) return T(20) * q3(T(3) - q);
        return T(0);
    } else if constexpr (K == PH_KERNEL_C2) {  // k = 4, g = 2q + 1
        if (q < T(2)) return q2(t) * (T(3) * (T(2) * q + T(1)) - T(8) * t);
        return T(0);
    } else if constexpr (K == PH_KERNEL_C4) {  // k = 6, g = 35/12 q^2 + 3q + 1
        if (!(q < T(2))) return T(0);
        const T g = (T(35) / T(12)) * q2(q) + T(3) * q + T(1), g1 = (T(35) / T(6)) * q + T(3);
        return q4(t) * (q2(t) * (T(35) / T(6)) - T(6) * t * g1 + T(7.5) * g);
    } else {  // k = 8, g = 4q^3 + 6.25q^2 + 4q + 1
        if (!(q < T(2))) return T(0);
        const T g = T(4) * q3(q) + T(6.25) * q2(q) + T(4) * q + T(1), g1 = T(12) * q2(q) + T(12.5) * q + T(4),
                g2 = T(24) * q + T(12.5);
        return q6(t) * (q2(t) * g2 - T(8) * t * g1 + T(14) * g);
    }
}

template <int K> __host__ __device__ constexpr double support_of() {
    return K == PH_KERNEL_M5 ? 2.5 : (K == PH_KERNEL_M6 ? 3.0 : 2.0);
}

template <class T, int K> __host__ __device__ inline T sigma3() {
    const T pi = T(3.14159265358979323846);
    if constexpr (K == PH_KERNEL_M4) return T(1) / pi;
    else if constexpr (K == PH_KERNEL_M5) return T(0.05) / pi;
    else if constexpr (K == PH_KERNEL_M6) return T(1) / (T(120) * pi);
    else if constexpr (K == PH_KERNEL_C2) return T(21) / (T(16) * pi);
    else if constexpr (K == PH_KERNEL_C4) return T(495) / (T(256) * pi);
    else return T(1365) / (T(512) * pi);
}

// W(r, h) in 3D.
template <class T, int K> __device__ inline T kernel_W(T r, T h) {
    T invh = T(1) / h;
    T inv_hD = invh * invh * invh;
    T q = r * invh;
    return inv_hD * (sigma3<T, K>() * shape_f<T, K>(q));
}

// grad W (dx, dy, dz, h) in 3D; zero when r^2 == 0.
template <class T, int K> __device__ inline void kernel_gradW(T dx, T dy, T dz, T h, T& gx, T& gy, T& gz) {
    T invh = T(1) / h;
    T inv_hDp1 = invh * invh * invh * invh;
    T r2 = dx * dx + dy * dy + dz * dz;
    T ax = T(0), ay = T(0), az = T(0);
    if (!(r2 == T(0))) {
        T r = sqrt(r2);
        T invr = T(1) / r;
        T q = r / h;
        T coeff = shape_df<T, K>(q) * sigma3<T, K>() * invr;
        ax = dx * coeff;
        ay = dy * coeff;
        az = dz * coeff;
    }
    gx = inv_hDp1 * ax;
    gy = inv_hDp1 * ay;
    gz = inv_hDp1 * az;
}

// Line-integrated kernel value (1/h) * I(r/h) with the 4096-entry table.
template <class T, int K> __device__ inline T line_W(T r, T h, const T* __restrict__ inorm, T dq) {
    T invh = T(1) / h;
    T q = r * invh;
    T I = T(0);
    if (!(q >= T(support_of<K>()))) {
        T idxf = q / dq + T(1);
        i32 i = (i32)idxf;
        i = i < 1 ? 1 : (i > PH_LUT_N_ENTRIES - 1 ? PH_LUT_N_ENTRIES - 1 : i);
        T t = idxf - T(i);
        I = inorm[i - 1] * (T(1) - t) + inorm[i] * t;
    }
    return invh * I;
}

// line_W without the branch, from invh = T(1) / h: the table index is clamped
// for every q (a NaN or huge q reads entry 1 or N-1), and the cutoff selects 0,
// so the value equals line_W for every q.
template <class T, int K> __device__ inline T line_W_inv(T r, T invh, const T* __restrict__ inorm, T dq) {
    T q = r * invh;
    T idxf = q / dq + T(1);
    const T lo = T(1), hi = T(PH_LUT_N_ENTRIES - 1);
    const T cf = idxf < lo ? lo : (idxf > hi ? hi : idxf);  // NaN stays NaN: converted to 0 below, then clamped
    i32 i = (i32)cf;
    i = i < 1 ? 1 : (i > PH_LUT_N_ENTRIES - 1 ? PH_LUT_N_ENTRIES - 1 : i);
    T t = idxf - T(i);
    T I = inorm[i - 1] * (T(1) - t) + inorm[i] * t;
    I = (q >= T(support_of<K>())) ? T(0) : I;
    return invh * I;
}
template <class T, int K> __device__ inline T line_W_nb(T r, T h, const T* __restrict__ inorm, T dq) {
    return line_W_inv<T, K>(r, T(1) / h, inorm, dq);
}
// line_W_inv with the float32 table argument q / dq + 1 formed without the
// division: rdq = RN(1 / dq), q0 = q rdq, q0 + (q - q0 dq) rdq (two fmas;
// Markstein's correction).  For the three float32 table steps (supports 2,
// 2.5 and 3) fl(that + 1) equals fl(q / dq + 1) for every float32 q with
// |q| <= 16 (checked exhaustively on the host), and an accepted neighbour has
// q = r / h < 2 K <= 6 (r^2 <= (K h)^2, or (K h)^2 overflows and r^2 does
// not); other lanes' values are discarded.  Float64 keeps the division.
template <class T, int K>
__device__ inline T line_W_rdq(T r, T invh, const T* __restrict__ inorm, T dq, T rdq) {
    if constexpr (!std::is_same<T, float>::value) {
        (void)rdq;
        return line_W_inv<T, K>(r, invh, inorm, dq);
    } else {
        const float q = r * invh;
        const float q0 = q * rdq;
        const float e = __builtin_fmaf(-q0, dq, q);
        const float idxf = __builtin_fmaf(e, rdq, q0) + 1.f;
        // line_W_inv's two clamps give the same i in one each: truncation is
        // monotonic and both bounds are integers (NaN: i = 1 there, here the
        // conversion's 0 raised to 1 or a bound; t and I are NaN either way)
        const float cf = __builtin_amdgcn_fmed3f(idxf, 1.f, float(PH_LUT_N_ENTRIES - 1));
        const i32 i = max((i32)cf, 1);
        const float t = idxf - float(i);
        float I = inorm[i - 1] * (1.f - t) + inorm[i] * t;
        I = (q >= float(support_of<K>())) ? 0.f : I;
        return invh * I;
    }
}

}  // namespace ph
