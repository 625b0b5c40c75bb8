// Complex arithmetic and compile-time trigonometry shared by the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fpcc {

template <typename R>
struct alignas(2 * sizeof(R)) C2 {
    R x, y;
};
using cf32 = C2<float>;
using cf64 = C2<double>;

template <typename R> __host__ __device__ __forceinline__ C2<R> mk(R a, R b) { return C2<R>{a, b}; }
template <typename R> __host__ __device__ __forceinline__ C2<R> operator+(C2<R> a, C2<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __host__ __device__ __forceinline__ C2<R> operator-(C2<R> a, C2<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __host__ __device__ __forceinline__ C2<R> operator*(R s, C2<R> a) { return {s * a.x, s * a.y}; }
template <typename R> __host__ __device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b)
{
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// conj(a) * b
template <typename R> __host__ __device__ __forceinline__ C2<R> cmulc(C2<R> a, C2<R> b)
{
    return {a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x};
}
template <typename R> __host__ __device__ __forceinline__ C2<R> conj(C2<R> a) { return {a.x, -a.y}; }
// multiply by DIR*i (DIR = -1 forward, +1 inverse): the radix-4 rotation
template <int DIR, typename R> __host__ __device__ __forceinline__ C2<R> rot(C2<R> a)
{
    return DIR < 0 ? C2<R>{a.y, -a.x} : C2<R>{-a.y, a.x};
}
template <typename R> __host__ __device__ __forceinline__ R norm2(C2<R> a) { return a.x * a.x + a.y * a.y; }

// ---- compile-time cos/sin of 2*pi*num/den (Taylor series on [-pi, pi]) ----
constexpr double kPi = 3.14159265358979323846;
constexpr double ct_sin_red(double x)
{
    double term = x, sum = x;
    for (int n = 1; n < 24; ++n) {
        term *= -x * x / ((2.0 * n) * (2.0 * n + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos_red(double x)
{
    double term = 1.0, sum = 1.0;
    for (int n = 1; n < 24; ++n) {
        term *= -x * x / ((2.0 * n - 1.0) * (2.0 * n));
        sum += term;
    }
    return sum;
}
constexpr double ct_angle(long num, long den)
{
    long m = ((num % den) + den) % den;
    double a = 2.0 * kPi * (double)m / (double)den;
    return a > kPi ? a - 2.0 * kPi : a;
}
constexpr double ct_cos2pi(long num, long den) { return ct_cos_red(ct_angle(num, den)); }
constexpr double ct_sin2pi(long num, long den) { return ct_sin_red(ct_angle(num, den)); }

}  // namespace fpcc
