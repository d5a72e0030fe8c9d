// Common device/host utilities for the DGFS HIP kernels (gfx950).
//
//  * cplx<T>: interleaved complex value (8 B for float, 16 B for double).
//  * Compile-time sine/cosine of rational multiples of 2*pi (own Taylor series
//    after exact octant reduction), so FFT twiddles are literal constants.
//  * sfor<B,E>(f): compile-time loop; f receives std::integral_constant<int,i>.
//  * Error plumbing for the C ABI: every entry point returns int status,
//    message retrievable through dgfs_last_error().
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <type_traits>
#include <utility>

#define DGFS_DEV __device__ __forceinline__
#define DGFS_HD __host__ __device__ __forceinline__

namespace dgfs {

// ---------------------------------------------------------------- complex
template <typename T>
struct alignas(2 * sizeof(T)) cplx {
    T x, y;
};

template <typename T> DGFS_HD cplx<T> mk(T a, T b) { cplx<T> r; r.x = a; r.y = b; return r; }
template <typename T> DGFS_HD cplx<T> operator+(cplx<T> a, cplx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T> DGFS_HD cplx<T> operator-(cplx<T> a, cplx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T> DGFS_HD cplx<T> operator*(cplx<T> a, cplx<T> b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename T> DGFS_HD cplx<T> operator*(T s, cplx<T> a) { return mk(s * a.x, s * a.y); }
template <typename T> DGFS_HD cplx<T> conj(cplx<T> a) { return mk(a.x, -a.y); }
// multiply by (sgn * i):  i*(x+iy) = -y + ix
template <int SGN, typename T> DGFS_HD cplx<T> mul_i(cplx<T> a) {
    return SGN > 0 ? mk(-a.y, a.x) : mk(a.y, -a.x);
}
template <typename T> DGFS_HD cplx<T> cmul_conj_if(cplx<T> a, cplx<T> w, bool c) {
    // a * (c ? conj(w) : w)
    T wy = c ? -w.y : w.y;
    return mk(a.x * w.x - a.y * wy, a.x * wy + a.y * w.x);
}

// ------------------------------------------------- compile-time trigonometry
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_small(double x) {  // |x| <= pi/4, Taylor to 1 ulp
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / double((2 * i) * (2 * i + 1));
        sum += term;
    }
    return sum;
}
constexpr double cx_cos_small(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / double((2 * i - 1) * (2 * i));
        sum += term;
    }
    return sum;
}
// cos/sin of 2*pi*k/L (exact octant reduction on the integer pair (k, L)).
struct cs_pair { double c, s; };
constexpr cs_pair cx_unit(long k, long L) {
    k %= L;
    if (k < 0) k += L;
    long t = 8 * k, m = t / L, rem = t - m * L;   // angle = (m + rem/L) pi/4
    double cb = 0, sb = 0;
    if (m % 2 == 0) {
        double phi = double(rem) * (kPi / (4.0 * double(L)));
        cb = cx_cos_small(phi);
        sb = cx_sin_small(phi);
    } else {
        double phi = double(L - rem) * (kPi / (4.0 * double(L)));
        cb = cx_sin_small(phi);
        sb = cx_cos_small(phi);
        m -= 1;
    }
    switch (m) {
        case 0: return {cb, sb};
        case 2: return {-sb, cb};
        case 4: return {-cb, -sb};
        default: return {sb, -cb};  // m == 6
    }
}

// ---------------------------------------------------------- compile-time loop
template <int B, int E, typename F>
DGFS_DEV void sfor(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        sfor<B + 1, E>(std::forward<F>(f));
    }
}

constexpr int smallest_radix(int L) {
    // prefer radix 4, then 2, 3, 5, 7 (any other prime: direct DFT)
    if (L % 4 == 0) return 4;
    if (L % 2 == 0) return 2;
    if (L % 3 == 0) return 3;
    if (L % 5 == 0) return 5;
    if (L % 7 == 0) return 7;
    return L;
}

// XCD-aware block remap: consecutive logical ids land on one XCD (8 XCDs,
// round-robin dispatch), so neighbouring work-groups share an L2.
DGFS_DEV int xcd_remap(int bid, int nblk) {
    const int per = nblk / 8;
    if (bid < per * 8) return (bid % 8) * per + bid / 8;
    return bid;
}

}  // namespace dgfs

// ---------------------------------------------------------------- C-ABI errors
namespace dgfs {
void set_error(const char* fmt, ...);
}

#define DGFS_HIP_CHECK(expr)                                                   \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            ::dgfs::set_error("%s failed: %s (%s:%d)", #expr,                  \
                              hipGetErrorString(_e), __FILE__, __LINE__);      \
            return 1;                                                          \
        }                                                                      \
    } while (0)

#define DGFS_REQUIRE(cond, ...)                                                \
    do {                                                                       \
        if (!(cond)) {                                                         \
            ::dgfs::set_error(__VA_ARGS__);                                    \
            return 2;                                                          \
        }                                                                      \
    } while (0)
