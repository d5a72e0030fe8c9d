// Linear minmod slope limiter of the IMEX driver (reference limitLin,
// astd/kernels/limiters.mako:17-76, applied by abstd.py:402-416 after each
// step when [time-integrator] limiter is set).
//
// Per velocity point v and interior element e (the first and last elements
// are copied), with nodal values u[k][e][v] (k = 0 the left and K-1 the right
// node), cell means a_e = uavg[0][e][v], forward/backward mean differences
// d+ = a_{e+1} - a_e, d- = a_e - a_{e-1}, and the linear slope s = ulx[0][e][v]:
//   flag = |a - mm(a - u_0, d+, d-) - u_0| + |a + mm(u_{K-1} - a, d+, d-) - u_{K-1}| > 1e-6
//   ulim[k] = flag ? a + (x_k - x_c) mm(s, d+/h, d-/h) : u[k]
// with mm the three-argument minmod and x_k - x_c the node offsets from the
// cell centre.  One thread per (e, v); ulim may alias u.
#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {
namespace {

DGFS_DEV int sgn(double x) { return x > 0 ? 1 : (x < 0 ? -1 : 0); }

DGFS_DEV double minmod3(double a, double b, double c) {
    const int m = (sgn(a) + sgn(b) + sgn(c)) / 3;
    return (m == 1 || m == -1) ? m * fmin(fmin(fabs(a), fabs(b)), fabs(c)) : 0.0;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_limit_linear(int K, int Ne, int Nv, const double* __restrict__ xrel, const double* __restrict__ h,
               const T* __restrict__ u, const T* __restrict__ uavg, const T* __restrict__ ulx,
               T* __restrict__ ulim) {
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= long(Ne) * Nv) return;
    const int e = int(gid / Nv), v = int(gid - long(e) * Nv);
    const size_t ev = size_t(e) * Nv + v, KNe = size_t(Ne) * Nv;
    if (e == 0 || e == Ne - 1) {
        for (int k = 0; k < K; ++k) ulim[k * KNe + ev] = u[k * KNe + ev];
        return;
    }
    // scalar arithmetic in the dtype, as the reference kernel computes it
    const T a = uavg[ev], ap = uavg[ev + Nv], am = uavg[ev - Nv];
    const T dp = ap - a, dm = a - am;
    const T ul = u[ev], ur = u[(K - 1) * KNe + ev];
    const double fl = fabs(double(a - T(minmod3(double(a - ul), double(dp), double(dm))) - ul)) +
                      fabs(double(a + T(minmod3(double(ur - a), double(dp), double(dm))) - ur));
    if (fl > 1e-6) {
        // the reference's h and node offsets are double literals: the
        // divisions and the reconstruction are evaluated in double
        const double he = h[e];
        const T s = T(minmod3(double(ulx[ev]), double(T(double(dp) / he)),
                              double(T(double(dm) / he))));
        for (int k = 0; k < K; ++k)
            ulim[k * KNe + ev] = T(double(a) + xrel[k * Ne + e] * double(s));
    } else {
        for (int k = 0; k < K; ++k) ulim[k * KNe + ev] = u[k * KNe + ev];
    }
}

// Positivity limiter of the upwind driver (reference limitPos,
// limiters.mako:178-203, applied by abstdu.py:432-442 after limitLin): per
// element, theta = min(1, |(a - 1e-13) / (a - min_k u_k)|) and
// ulim_k = a + theta (u_k - a).  The reference takes the absolute value with
// fabsf, so theta is rounded to float; a NaN quotient leaves theta = 1.
template <typename T>
__global__ void __launch_bounds__(256)
k_limit_positivity(int K, int Ne, int Nv, const T* __restrict__ u, const T* __restrict__ uavg,
                   T* __restrict__ ulim) {
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= long(Ne) * Nv) return;
    const size_t ev = size_t(gid), KNe = size_t(Ne) * Nv;
    double fmn = 1e10;
    for (int k = 0; k < K; ++k) fmn = fmin(fmn, double(u[k * KNe + ev]));
    const double a = double(uavg[ev]);
    const double th = fmin(1.0, double(fabsf(float((a - 1e-13) / (a - fmn)))));
    const T at = uavg[ev];
    for (int k = 0; k < K; ++k) {
        const T uk = u[k * KNe + ev];
        ulim[k * KNe + ev] = T(double(at) + th * double(uk - at));
    }
}

}  // namespace
}  // namespace dgfs

extern "C" int dgfs_limit_positivity(int dtype, int K, int Ne, int Nv, const void* u,
                                     const void* uavg, void* ulim, hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "limit_positivity: bad dtype");
    DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0, "limit_positivity: bad sizes");
    DGFS_REQUIRE(u && uavg && ulim, "limit_positivity: null pointer");
    const long nt = long(Ne) * Nv;
    const dim3 grid((nt + 255) / 256);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(dgfs::k_limit_positivity<double>, grid, dim3(256), 0, stream, K, Ne, Nv,
                           (const double*)u, (const double*)uavg, (double*)ulim);
    else
        hipLaunchKernelGGL(dgfs::k_limit_positivity<float>, grid, dim3(256), 0, stream, K, Ne, Nv,
                           (const float*)u, (const float*)uavg, (float*)ulim);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_limit_linear(int dtype, int K, int Ne, int Nv, const double* xrel,
                                 const double* h, const void* u, const void* uavg,
                                 const void* ulx, void* ulim, hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "limit_linear: bad dtype");
    DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0, "limit_linear: bad sizes");
    DGFS_REQUIRE(xrel && h && u && uavg && ulx && ulim, "limit_linear: null pointer");
    const long nt = long(Ne) * Nv;
    const dim3 grid((nt + 255) / 256);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(dgfs::k_limit_linear<double>, grid, dim3(256), 0, stream, K, Ne, Nv,
                           xrel, h, (const double*)u, (const double*)uavg, (const double*)ulx,
                           (double*)ulim);
    else
        hipLaunchKernelGGL(dgfs::k_limit_linear<float>, grid, dim3(256), 0, stream, K, Ne, Nv,
                           xrel, h, (const float*)u, (const float*)uavg, (const float*)ulx,
                           (float*)ulim);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}
