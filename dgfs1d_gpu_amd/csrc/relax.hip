// Relaxation-type collision terms and the FP64 conservation projection for
// the DGFS 1D3V single-species path (gfx950).
//
// Reference behaviour (jaisw7/dgfs1D_gpu):
//   Shakov model 'shakov-direct-gll'   dgfs1D/std/scattering.py:1636-1833
//     moments, f_S, output              std/kernels/scattering/shakov-direct-gll.mako:15-167
//   conservation projection             no counterpart (the reference's only discrete
//                                       conservation mechanism is the BGK Newton solve,
//                                       std/scattering.py:647-695)
//
// Both operators work on whole velocity slices (n = N^3 points of one
// (mode, element) pair).  Reductions are two-kernel and deterministic: a
// partial kernel writes one FP64 vector per (slice, 256-point block), and the
// apply kernel re-reduces those partials in a fixed order inside every block
// that needs them (no atomics, no host synchronisation).
#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {
namespace {

constexpr int kRB = 256;

// block-wide sum of NV doubles; every thread receives the totals
template <int NV>
DGFS_DEV void block_sum_all(double (&v)[NV], double* sh /* [NV][kRB] */) {
    const int tid = threadIdx.x;
    for (int i = 0; i < NV; ++i) sh[i * kRB + tid] = v[i];
    __syncthreads();
    for (int s = kRB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int i = 0; i < NV; ++i) sh[i * kRB + tid] += sh[i * kRB + tid + s];
        __syncthreads();
    }
    for (int i = 0; i < NV; ++i) v[i] = sh[i * kRB];
    __syncthreads();
}

// sum of part[b][NV], b < nb, in a fixed order; result in every thread
template <int NV>
DGFS_DEV void reduce_partials(const double* __restrict__ part, int nb, double (&v)[NV],
                              double* sh) {
    for (int i = 0; i < NV; ++i) v[i] = 0.0;
    for (int b = threadIdx.x; b < nb; b += kRB)
        for (int i = 0; i < NV; ++i) v[i] += part[size_t(b) * NV + i];
    block_sum_all<NV>(v, sh);
}

// ------------------------------------------------ conservation projection
// Q_c = Q - Phi^T G Phi Q with Phi = {1, c_x, c_y, c_z, |c|^2} (5 x n) and
// G = (Phi Phi^T)^{-1}: the Euclidean-nearest slice to Q whose mass, momentum
// and energy moments vanish.
struct Gram5 { double g[25]; };

template <typename T>
__global__ void __launch_bounds__(kRB)
k_proj_partial(int n, const T* __restrict__ Q, const T* __restrict__ cx,
               const T* __restrict__ cy, const T* __restrict__ cz,
               double* __restrict__ part /*[slice][nb][5]*/) {
    __shared__ double sh[5 * kRB];
    const int nb = gridDim.x, s = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    double acc[5] = {0, 0, 0, 0, 0};
    if (v < n) {
        const double q = Q[size_t(s) * n + v];
        const double x = cx[v], y = cy[v], z = cz[v];
        acc[0] = q;
        acc[1] = q * x;
        acc[2] = q * y;
        acc[3] = q * z;
        acc[4] = q * (x * x + y * y + z * z);
    }
    block_sum_all<5>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 5; ++i) part[(size_t(s) * nb + blockIdx.x) * 5 + i] = acc[i];
}

template <typename T>
__global__ void __launch_bounds__(kRB)
k_proj_apply(int n, Gram5 G, const T* __restrict__ Q, const T* __restrict__ cx,
             const T* __restrict__ cy, const T* __restrict__ cz,
             const double* __restrict__ part, T* __restrict__ out, int accumulate) {
    __shared__ double sh[5 * kRB];
    const int nb = gridDim.x, s = blockIdx.y;
    double m[5];
    reduce_partials<5>(part + size_t(s) * nb * 5, nb, m, sh);
    double lam[5];
    for (int i = 0; i < 5; ++i) {
        double a = 0.0;
        for (int j = 0; j < 5; ++j) a += G.g[i * 5 + j] * m[j];
        lam[i] = a;
    }
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    const double x = cx[v], y = cy[v], z = cz[v];
    const size_t o = size_t(s) * n + v;
    const double qc = double(Q[o]) -
                      (lam[0] + lam[1] * x + lam[2] * y + lam[3] * z +
                       lam[4] * (x * x + y * y + z * z));
    if (accumulate) out[o] = T(double(out[o]) + qc);
    else out[o] = T(qc);
}

// ------------------------------------------------------------ Shakov model
// Raw FP64 moments of f_t per 256-point block:
//   0: sum f          1-3: sum f c_i        4-9: sum f c_i c_j (xx yy zz xy yz zx)
//   10-12: sum f |c|^2 c_i
// The reference's two-pass central moments (mako:31-104) follow from these
// exactly in real arithmetic.
constexpr int kShkMom = 13;

template <typename T>
DGFS_DEV double shk_task_value(const int* __restrict__ nterm, const int* __restrict__ slice,
                               const T* __restrict__ coef, int K, int n, const T* __restrict__ u,
                               int t, int v) {
    const int nt = nterm[2 * t];
    const int* sl = slice + size_t(t) * 2 * K;
    const T* cf = coef + size_t(t) * 2 * K;
    T acc = T(0);
    for (int a = 0; a < nt; ++a) acc += cf[a] * u[size_t(sl[a]) * n + v];
    return double(acc);
}

template <typename T>
__global__ void __launch_bounds__(kRB)
k_shk_partial(int K, int n, const int* __restrict__ nterm, const int* __restrict__ slice,
              const T* __restrict__ coef, const T* __restrict__ u, const T* __restrict__ cx,
              const T* __restrict__ cy, const T* __restrict__ cz,
              double* __restrict__ part /*[task][nb][13]*/) {
    __shared__ double sh[kShkMom * kRB];
    const int nb = gridDim.x, t = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    double a[kShkMom];
    for (int i = 0; i < kShkMom; ++i) a[i] = 0.0;
    if (v < n) {
        const double f = shk_task_value(nterm, slice, coef, K, n, u, t, v);
        const double x = cx[v], y = cy[v], z = cz[v];
        const double c2 = x * x + y * y + z * z;
        a[0] = f;
        a[1] = f * x; a[2] = f * y; a[3] = f * z;
        a[4] = f * x * x; a[5] = f * y * y; a[6] = f * z * z;
        a[7] = f * x * y; a[8] = f * y * z; a[9] = f * z * x;
        a[10] = f * c2 * x; a[11] = f * c2 * y; a[12] = f * c2 * z;
    }
    block_sum_all<kShkMom>(a, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < kShkMom; ++i) part[(size_t(t) * nb + blockIdx.x) * kShkMom + i] = a[i];
}

struct ShkParams {
    double prefac, omega, Pr, cw;
};

// macroscopic state of one task: rho, u (3), T, q (3)  (mako:50-104)
struct ShkState {
    double rho, ux, uy, uz, T, qx, qy, qz;
};

DGFS_DEV ShkState shk_state(const double (&m)[kShkMom], double cw) {
    ShkState s;
    const double m0 = m[0];
    s.ux = m[1] / m0;
    s.uy = m[2] / m0;
    s.uz = m[3] / m0;
    s.rho = m0 * cw;
    const double ux = s.ux, uy = s.uy, uz = s.uz;
    const double u2 = ux * ux + uy * uy + uz * uz;
    const double tr = m[4] + m[5] + m[6];
    const double udm1 = ux * m[1] + uy * m[2] + uz * m[3];
    const double S2 = tr - 2.0 * udm1 + u2 * m0;               // sum f |c-u|^2
    s.T = S2 * cw / (1.5 * s.rho);
    // sum f |c-u|^2 (c-u)_i = m3_i - 2 (M2 u)_i + |u|^2 m1_i - u_i S2
    const double M2u_x = m[4] * ux + m[7] * uy + m[9] * uz;
    const double M2u_y = m[7] * ux + m[5] * uy + m[8] * uz;
    const double M2u_z = m[9] * ux + m[8] * uy + m[6] * uz;
    s.qx = cw * (m[10] - 2.0 * M2u_x + u2 * m[1] - ux * S2);
    s.qy = cw * (m[11] - 2.0 * M2u_y + u2 * m[2] - uy * S2);
    s.qz = cw * (m[12] - 2.0 * M2u_z + u2 * m[3] - uz * S2);
    return s;
}

// Q[gslice[g]] += sum_{t in group g} prefac rho T^{1-omega} (f_S(t) - f_t)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_shk_apply(int K, int n, const int* __restrict__ nterm, const int* __restrict__ slice,
            const T* __restrict__ coef, const int* __restrict__ gstart,
            const int* __restrict__ gslice, const T* __restrict__ u, const T* __restrict__ cx,
            const T* __restrict__ cy, const T* __restrict__ cz,
            const double* __restrict__ part, ShkParams P, T* __restrict__ Q) {
    __shared__ double sh[kShkMom * kRB];
    const int nb = gridDim.x, g = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    const bool in = v < n;
    const double x = in ? double(cx[v]) : 0.0, y = in ? double(cy[v]) : 0.0,
                 z = in ? double(cz[v]) : 0.0;
    double acc = 0.0;
    for (int t = gstart[g]; t < gstart[g + 1]; ++t) {
        double m[kShkMom];
        reduce_partials<kShkMom>(part + size_t(t) * nb * kShkMom, nb, m, sh);
        if (!in) continue;
        const ShkState s = shk_state(m, P.cw);
        const double Cx = x - s.ux, Cy = y - s.uy, Cz = z - s.uz;
        const double C2 = Cx * Cx + Cy * Cy + Cz * Cz;
        const double fe = s.rho / pow(kPi * s.T, 1.5) * exp(-C2 / s.T);
        // (1-Pr)/5 * 4 (q.C)/(rho T^2) (C^2/T - 5/2)   (mako:123-141)
        const double corr = (1.0 - P.Pr) / 5.0 * 4.0 * (s.qx * Cx + s.qy * Cy + s.qz * Cz) /
                            (s.rho * s.T * s.T) * (C2 / s.T - 2.5);
        const double fS = fe * (1.0 + corr);
        const double f = shk_task_value(nterm, slice, coef, K, n, u, t, v);
        acc += P.prefac * (s.rho / pow(s.T, P.omega - 1.0)) * (fS - f);
    }
    if (!in) return;
    const size_t o = size_t(gslice[g]) * n + v;
    Q[o] = T(double(Q[o]) + acc);
}

}  // namespace
}  // namespace dgfs

using namespace dgfs;

extern "C" size_t dgfs_projection_work_bytes(int nslice, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return size_t(nslice) * nb * 5 * sizeof(double);
}

extern "C" int dgfs_project_conservative(int dtype, int nslice, int Nv, const void* cx,
                                         const void* cy, const void* cz, const double* gram_inv,
                                         const void* Q, void* out, int accumulate, double* work,
                                         hipStream_t stream) {
    {
        DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
        DGFS_REQUIRE(nslice >= 0 && Nv > 0, "bad sizes");
        DGFS_REQUIRE(gram_inv && Q && out && work && cx && cy && cz, "null pointer");
        if (nslice == 0) return 0;
        DGFS_REQUIRE(nslice <= 65535, "too many slices");
        Gram5 G;
        for (int i = 0; i < 25; ++i) G.g[i] = gram_inv[i];
        const int nb = (Nv + kRB - 1) / kRB;
        const dim3 grid(nb, nslice);
        if (dtype == DGFS_F64) {
            hipLaunchKernelGGL(k_proj_partial<double>, grid, dim3(kRB), 0, stream, Nv,
                               (const double*)Q, (const double*)cx, (const double*)cy,
                               (const double*)cz, work);
            DGFS_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_proj_apply<double>, grid, dim3(kRB), 0, stream, Nv, G,
                               (const double*)Q, (const double*)cx, (const double*)cy,
                               (const double*)cz, (const double*)work, (double*)out, accumulate);
        } else {
            hipLaunchKernelGGL(k_proj_partial<float>, grid, dim3(kRB), 0, stream, Nv,
                               (const float*)Q, (const float*)cx, (const float*)cy,
                               (const float*)cz, work);
            DGFS_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_proj_apply<float>, grid, dim3(kRB), 0, stream, Nv, G,
                               (const float*)Q, (const float*)cx, (const float*)cy,
                               (const float*)cz, (const double*)work, (float*)out, accumulate);
        }
        DGFS_HIP_CHECK(hipGetLastError());
        return 0;
    }
}

extern "C" size_t dgfs_shakov_work_bytes(int ntask, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return size_t(ntask) * nb * kShkMom * sizeof(double);
}

extern "C" int dgfs_shakov(int dtype, int K, int Ne, int Nv, int ntask, const int* task_nterm,
                           const int* task_slice, const void* task_coef, int ngroup,
                           const int* group_start, const int* group_slice, const void* u,
                           const void* cx, const void* cy, const void* cz, double cw,
                           double prefac, double omega, double Pr, void* Q, double* work,
                           hipStream_t stream) {
    {
        DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
        DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0 && ntask >= 0 && ngroup >= 0, "bad sizes");
        if (ntask == 0 || ngroup == 0) return 0;
        DGFS_REQUIRE(ntask <= 65535 && ngroup <= 65535, "too many tasks");
        DGFS_REQUIRE(task_nterm && task_slice && task_coef && group_start && group_slice && u &&
                         cx && cy && cz && Q && work,
                     "null pointer");
        const int nb = (Nv + kRB - 1) / kRB;
        const ShkParams P{prefac, omega, Pr, cw};
        if (dtype == DGFS_F64) {
            hipLaunchKernelGGL(k_shk_partial<double>, dim3(nb, ntask), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const double*)task_coef,
                               (const double*)u, (const double*)cx, (const double*)cy,
                               (const double*)cz, work);
            DGFS_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_shk_apply<double>, dim3(nb, ngroup), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const double*)task_coef, group_start,
                               group_slice, (const double*)u, (const double*)cx,
                               (const double*)cy, (const double*)cz, (const double*)work, P,
                               (double*)Q);
        } else {
            hipLaunchKernelGGL(k_shk_partial<float>, dim3(nb, ntask), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const float*)task_coef,
                               (const float*)u, (const float*)cx, (const float*)cy,
                               (const float*)cz, work);
            DGFS_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL(k_shk_apply<float>, dim3(nb, ngroup), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const float*)task_coef, group_start,
                               group_slice, (const float*)u, (const float*)cx,
                               (const float*)cy, (const float*)cz, (const double*)work, P,
                               (float*)Q);
        }
        DGFS_HIP_CHECK(hipGetLastError());
        (void)Ne;
        return 0;
    }
}
