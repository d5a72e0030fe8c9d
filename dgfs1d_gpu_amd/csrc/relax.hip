// Relaxation-type collision terms and the FP64 conservation projection for
// the DGFS 1D3V single-species path (gfx950).
//
// Reference behaviour (jaisw7/dgfs1D_gpu):
//   Shakov model 'shakov-direct-gll'   dgfs1D/std/scattering.py:1636-1833
//     moments, f_S, output              std/kernels/scattering/shakov-direct-gll.mako:15-167
//   ES-BGK model 'esbgk-direct-gll'    dgfs1D/std/scattering.py:1372-1632
//     moments, tensor, Gaussian, output std/kernels/scattering/esbgk-direct-gll.mako:15-229
//   conservation projection             no counterpart (the reference's only discrete
//                                       conservation mechanism is the BGK Newton solve,
//                                       std/scattering.py:647-695)
//
// Both operators work on whole velocity slices (n = N^3 points of one
// (mode, element) pair).  Reductions are deterministic and three-kernel: a
// partial kernel writes one FP64 vector per (slice, 256-point block), a
// finalize kernel (one work-group per slice) sums those partials in a fixed
// order and turns them into the slice's constants (Lagrange multipliers,
// Shakov state), and the apply kernel reads the constants (no atomics, no
// host synchronisation).
#include <algorithm>
#include <mutex>

#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {
namespace {

constexpr int kRB = 256;
constexpr int kNchMax = 64;    // work-groups per task of the Newton solves, at most

// block-wide sum of NV doubles in a fixed order (wave shuffles, then the four
// wave totals); every thread receives the totals.  sh holds NV * 4 doubles.
template <int NV>
DGFS_DEV void block_sum_all(double (&v)[NV], double* sh) {
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < NV; ++i)
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v[i] += __shfl_xor(v[i], o);
    if (l == 0)
#pragma unroll
        for (int i = 0; i < NV; ++i) sh[i * 4 + w] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = (sh[i * 4] + sh[i * 4 + 1]) + (sh[i * 4 + 2] + sh[i * 4 + 3]);
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
    __shared__ double sh[5 * 4];
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


// mixture: per slice the two species' moments m0, m1 -> residuals of the
// six constraints r = (m0[0], m1[0], M0 m0[1..3] + M1 m1[1..3],
// M0 m0[4] + M1 m1[4]), multipliers l = G6 r, and each species' correction
// coefficients in the basis {1, c, |c|^2}: lam_p = (l[p], M_p l[2..5]).
struct Gram6 { double g[36]; };

__global__ void __launch_bounds__(kRB)
k_proj_final_mix(int nb, Gram6 G, double M0, double M1, const double* __restrict__ part0,
                 const double* __restrict__ part1, double* __restrict__ lam0,
                 double* __restrict__ lam1) {
    __shared__ double sh[5 * 4];
    const int s = blockIdx.x;
    double a[5], b[5];
    reduce_partials<5>(part0 + size_t(s) * nb * 5, nb, a, sh);
    __syncthreads();
    reduce_partials<5>(part1 + size_t(s) * nb * 5, nb, b, sh);
    if (threadIdx.x == 0) {
        const double r[6] = {a[0], b[0], M0 * a[1] + M1 * b[1], M0 * a[2] + M1 * b[2],
                             M0 * a[3] + M1 * b[3], M0 * a[4] + M1 * b[4]};
        double l[6];
        for (int i = 0; i < 6; ++i) {
            double acc = 0.0;
            for (int j = 0; j < 6; ++j) acc += G.g[i * 6 + j] * r[j];
            l[i] = acc;
        }
        double* L0 = lam0 + size_t(s) * 8;
        double* L1 = lam1 + size_t(s) * 8;
        L0[0] = l[0];
        L1[0] = l[1];
        for (int i = 1; i < 5; ++i) {
            L0[i] = M0 * l[i + 1];
            L1[i] = M1 * l[i + 1];
        }
    }
}

// out[s][v] (+)= Q[s][v] - Phi(v)^T L for one point
template <typename T>
DGFS_DEV void proj_apply_point(int n, const T* __restrict__ Q, const T* __restrict__ cx,
                               const T* __restrict__ cy, const T* __restrict__ cz,
                               const double* L, T* __restrict__ out, int accumulate, int s,
                               int v) {
    const double x = cx[v], y = cy[v], z = cz[v];
    const size_t o = size_t(s) * n + v;
    const double qc = double(Q[o]) -
                      (L[0] + L[1] * x + L[2] * y + L[3] * z + L[4] * (x * x + y * y + z * z));
    if (accumulate) out[o] = T(double(out[o]) + qc);
    else out[o] = T(qc);
}

// multipliers lam[s][8] (first 5 used) from memory (the mixture's)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_proj_apply(int n, const T* __restrict__ Q, const T* __restrict__ cx,
             const T* __restrict__ cy, const T* __restrict__ cz,
             const double* __restrict__ lam, T* __restrict__ out, int accumulate) {
    const int s = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    proj_apply_point<T>(n, Q, cx, cy, cz, lam + size_t(s) * 8, out, accumulate, s, v);
}

// lam = G (Phi Q) of slice s formed in every work-group: each sums the
// slice's block partials in the same fixed order (no separate launch for
// the 5 x 5 product)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_proj_apply_fused(int n, const T* __restrict__ Q, const T* __restrict__ cx,
                   const T* __restrict__ cy, const T* __restrict__ cz, int nb, Gram5 G,
                   const double* __restrict__ part, T* __restrict__ out, int accumulate) {
    __shared__ double sh[5 * 4];
    const int s = blockIdx.y;
    double m[5], L[5];
    reduce_partials<5>(part + size_t(s) * nb * 5, nb, m, sh);
    for (int i = 0; i < 5; ++i) {
        double a = 0.0;
        for (int j = 0; j < 5; ++j) a += G.g[i * 5 + j] * m[j];
        L[i] = a;
    }
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    proj_apply_point<T>(n, Q, cx, cy, cz, L, out, accumulate, s, v);
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
    __shared__ double sh[kShkMom * 4];
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

// per-task constants of f_S and the relaxation rate, from the summed raw
// moments: cst[t][16] = {ux, uy, uz, 1/T, rho/(pi T)^1.5,
// (1-Pr)/5 * 4/(rho T^2), qx, qy, qz, prefac rho T^(1-omega),
// sum (c-u)_i (c-u)_j f (xx yy zz xy yz zx)}  (mako:50-141)
constexpr int kShkCst = 16;

// With flags non-null it also zeroes the task's work-group flags
// flags[i][t][0 .. kNchMax), i < nflag, and its status word (read by the
// Newton solves that follow on the stream).
__global__ void __launch_bounds__(kRB)
k_shk_final(int nb, ShkParams P, const double* __restrict__ part, double* __restrict__ cst,
            unsigned* __restrict__ flags = nullptr, int nflag = 0,
            int* __restrict__ status = nullptr) {
    __shared__ double sh[kShkMom * 4];
    const int t = blockIdx.x;
    double m[kShkMom];
    reduce_partials<kShkMom>(part + size_t(t) * nb * kShkMom, nb, m, sh);
    if (flags) {
        for (int i = 0; i < nflag; ++i)
            for (int j = threadIdx.x; j < kNchMax; j += kRB)
                flags[(size_t(i) * gridDim.x + t) * kNchMax + j] = 0u;
        if (threadIdx.x == 0) status[t] = 0;
    }
    if (threadIdx.x != 0) return;
    const ShkState st = shk_state(m, P.cw);
    double* c = cst + size_t(t) * kShkCst;
    c[0] = st.ux;
    c[1] = st.uy;
    c[2] = st.uz;
    c[3] = 1.0 / st.T;
    c[4] = st.rho / pow(kPi * st.T, 1.5);
    c[5] = (1.0 - P.Pr) / 5.0 * 4.0 / (st.rho * st.T * st.T);
    c[6] = st.qx;
    c[7] = st.qy;
    c[8] = st.qz;
    c[9] = P.prefac * (st.rho / pow(st.T, P.omega - 1.0));
    // sum (c-u)_i (c-u)_j f (xx yy zz xy yz zx), for the ES-BGK solve
    c[10] = m[4] - m[0] * st.ux * st.ux;
    c[11] = m[5] - m[0] * st.uy * st.uy;
    c[12] = m[6] - m[0] * st.uz * st.uz;
    c[13] = m[7] - m[0] * st.ux * st.uy;
    c[14] = m[8] - m[0] * st.uy * st.uz;
    c[15] = m[9] - m[0] * st.uz * st.ux;
}

// Q[gslice[g]] += sum_{t in group g} prefac rho T^{1-omega} (f_S(t) - f_t)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_shk_apply(int K, int n, const int* __restrict__ nterm, const int* __restrict__ slice,
            const T* __restrict__ coef, const int* __restrict__ gstart,
            const int* __restrict__ gslice, const T* __restrict__ u, const T* __restrict__ cx,
            const T* __restrict__ cy, const T* __restrict__ cz,
            const double* __restrict__ cst, T* __restrict__ Q) {
    const int g = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    const double x = cx[v], y = cy[v], z = cz[v];
    double acc = 0.0;
    for (int t = gstart[g]; t < gstart[g + 1]; ++t) {
        const double* c = cst + size_t(t) * kShkCst;
        const double Cx = x - c[0], Cy = y - c[1], Cz = z - c[2];
        const double C2 = Cx * Cx + Cy * Cy + Cz * Cz;
        const double C2T = C2 * c[3];
        // f_S = f_M (1 + (1-Pr)/5 * 4 (q.C)/(rho T^2) (C^2/T - 5/2))   (mako:123-141)
        const double fS = c[4] * exp(-C2T) *
                          (1.0 + c[5] * (c[6] * Cx + c[7] * Cy + c[8] * Cz) * (C2T - 2.5));
        const double f = shk_task_value(nterm, slice, coef, K, n, u, t, v);
        acc += c[9] * (fS - f);
    }
    const size_t o = size_t(gslice[g]) * n + v;
    Q[o] = T(double(Q[o]) + acc);
}

// ------------------------------------------------------------ ES-BGK model
// 'esbgk-direct-gll' (std/scattering.py:1372-1632, esbgk-direct-gll.mako):
// target G = rho / ((2 pi)^1.5 sqrt(det L)) exp(-C^T L^{-1} C / 2) with
// L = cw/rho ((1 - 1/Pr) sum f C C^T + (1/Pr) sum f_M C C^T), f_M the node's
// discrete Maxwellian of (rho, u, T); rate prefac rho T^{1-omega}.  The f part
// of L comes from the first pass's raw moments; the f_M part is a second pass
// over the grid (the reference sums the discrete Maxwellian, mako:97-162).
constexpr int kEsMom = 6;     // xx yy zz xy yz zx
constexpr int kEsCst = 16;

// per 256-point block: sum f_M C_i C_j with f_M = rho/(pi T)^1.5 exp(-C^2/T)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_es_partial(int n, const T* __restrict__ cx, const T* __restrict__ cy,
             const T* __restrict__ cz, const double* __restrict__ cst,
             double* __restrict__ part /*[task][nb][6]*/) {
    __shared__ double sh[kEsMom * 4];
    const int nb = gridDim.x, t = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    double a[kEsMom];
    for (int i = 0; i < kEsMom; ++i) a[i] = 0.0;
    if (v < n) {
        const double* c = cst + size_t(t) * kShkCst;
        const double Cx = double(cx[v]) - c[0], Cy = double(cy[v]) - c[1],
                     Cz = double(cz[v]) - c[2];
        const double fm = c[4] * exp(-(Cx * Cx + Cy * Cy + Cz * Cz) * c[3]);
        a[0] = fm * Cx * Cx; a[1] = fm * Cy * Cy; a[2] = fm * Cz * Cz;
        a[3] = fm * Cx * Cy; a[4] = fm * Cy * Cz; a[5] = fm * Cz * Cx;
    }
    block_sum_all<kEsMom>(a, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < kEsMom; ++i) part[(size_t(t) * nb + blockIdx.x) * kEsMom + i] = a[i];
}

// est[t] = {ux, uy, uz, A_xx, A_yy, A_zz, A_xy, A_yz, A_zx (A = L^{-1}),
//           rho/((2 pi)^1.5 sqrt(det L)), prefac rho T^(1-omega)}  (mako:143-229)
__global__ void __launch_bounds__(kRB)
k_es_final(int nb, ShkParams P, const double* __restrict__ part1,
           const double* __restrict__ part2, const double* __restrict__ cst,
           double* __restrict__ est) {
    __shared__ double sh[kShkMom * 4];
    const int t = blockIdx.x;
    double m[kShkMom], b[kEsMom];
    reduce_partials<kShkMom>(part1 + size_t(t) * nb * kShkMom, nb, m, sh);
    __syncthreads();
    reduce_partials<kEsMom>(part2 + size_t(t) * nb * kEsMom, nb, b, sh);
    if (threadIdx.x != 0) return;
    const double* c = cst + size_t(t) * kShkCst;
    const double ux = c[0], uy = c[1], uz = c[2], m0 = m[0];
    const double rho = m0 * P.cw;
    // sum f C_i C_j = M2_ij - u_i m1_j - u_j m1_i + u_i u_j m0
    const double F[kEsMom] = {m[4] - 2.0 * ux * m[1] + ux * ux * m0,
                              m[5] - 2.0 * uy * m[2] + uy * uy * m0,
                              m[6] - 2.0 * uz * m[3] + uz * uz * m0,
                              m[7] - ux * m[2] - uy * m[1] + ux * uy * m0,
                              m[8] - uy * m[3] - uz * m[2] + uy * uz * m0,
                              m[9] - uz * m[1] - ux * m[3] + uz * ux * m0};
    const double wf = 1.0 - 1.0 / P.Pr, wm = 1.0 / P.Pr, sc = P.cw / rho;
    double L[kEsMom];
    for (int i = 0; i < kEsMom; ++i) L[i] = sc * (wf * F[i] + wm * b[i]);
    const double xx = L[0], yy = L[1], zz = L[2], xy = L[3], yz = L[4], zx = L[5];
    const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * zx) +
                       zx * (xy * yz - yy * zx);
    double* e = est + size_t(t) * kEsCst;
    e[0] = ux;
    e[1] = uy;
    e[2] = uz;
    e[3] = (yy * zz - yz * yz) / det;
    e[4] = (xx * zz - zx * zx) / det;
    e[5] = (xx * yy - xy * xy) / det;
    e[6] = (yz * zx - zz * xy) / det;
    e[7] = (xy * zx - xx * yz) / det;
    e[8] = (xy * yz - yy * zx) / det;
    e[9] = rho / (pow(2.0 * kPi, 1.5) * sqrt(det));
    e[10] = c[9];
}

// Q[gslice[g]] += sum_{t in group g} prefac rho T^{1-omega} (G(t) - f_t)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_es_apply(int K, int n, const int* __restrict__ nterm, const int* __restrict__ slice,
           const T* __restrict__ coef, const int* __restrict__ gstart,
           const int* __restrict__ gslice, const T* __restrict__ u, const T* __restrict__ cx,
           const T* __restrict__ cy, const T* __restrict__ cz,
           const double* __restrict__ est, T* __restrict__ Q) {
    const int g = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    const double x = cx[v], y = cy[v], z = cz[v];
    double acc = 0.0;
    for (int t = gstart[g]; t < gstart[g + 1]; ++t) {
        const double* e = est + size_t(t) * kEsCst;
        const double Cx = x - e[0], Cy = y - e[1], Cz = z - e[2];
        const double qf = e[3] * Cx * Cx + e[4] * Cy * Cy + e[5] * Cz * Cz +
                          2.0 * (e[6] * Cx * Cy + e[7] * Cy * Cz + e[8] * Cz * Cx);
        const double G = e[9] * exp(-0.5 * qf);
        const double f = shk_task_value(nterm, slice, coef, K, n, u, t, v);
        acc += e[10] * (G - f);
    }
    const size_t o = size_t(gslice[g]) * n + v;
    Q[o] = T(double(Q[o]) + acc);
}

// ---------------------------------------------- multi-work-group Newton solves
// The Newton solves of 'bgk-gll' and 'esbgk-gll' split each task's velocity
// sums over nch work-groups: grid (nch, ntask).  Each thread keeps up to
// kNewtonPts of its velocity points in registers across the iterations.
// Pass k writes one partial vector per work-group into npart[t][k & 1][b][NV]
// and raises the work-group's flag flags[t][b] to k + 1 (a release store);
// the task's work-groups wait until all nch flags of the task have reached
// k + 1, acquire once, and every one of them adds the task's partials in the
// same order and solves the same small system, so all hold identical iterates
// and leave the loop together.  The flags are zeroed by k_shk_final before
// the launch and only grow inside it.  The partials are double-buffered:
// buffer k & 1 is written again only at pass k + 2, which a work-group
// reaches after the meeting of pass k + 1, and a work-group raises its flag
// to k + 2 only after it has read pass k's partials.
// Progress needs every work-group of a task resident at once: the host sizes
// each launch to the device's co-resident capacity for the kernel and
// launches it plainly (a cooperative launch adds about 17 us per launch on
// the MI355X; a grid within the capacity is resident as soon as the device is
// free of other work, and work-groups that kernels of other streams keep out
// start once those finish).  The wait is bounded by kMeetTicks of the 100 MHz
// real-time clock: a meeting that does not fill in time sets kStMeetTimeout
// and ends the solve, so every wave exits.
constexpr int kNewtonPts = 8;
constexpr unsigned long long kMeetTicks = 100000000ull;   // 1 s

// per-task status bits written by the Newton kernels
constexpr int kStNonFinite = 1;      // a residual or iterate became NaN or infinite
constexpr int kStNotConverged = 2;   // the iteration limit was reached above tolerance
constexpr int kStMeetTimeout = 4;    // a work-group meeting gave up waiting

// Raise this work-group's flag to pass1, then wait until all nch flags of the
// task have reached it: wave 0 polls (one flag per lane, without acquiring:
// an acquire at device scope invalidates the L2 of the XCD each time) and
// acquires once; the other waves follow through the barrier.  false on
// timeout.
DGFS_DEV bool task_meet(unsigned* flags, unsigned pass1, int nch) {
    __shared__ int ok;
    __syncthreads();                      // the partial's stores (wave 0) come first
    if (threadIdx.x < 64) {
        const int l = threadIdx.x;
        if (l == 0)
            __hip_atomic_store(flags + blockIdx.x, pass1, __ATOMIC_RELEASE,
                               __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long start = wall_clock64();
        bool done = false;
        for (;;) {
            bool mine = true;
            for (int j = l; j < nch; j += 64)
                mine = mine && __hip_atomic_load(flags + j, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT) >= pass1;
            if (__all(mine)) {
                done = true;
                break;
            }
            if (wall_clock64() - start > kMeetTicks) break;
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (l == 0) ok = done;
    }
    __syncthreads();
    return ok != 0;
}

// Wave reduce-scatter of NV <= 64 values: lane l returns the sum over the
// wave's 64 lanes of value l % NP (NP = NV rounded up to a power of two;
// zero for l % NP >= NV).  Each step sends half of the remaining values to
// the partner lane, NP - 1 + log2(64 / NP) shuffles in all instead of
// 6 NV for a butterfly per value.
template <int NV>
DGFS_DEV double wave_reduce_scatter(const double (&v)[NV]) {
    constexpr int NP = NV <= 1 ? 1 : NV <= 2 ? 2 : NV <= 4 ? 4 : NV <= 8 ? 8 : NV <= 16 ? 16
                     : NV <= 32 ? 32 : 64;
    static_assert(NV <= 64, "at most 64 values");
    const int l = threadIdx.x & 63;
    double a[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) a[i] = i < NV ? v[i] : 0.0;
#pragma unroll
    for (int h = NP / 2; h >= 1; h >>= 1) {
        const bool up = (l & h) != 0;
#pragma unroll
        for (int j = 0; j < h; ++j) {
            const double send = up ? a[j] : a[j + h];
            const double keep = up ? a[j + h] : a[j];
            a[j] = keep + __shfl_xor(send, h);
        }
    }
#pragma unroll
    for (int h = NP; h < 64; h <<= 1) a[0] += __shfl_xor(a[0], h);
    return a[0];
}

// Sum v over the task's work-groups (pass `pass`); every thread receives the
// task totals, added in a fixed order.  base: the task's npart block
// [2][nch][NV]; flags: the task's kNchMax flags; sh: kRB doubles of LDS; pl:
// kNchMax * NV doubles of LDS (the task's partials, read in one round trip).
template <int NV>
DGFS_DEV bool task_sum(double (&v)[NV], double* sh, double* pl, double* base, int nch,
                       int pass, unsigned* flags) {
    constexpr int NW = kRB / 64;
    constexpr int NL = (kNchMax * NV + kRB - 1) / kRB;   // partial loads per thread
    const double r = wave_reduce_scatter<NV>(v);
    sh[threadIdx.x] = r;                  // [wave][lane]
    __syncthreads();
    double* buf = base + size_t(pass & 1) * nch * NV;
    if (threadIdx.x < NV) {               // wave 0, the wave that raises the flag
        double s = sh[threadIdx.x];
#pragma unroll
        for (int w = 1; w < NW; ++w) s += sh[w * 64 + threadIdx.x];
        buf[size_t(blockIdx.x) * NV + threadIdx.x] = s;
    }
    const bool ok = task_meet(flags, unsigned(pass + 1), nch);
    const int tot = nch * NV;
    double x[NL];
#pragma unroll
    for (int k = 0; k < NL; ++k) {        // every load in flight before the first use
        const int j = threadIdx.x + k * kRB;
        x[k] = j < tot ? buf[j] : 0.0;
    }
#pragma unroll
    for (int k = 0; k < NL; ++k) {
        const int j = threadIdx.x + k * kRB;
        if (j < tot) pl[j] = x[k];
    }
    __syncthreads();
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int b = 0; b < nch; ++b) s += pl[b * NV + threadIdx.x];
        sh[threadIdx.x] = s;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) v[i] = sh[i];
    __syncthreads();
    return ok;
}

// residual max that keeps a NaN (fmax would drop it)
DGFS_DEV double res_max(double res, double r) { return (r > res || r != r) ? r : res; }

// 1/x from the hardware reciprocal and two Newton-Raphson steps (each
// doubles the correct bits): a short dependency chain for the small solves
// below, whose Newton iterations correct any last-bit difference.
DGFS_DEV double rcp_f64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(r, e, r);
}

// v of lane `lane` (a constant) in every lane of the wave
DGFS_DEV double readlane_f64(double v, int lane) {
    const unsigned long long b = __double_as_longlong(v);
    const unsigned lo = __builtin_amdgcn_readlane(unsigned(b), lane);
    const unsigned hi = __builtin_amdgcn_readlane(unsigned(b >> 32), lane);
    return __longlong_as_double((static_cast<long long>(hi) << 32) | lo);
}

// Solves J x = J[:, R] (J is R x (R + 1), the right-hand side in the last
// column) by Gaussian elimination with partial pivoting.  The pivot row is
// found and moved by compare-and-swap against every row below it, so that
// every index is a compile-time constant and J stays in registers; the pivot
// is the first row of largest magnitude, as a scan with a strict comparison
// picks it.  J is destroyed.
template <int R>
DGFS_DEV void solve_pivoted(double (&J)[R][R + 1], double (&x)[R]) {
    double rp[R];                         // reciprocals of the pivots
#pragma unroll
    for (int k = 0; k < R; ++k) {
#pragma unroll
        for (int i = k + 1; i < R; ++i) {
            const bool sw = fabs(J[i][k]) > fabs(J[k][k]);
#pragma unroll
            for (int j = k; j <= R; ++j) {
                const double a = J[k][j], b = J[i][j];
                J[k][j] = sw ? b : a;
                J[i][j] = sw ? a : b;
            }
        }
        rp[k] = rcp_f64(J[k][k]);
#pragma unroll
        for (int i = k + 1; i < R; ++i) {
            const double f = J[i][k] * rp[k];
#pragma unroll
            for (int j = k; j <= R; ++j) J[i][j] -= f * J[k][j];
        }
    }
#pragma unroll
    for (int i = R - 1; i >= 0; --i) {
        double acc = J[i][R];
#pragma unroll
        for (int j = i + 1; j < R; ++j) acc -= J[i][j] * x[j];
        x[i] = acc * rp[i];
    }
}

// Solves A x = b for a symmetric definite A (positive or negative) given by
// its upper triangle packed row by row (a[0] = A00, a[1] = A01, ...), by the
// LDL^T factorisation, which needs no pivoting for a definite matrix.  The
// factor overwrites a (L_ij, i > j, in the place of A_ji; D on the diagonal).
// All indices are compile-time constants.
template <int R>
DGFS_DEV constexpr int packed(int i, int j) { return i * R - i * (i - 1) / 2 + (j - i); }

template <int R>
DGFS_DEV void solve_ldlt(double (&a)[R * (R + 1) / 2], const double (&b)[R], double (&x)[R]) {
    double rd[R];
#pragma unroll
    for (int j = 0; j < R; ++j) {
        double e[R];                                   // e_k = L_jk D_k
        double dj = a[packed<R>(j, j)];
#pragma unroll
        for (int k = 0; k < j; ++k) {
            e[k] = a[packed<R>(k, j)] * a[packed<R>(k, k)];
            dj -= a[packed<R>(k, j)] * e[k];
        }
        a[packed<R>(j, j)] = dj;
        rd[j] = rcp_f64(dj);
#pragma unroll
        for (int i = j + 1; i < R; ++i) {
            double s = a[packed<R>(j, i)];
#pragma unroll
            for (int k = 0; k < j; ++k) s -= a[packed<R>(k, i)] * e[k];
            a[packed<R>(j, i)] = s * rd[j];
        }
    }
    double y[R];
#pragma unroll
    for (int i = 0; i < R; ++i) {
        double s = b[i];
#pragma unroll
        for (int k = 0; k < i; ++k) s -= a[packed<R>(k, i)] * y[k];
        y[i] = s;
    }
#pragma unroll
    for (int i = R - 1; i >= 0; --i) {
        double s = y[i] * rd[i];
#pragma unroll
        for (int k = i + 1; k < R; ++k) s -= a[packed<R>(i, k)] * x[k];
        x[i] = s;
    }
}

// --------------------------------------------------------- conservative BGK
// 'bgk-gll' (std/scattering.py:445-706, bgk-gll.mako:83-229): per task the
// discrete Maxwellian fe = a0 exp(-a1 |C|^2 + a2 Cx + a3 Cy + a4 Cz), C = c - u,
// whose discrete mass, momentum and energy (sums times cw) equal f's, found
// by Newton's method from a = (rho/(pi T)^1.5, 1/T, 0, 0, 0); then
// Q += prefac rho T^{1-omega} (fe - f).  At most kBgkIters iterations, as the
// reference's loop (scattering.py:647-695), which raises on a NaN residual.
// Each pass forms the sums v[i] = sum phi_i fe and v[5 + 5 i + j] =
// sum phi_i d_j, phi = {1, c_x, c_y, c_z, |c|^2}, d = dfe/da =
// fe {1/a0, -|C|^2, C_x, C_y, C_z}; for the ES-BGK model (ES = true) also
// v[30 + k] = sum C_i C_j fe (xx yy zz xy yz zx), which the converged pass
// leaves in est[t][10 + k] for the ES-BGK solve.
constexpr int kBgkIters = 20;
constexpr double kBgkTol = 1e-12;
constexpr int kBgkSums = 30;

template <bool ES>
struct BgkSums { static constexpr int value = ES ? kBgkSums + 6 : kBgkSums; };

// Per-task constants of the BGK solve from cst (k_shk_final's).
struct BgkTask {
    double ux, uy, uz, m[5], sc[5], a[5];
    DGFS_DEV explicit BgkTask(const double* c) {
        ux = c[0];
        uy = c[1];
        uz = c[2];
        const double Tn = 1.0 / c[3];
        const double rho = c[4] * pow(kPi * Tn, 1.5);
        m[0] = rho;
        m[1] = rho * ux;
        m[2] = rho * uy;
        m[3] = rho * uz;
        m[4] = 1.5 * rho * Tn + rho * (ux * ux + uy * uy + uz * uz);
        // residual scales: |rho| (a task's f may enter with a negative
        // coefficient, e.g. the basis' U/V terms, so that rho, a0 and the rate
        // are negative)
        const double ar = fabs(rho);
        sc[0] = ar;
        sc[1] = sc[2] = sc[3] = ar * sqrt(Tn);
        sc[4] = ar * Tn;
        a[0] = c[4];
        a[1] = c[3];
        a[2] = a[3] = a[4] = 0.0;
    }
};

// One Newton update of a from the pass's sums: 1 = converged (a unchanged),
// -1 = non-finite residual, 0 = updated.
template <int NV>
DGFS_DEV int bgk_update(const double (&v)[NV], const BgkTask& k, double cw, double (&a)[5]) {
    double F[5], J[5][6], res = 0.0;
#pragma unroll
    for (int i = 0; i < 5; ++i) {
        F[i] = cw * v[i] - k.m[i];
        res = res_max(res, fabs(F[i]) / k.sc[i]);
#pragma unroll
        for (int j = 0; j < 5; ++j) J[i][j] = cw * v[5 + 5 * i + j];
        J[i][5] = -F[i];
    }
    if (!isfinite(res)) return -1;
    if (res < kBgkTol) return 1;
    double dlt[5];
    solve_pivoted<5>(J, dlt);
#pragma unroll
    for (int i = 0; i < 5; ++i) a[i] += dlt[i];
    return 0;
}

// est[t] = {ux, uy, uz, a0..a4, rate, iterations, (ES) sum C_i C_j fe}
template <int NV>
DGFS_DEV void bgk_finish(const BgkTask& k, const double (&a)[5], const double (&v)[NV], int it,
                         bool conv, int st, const double* c, double* e, int* status) {
    if (!conv && st == 0) st = kStNotConverged;
    for (int i = 0; i < 5; ++i)
        if (!isfinite(a[i])) st |= kStNonFinite;
    e[0] = k.ux;
    e[1] = k.uy;
    e[2] = k.uz;
    for (int i = 0; i < 5; ++i) e[3 + i] = a[i];
    e[8] = c[9];
    e[9] = double(it);
    for (int i = kBgkSums; i < NV; ++i) e[10 + i - kBgkSums] = v[i];
    *status |= st;
}

template <typename T, bool ES>
__global__ void __launch_bounds__(kRB)
k_bgk_newton(int n, int t0, const T* __restrict__ cx, const T* __restrict__ cy,
             const T* __restrict__ cz, double cw, const double* __restrict__ cst,
             double* __restrict__ npart /*[task][2][nch][NV]*/, unsigned* __restrict__ flags,
             double* __restrict__ est /*[task][16]*/, int* __restrict__ status) {
    constexpr int NV = BgkSums<ES>::value;
    __shared__ double sh[kRB];
    __shared__ double pl[kNchMax * NV];
    const int nch = gridDim.x, t = t0 + blockIdx.y;
    const double* c = cst + size_t(t) * kShkCst;
    const BgkTask tk(c);
    const double ux = tk.ux, uy = tk.uy, uz = tk.uz;
    double a[5] = {tk.a[0], tk.a[1], tk.a[2], tk.a[3], tk.a[4]};
    double* base = npart + size_t(t) * 2 * nch * NV;
    unsigned* fl = flags + size_t(t) * kNchMax;
    // this thread's points p0 + k * stride; the first kNewtonPts stay in registers
    const int p0 = blockIdx.x * kRB + threadIdx.x, stride = nch * kRB;
    T px[kNewtonPts], py[kNewtonPts], pz[kNewtonPts];
#pragma unroll
    for (int k = 0; k < kNewtonPts; ++k) {
        const int p = p0 + k * stride;
        px[k] = p < n ? cx[p] : T(0);
        py[k] = p < n ? cy[p] : T(0);
        pz[k] = p < n ? cz[p] : T(0);
    }
    int it = 0, st = 0;
    bool conv = false;
    double v[NV];
    for (; it < kBgkIters; ++it) {
#pragma unroll
        for (int i = 0; i < NV; ++i) v[i] = 0.0;
        auto point = [&](double x, double y, double z) {
            const double Cx = x - ux, Cy = y - uy, Cz = z - uz;
            const double C2 = Cx * Cx + Cy * Cy + Cz * Cz;
            const double fe = a[0] * exp(-a[1] * C2 + a[2] * Cx + a[3] * Cy + a[4] * Cz);
            const double phi[5] = {1.0, x, y, z, x * x + y * y + z * z};
            const double d[5] = {fe / a[0], -fe * C2, fe * Cx, fe * Cy, fe * Cz};
#pragma unroll
            for (int i = 0; i < 5; ++i) {
                v[i] += phi[i] * fe;
#pragma unroll
                for (int j = 0; j < 5; ++j) v[5 + 5 * i + j] += phi[i] * d[j];
            }
            if constexpr (ES) {
                v[30] += fe * Cx * Cx;
                v[31] += fe * Cy * Cy;
                v[32] += fe * Cz * Cz;
                v[33] += fe * Cx * Cy;
                v[34] += fe * Cy * Cz;
                v[35] += fe * Cz * Cx;
            }
        };
#pragma unroll
        for (int k = 0; k < kNewtonPts; ++k)
            if (p0 + k * stride < n) point(px[k], py[k], pz[k]);
        for (int p = p0 + kNewtonPts * stride; p < n; p += stride) point(cx[p], cy[p], cz[p]);
        if (!task_sum<NV>(v, sh, pl, base, nch, it, fl)) {
            st |= kStMeetTimeout;
            break;
        }
        const int r = bgk_update<NV>(v, tk, cw, a);
        if (r < 0) {
            st |= kStNonFinite;
            break;
        }
        if (r > 0) {
            conv = true;
            break;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        bgk_finish<NV>(tk, a, v, it, conv, st, c, est + size_t(t) * kEsCst, status + t);
}

// 'bgk-gll' on a tensor velocity grid, c(p) = (c0[i], c0[j], c0[k]) at
// p = (i N + j) N + k: fe = a0 Ex(c_x) Ey(c_y) Ez(c_z) with
// Ex(x) = exp(-a1 (x - ux)^2 + a2 (x - ux)) (Ey, Ez alike), so the sum over
// the N^3 points of any polynomial in C = c - u times fe is a0 times products
// of the one-dimensional moments X_p = sum_i (c0_i - ux)^p Ex(c0_i), p <= 4
// (and Y_p, Z_p).  One wave per task forms the 15 moments (grid_n <= 64),
// rebuilds the sums of k_bgk_newton from them (phi and d written in C) and
// takes the same Newton steps: no pass over N^3 points and no meeting
// between work-groups.  The sums equal the point sums up to rounding.
constexpr int kTensorMaxN = 64;   // largest grid_n of the tensor kernel

template <typename T, bool ES>
__global__ void __launch_bounds__(64)
k_bgk_newton_tensor(int N, int t0, const T* __restrict__ c0, double cw,
                    const double* __restrict__ cst, double* __restrict__ est,
                    int* __restrict__ status) {
    constexpr int NV = BgkSums<ES>::value;
    __shared__ double sC[3][kTensorMaxN], sE[3][kTensorMaxN];
    const int t = t0 + blockIdx.x, l = threadIdx.x;
    const double* c = cst + size_t(t) * kShkCst;
    const BgkTask tk(c);
    const double ux = tk.ux, uy = tk.uy, uz = tk.uz;
    const double u2 = ux * ux + uy * uy + uz * uz;
    double a[5] = {tk.a[0], tk.a[1], tk.a[2], tk.a[3], tk.a[4]};
    int it = 0, st = 0;
    bool conv = false;
    double v[NV];
    for (; it < kBgkIters; ++it) {
        // per node and axis the offset C and weight E = exp(-a1 C^2 + a_(2+d) C)
        // to LDS; lane 4 q + g (q < 15: moment (d, p) = (q / 5, q % 5);
        // g < 4: nodes g, g + 4, ...) sums C^p E; the quad adds its four
        // parts and lane 4 q hands moment q to every lane
        for (int i = l; i < N; i += 64) {
            const double x = double(c0[i]);
#pragma unroll
            for (int d = 0; d < 3; ++d) {
                const double C = x - (d == 0 ? ux : d == 1 ? uy : uz);
                sC[d][i] = C;
                sE[d][i] = exp(-a[1] * C * C + a[2 + d] * C);
            }
        }
        __syncthreads();
        double part = 0.0;
        {
            const int q = l >> 2, g = l & 3;
            const int d = q < 15 ? q / 5 : 0, pw = q % 5;
            if (q < 15) {
                double s0 = 0.0, s1 = 0.0;
                int i = g;
                for (; i + 4 < N; i += 8) {
                    const double C0 = sC[d][i], C1 = sC[d][i + 4];
                    double w0 = sE[d][i], w1 = sE[d][i + 4];
                    if (pw > 0) { w0 *= C0; w1 *= C1; }
                    if (pw > 1) { w0 *= C0; w1 *= C1; }
                    if (pw > 2) { w0 *= C0; w1 *= C1; }
                    if (pw > 3) { w0 *= C0; w1 *= C1; }
                    s0 += w0;
                    s1 += w1;
                }
                if (i < N) {
                    const double C0 = sC[d][i];
                    double w0 = sE[d][i];
                    if (pw > 0) w0 *= C0;
                    if (pw > 1) w0 *= C0;
                    if (pw > 2) w0 *= C0;
                    if (pw > 3) w0 *= C0;
                    s0 += w0;
                }
                part = s0 + s1;
            }
            part += __shfl_xor(part, 1);
            part += __shfl_xor(part, 2);
        }
        __syncthreads();                    // sC, sE are rewritten by the next pass
        double m[15];                       // X_0..4, Y_0..4, Z_0..4
#pragma unroll
        for (int q = 0; q < 15; ++q) m[q] = readlane_f64(part, 4 * q);
        const double* X = m;
        const double* Y = m + 5;
        const double* Z = m + 10;
        // centred monomial sums sum Cx^a Cy^b Cz^c fe = a0 X_a Y_b Z_c
        auto M = [&](int i, int j, int k) { return a[0] * X[i] * Y[j] * Z[k]; };
        const double n0 = M(0, 0, 0);
        const double nx = M(1, 0, 0), ny = M(0, 1, 0), nz = M(0, 0, 1);
        const double nxx = M(2, 0, 0), nyy = M(0, 2, 0), nzz = M(0, 0, 2);
        const double nxy = M(1, 1, 0), nyz = M(0, 1, 1), nzx = M(1, 0, 1);
        const double c2 = nxx + nyy + nzz;                          // sum C^2 fe
        const double c2x = M(3, 0, 0) + M(1, 2, 0) + M(1, 0, 2);    // sum C^2 Cx fe
        const double c2y = M(2, 1, 0) + M(0, 3, 0) + M(0, 1, 2);
        const double c2z = M(2, 0, 1) + M(0, 2, 1) + M(0, 0, 3);
        const double c4 = M(4, 0, 0) + M(0, 4, 0) + M(0, 0, 4) +
                          2.0 * (M(2, 2, 0) + M(0, 2, 2) + M(2, 0, 2));
        // phi = {1, Cx + ux, Cy + uy, Cz + uz, C^2 + 2 u.C + |u|^2}
        v[0] = n0;
        v[1] = nx + ux * n0;
        v[2] = ny + uy * n0;
        v[3] = nz + uz * n0;
        v[4] = c2 + 2.0 * (ux * nx + uy * ny + uz * nz) + u2 * n0;
        // d_0 = fe / a0
        const double ra0 = rcp_f64(a[0]);
#pragma unroll
        for (int i = 0; i < 5; ++i) v[5 + 5 * i] = v[i] * ra0;
        // d_1 = -C^2 fe
        v[6] = -c2;
        v[11] = -(c2x + ux * c2);
        v[16] = -(c2y + uy * c2);
        v[21] = -(c2z + uz * c2);
        v[26] = -(c4 + 2.0 * (ux * c2x + uy * c2y + uz * c2z) + u2 * c2);
        // d_2 = Cx fe, d_3 = Cy fe, d_4 = Cz fe
        v[7] = nx;
        v[12] = nxx + ux * nx;
        v[17] = nxy + uy * nx;
        v[22] = nzx + uz * nx;
        v[27] = c2x + 2.0 * (ux * nxx + uy * nxy + uz * nzx) + u2 * nx;
        v[8] = ny;
        v[13] = nxy + ux * ny;
        v[18] = nyy + uy * ny;
        v[23] = nyz + uz * ny;
        v[28] = c2y + 2.0 * (ux * nxy + uy * nyy + uz * nyz) + u2 * ny;
        v[9] = nz;
        v[14] = nzx + ux * nz;
        v[19] = nyz + uy * nz;
        v[24] = nzz + uz * nz;
        v[29] = c2z + 2.0 * (ux * nzx + uy * nyz + uz * nzz) + u2 * nz;
        if constexpr (ES) {
            v[30] = nxx;
            v[31] = nyy;
            v[32] = nzz;
            v[33] = nxy;
            v[34] = nyz;
            v[35] = nzx;
        }
        const int r = bgk_update<NV>(v, tk, cw, a);
        if (r < 0) {
            st |= kStNonFinite;
            break;
        }
        if (r > 0) {
            conv = true;
            break;
        }
    }
    if (l == 0) bgk_finish<NV>(tk, a, v, it, conv, st, c, est + size_t(t) * kEsCst, status + t);
}

// Q[gslice[g]] += sum_{t in group g} prefac rho T^{1-omega} (fe(t) - f_t)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_bgk_apply(int K, int n, const int* __restrict__ nterm, const int* __restrict__ slice,
            const T* __restrict__ coef, const int* __restrict__ gstart,
            const int* __restrict__ gslice, const T* __restrict__ u, const T* __restrict__ cx,
            const T* __restrict__ cy, const T* __restrict__ cz,
            const double* __restrict__ est, T* __restrict__ Q) {
    const int g = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    const double x = cx[v], y = cy[v], z = cz[v];
    double acc = 0.0;
    for (int t = gstart[g]; t < gstart[g + 1]; ++t) {
        const double* e = est + size_t(t) * kEsCst;
        const double Cx = x - e[0], Cy = y - e[1], Cz = z - e[2];
        const double fe =
            e[3] * exp(-e[4] * (Cx * Cx + Cy * Cy + Cz * Cz) + e[5] * Cx + e[6] * Cy + e[7] * Cz);
        const double f = shk_task_value(nterm, slice, coef, K, n, u, t, v);
        acc += e[8] * (fe - f);
    }
    const size_t o = size_t(gslice[g]) * n + v;
    Q[o] = T(double(Q[o]) + acc);
}

// ------------------------------------------------------- conservative ES-BGK
// 'esbgk-gll' (std/scattering.py:752-1185, esbgk-gll.mako:218-416): after the
// conservative BGK target fe (k_bgk_newton), the second moments
// T_ij = cw/rho sum C_i C_j ((1 - 1/Pr) f + (1/Pr) fe) set the targets
// {rho, rho u, rho (u_i u_j + T_ij)}; the discrete Gaussian
// G = p0 exp(sum_{l=1..9} p_l phi_l(c)), phi = {c_x, c_y, c_z, c_x^2, c_y^2,
// c_z^2, c_x c_y, c_y c_z, c_z c_x}, with those ten moments is found by
// Newton's method (at most 40 iterations, residual < 1e-12 of the scales).
// The Jacobian cw sum phi_k phi_l G (phi_0 = 1) is symmetric, and its 55
// entries are 35 distinct sums cw sum x^a y^b z^c G (a + b + c <= 4).
// The task's work-groups share the sums as k_bgk_newton's do (pass k:
// Newton iteration k), and every thread solves the 10 x 10 system in
// registers.  T_ij needs no pass of its own: f's part comes from the raw
// moments summed before the solves, fe's from the BGK solve's last pass.
constexpr int kEsgIters = 40;
constexpr int kEsgCst = 16;
constexpr int kEsgPairs = 55;      // (i, j), i <= j, of the ten monomials phi
constexpr int kEsgMono = 35;       // distinct products phi_i phi_j: x^a y^b z^c, a + b + c <= 4
// partial sums per work-group of the ES-BGK path's two solves, at most
constexpr int kEsNewtonSums = BgkSums<true>::value > kEsgMono ? BgkSums<true>::value : kEsgMono;

// The 35 monomials x^a y^b z^c of degree <= 4, by degree, then by a, b
// descending: (0,0,0), (1,0,0), (0,1,0), (0,0,1), (2,0,0), (1,1,0), ...
// Monomial i > 0 is monomial kMonoParent[i] times variable kMonoVar[i]
// (0 = x, 1 = y, 2 = z), its parent listed before it.
constexpr int kMonoParent[kEsgMono] = {0,  0,  0,  0,  1,  2,  3,  2,  3,  3,  4,  5,
                                       6,  7,  8,  9,  7,  8,  9,  9,  10, 11, 12, 13,
                                       14, 15, 16, 17, 18, 19, 16, 17, 18, 19, 19};
constexpr int kMonoVar[kEsgMono] = {0, 0, 1, 2, 0, 0, 0, 1, 1, 2, 0, 0, 0, 0, 0, 0, 1, 1,
                                    1, 2, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2};
// z exponent of each monomial
constexpr int kMonoZ[kEsgMono] = {0, 0, 0, 1, 0, 0, 1, 0, 1, 2, 0, 0, 1, 0, 1, 2, 0, 1,
                                  2, 3, 0, 0, 1, 0, 1, 2, 0, 1, 2, 3, 0, 1, 2, 3, 4};
// monomial of phi_i phi_j for the pairs (i, j), i <= j, row by row, with
// phi = {1, x, y, z, x^2, y^2, z^2, xy, yz, zx}
constexpr int kPairMono[kEsgPairs] = {
    0,  1,  2,  3,  4,  7,  9,  5,  8,  6,  4,  5,  6,  10, 13, 15, 11, 14, 12,
    7,  8,  11, 16, 18, 13, 17, 14, 9,  12, 17, 19, 14, 18, 15, 20, 23, 25, 21,
    24, 22, 30, 32, 26, 31, 27, 34, 28, 33, 29, 23, 27, 24, 32, 28, 25};

DGFS_DEV void esg_phi(double x, double y, double z, double (&ph)[10]) {
    ph[0] = 1.0; ph[1] = x; ph[2] = y; ph[3] = z;
    ph[4] = x * x; ph[5] = y * y; ph[6] = z * z;
    ph[7] = x * y; ph[8] = y * z; ph[9] = z * x;
}

// Per-task targets of the ES-BGK solve: m (the ten moments), their residual
// scales and the start q (the continuous Gaussian with the target's
// covariance Sig = cw S / rho, G = rho / ((2 pi)^1.5 sqrt(det Sig))
// exp(-(c-u)^T Sig^-1 (c-u) / 2) written in the monomials, whose discrete
// moments differ from the target by the quadrature error only; if Sig is not
// positive definite, the Maxwellian of (rho, u, T): the reference's start,
// esbgk-gll.mako:261-270).  S = sum C_i C_j ((1 - 1/Pr) f + (1/Pr) fe): f's
// part from the raw moments (cst[10..15]), fe's from the BGK solve's last
// pass (ebgk[10..15]).
struct EsgTask {
    double m[10], sc[10], q[10];
    DGFS_DEV void init(const double* c, const double* b, double cw, double Pr) {
        const double ux = c[0], uy = c[1], uz = c[2], Tn = 1.0 / c[3];
        const double rho = c[4] * pow(kPi * Tn, 1.5);
        double S[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) S[k] = (1.0 - 1.0 / Pr) * c[10 + k] + (1.0 / Pr) * b[10 + k];
        m[0] = rho;
        m[1] = rho * ux;
        m[2] = rho * uy;
        m[3] = rho * uz;
        m[4] = rho * ux * ux + cw * S[0];
        m[5] = rho * uy * uy + cw * S[1];
        m[6] = rho * uz * uz + cw * S[2];
        m[7] = rho * ux * uy + cw * S[3];
        m[8] = rho * uy * uz + cw * S[4];
        m[9] = rho * uz * ux + cw * S[5];
        const double ar = fabs(rho), sT = sqrt(Tn);
        sc[0] = ar;
        sc[1] = sc[2] = sc[3] = ar * sT;
#pragma unroll
        for (int k = 4; k < 10; ++k) sc[k] = ar * Tn;
        const double sxx = cw * S[0] / rho, syy = cw * S[1] / rho, szz = cw * S[2] / rho;
        const double sxy = cw * S[3] / rho, syz = cw * S[4] / rho, szx = cw * S[5] / rho;
        // adjugate of Sig (symmetric)
        const double Axx = syy * szz - syz * syz, Ayy = sxx * szz - szx * szx;
        const double Azz = sxx * syy - sxy * sxy, Axy = syz * szx - sxy * szz;
        const double Ayz = sxy * szx - sxx * syz, Azx = sxy * syz - syy * szx;
        const double det = sxx * Axx + sxy * Axy + szx * Azx;
        if (sxx > 0.0 && Azz > 0.0 && det > 0.0 && isfinite(det)) {
            const double Pxx = Axx / det, Pyy = Ayy / det, Pzz = Azz / det;
            const double Pxy = Axy / det, Pyz = Ayz / det, Pzx = Azx / det;
            const double Pux = Pxx * ux + Pxy * uy + Pzx * uz;
            const double Puy = Pxy * ux + Pyy * uy + Pyz * uz;
            const double Puz = Pzx * ux + Pyz * uy + Pzz * uz;
            q[0] = rho / (pow(2.0 * kPi, 1.5) * sqrt(det)) *
                   exp(-0.5 * (ux * Pux + uy * Puy + uz * Puz));
            q[1] = Pux;
            q[2] = Puy;
            q[3] = Puz;
            q[4] = -0.5 * Pxx;
            q[5] = -0.5 * Pyy;
            q[6] = -0.5 * Pzz;
            q[7] = -Pxy;
            q[8] = -Pyz;
            q[9] = -Pzx;
        } else {
            q[0] = c[4] * exp(-(ux * ux + uy * uy + uz * uz) / Tn);
            q[1] = 2.0 * ux / Tn;
            q[2] = 2.0 * uy / Tn;
            q[3] = 2.0 * uz / Tn;
            q[4] = q[5] = q[6] = -1.0 / Tn;
            q[7] = q[8] = q[9] = 0.0;
        }
    }
};

// One Newton update of q from the monomial sums v: 1 = converged (q
// unchanged), -1 = non-finite residual, 0 = updated.  J_ij = cw sum phi_i
// phi_j G is the Jacobian in (p0 / q0, p1, .., p9) (d G / d p0 = G / p0),
// symmetric and definite: the step solves J d' = -F and scales d'_0 by q0.
DGFS_DEV int esg_update(const double (&v)[kEsgMono], const EsgTask& k, double cw, double (&q)[10]) {
    double a[kEsgPairs], F[10], d[10], res = 0.0;
    sfor<0, kEsgPairs>([&](auto K) {
        constexpr int i = K;
        a[i] = cw * v[kPairMono[i]];
    });
#pragma unroll
    for (int i = 0; i < 10; ++i) {
        F[i] = a[i] - k.m[i];                             // cw sum phi_i G
        res = res_max(res, fabs(F[i]) / k.sc[i]);
        F[i] = -F[i];
    }
    if (!isfinite(res)) return -1;
    if (res < kBgkTol) return 1;
    solve_ldlt<10>(a, F, d);
    q[0] += q[0] * d[0];
#pragma unroll
    for (int i = 1; i < 10; ++i) q[i] += d[i];
    return 0;
}

DGFS_DEV void esg_finish(const double (&q)[10], int it, bool conv, int st, const double* c,
                         double* e, int* status) {
    if (!conv && st == 0) st = kStNotConverged;
    for (int i = 0; i < 10; ++i)
        if (!isfinite(q[i])) st |= kStNonFinite;
    for (int i = 0; i < 10; ++i) e[i] = q[i];
    e[10] = c[9];
    e[11] = double(it);
    *status |= st;
}

template <typename T>
__global__ void __launch_bounds__(kRB)
k_esg_newton(int n, int t0, const T* __restrict__ cx, const T* __restrict__ cy,
             const T* __restrict__ cz, double cw, double Pr, const double* __restrict__ cst,
             const double* __restrict__ ebgk, double* __restrict__ npart /*[task][2][nch][35]*/,
             unsigned* __restrict__ flags, double* __restrict__ eesg /*[task][16]*/,
             int* __restrict__ status) {
    __shared__ double sh[kRB];
    __shared__ double pl[kNchMax * kEsgMono];
    const int nch = gridDim.x, t = t0 + blockIdx.y;
    const double* c = cst + size_t(t) * kShkCst;
    EsgTask tk;
    tk.init(c, ebgk + size_t(t) * kEsCst, cw, Pr);
    double q[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) q[i] = tk.q[i];
    double* base = npart + size_t(t) * 2 * nch * kEsgMono;
    unsigned* fl = flags + size_t(t) * kNchMax;
    const int p0 = blockIdx.x * kRB + threadIdx.x, stride = nch * kRB;
    T px[kNewtonPts], py[kNewtonPts], pz[kNewtonPts];
#pragma unroll
    for (int k = 0; k < kNewtonPts; ++k) {
        const int p = p0 + k * stride;
        px[k] = p < n ? cx[p] : T(0);
        py[k] = p < n ? cy[p] : T(0);
        pz[k] = p < n ? cz[p] : T(0);
    }
    int it = 0, st = 0;
    bool conv = false;
    for (; it < kEsgIters; ++it) {
        // the sums sum_p x^a y^b z^c G(p) over the 35 monomials: every
        // product phi_i phi_j is one of them
        double v[kEsgMono];
#pragma unroll
        for (int i = 0; i < kEsgMono; ++i) v[i] = 0.0;
        auto point = [&](double x, double y, double z) {
            const double e = q[1] * x + q[2] * y + q[3] * z + q[4] * (x * x) + q[5] * (y * y) +
                             q[6] * (z * z) + q[7] * (x * y) + q[8] * (y * z) + q[9] * (z * x);
            const double xyz[3] = {x, y, z};
            double g[kEsgMono];
            g[0] = q[0] * exp(e);
            sfor<1, kEsgMono>([&](auto I) {
                constexpr int i = I;
                g[i] = g[kMonoParent[i]] * xyz[kMonoVar[i]];
            });
#pragma unroll
            for (int i = 0; i < kEsgMono; ++i) v[i] += g[i];
        };
#pragma unroll
        for (int k = 0; k < kNewtonPts; ++k)
            if (p0 + k * stride < n) point(px[k], py[k], pz[k]);
        for (int p = p0 + kNewtonPts * stride; p < n; p += stride) point(cx[p], cy[p], cz[p]);
        if (!task_sum<kEsgMono>(v, sh, pl, base, nch, it, fl)) {
            st |= kStMeetTimeout;
            break;
        }
        const int r = esg_update(v, tk, cw, q);
        if (r < 0) {
            st |= kStNonFinite;
            break;
        }
        if (r > 0) {
            conv = true;
            break;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0)
        esg_finish(q, it, conv, st, c, eesg + size_t(t) * kEsgCst, status + t);
}

// 'esbgk-gll' on a tensor velocity grid (grid_n <= kTensorMaxN; layout as
// k_bgk_newton_tensor's), one work-group per task and no meeting.  For a
// column (x_i, y_j) the Gaussian is G = q0 Ax(x) Ay(y) exp(q7 x y)
// Az(z) exp(beta z), Ax(x) = exp(q1 x + q4 x^2) (Ay, Az alike),
// beta = q8 y + q9 x.  With the nodes z_k = z0 + k h + dz_k (dz_k the
// rounding of the stored nodes off the uniform ones, zero up to rounding for
// FP64 grids), exp(beta z_k) = exp(beta z0) rho^k (1 + beta dz_k) to
// (beta dz_k)^2 / 2, rho = exp(beta h), so the column's moments
// W_c = sum_k z_k^c G are Horner sums in rho of the node tables
// z_k^c Az(z_k) and z_k^c Az(z_k) dz_k: N^3 fused multiply-adds and N^2 + 3N
// exponentials per pass instead of N^3 exponentials.  Then
// sum x^a y^b z^c G = sum over columns of x^a y^b W_c.
constexpr int kEsgTensorMaxN = 48;   // largest grid_n of the tensor ES-BGK kernel
constexpr int kEsgRedG = 7;          // groups per value in its block sums

template <typename T, int NT, int CB>   // NT threads; CB columns per thread and Horner sweep
__global__ void __launch_bounds__(NT)
k_esg_newton_tensor(int N, int t0, const T* __restrict__ c0, double cw, double Pr,
                    const double* __restrict__ cst, const double* __restrict__ ebgk,
                    double* __restrict__ eesg, int* __restrict__ status) {
    static_assert(kEsgMono * kEsgRedG <= NT, "one thread per (value, group)");
    static_assert(3 * kEsgTensorMaxN <= NT, "one thread per node and axis");
    __shared__ double zc[kEsgTensorMaxN], dz[kEsgTensorMaxN];
    __shared__ double Ax[kEsgTensorMaxN], Ay[kEsgTensorMaxN];
    __shared__ double BD[kEsgTensorMaxN][10];   // z^c Az, then z^c Az dz (c <= 4)
    __shared__ double red[kEsgMono][NT];        // per-thread monomial sums
    __shared__ double part[kEsgMono * kEsgRedG];
    __shared__ EsgTask tk;                      // the task's targets, kept out of registers
    const int t = t0 + blockIdx.x, tid = threadIdx.x;
    const double* c = cst + size_t(t) * kShkCst;
    if (tid == 0) tk.init(c, ebgk + size_t(t) * kEsCst, cw, Pr);
    for (int k = tid; k < N; k += NT) zc[k] = double(c0[k]);
    __syncthreads();
    double q[10];
#pragma unroll
    for (int i = 0; i < 10; ++i) q[i] = tk.q[i];
    const double z0 = zc[0], h = N > 1 ? (zc[N - 1] - zc[0]) / double(N - 1) : 0.0;
    for (int k = tid; k < N; k += NT) dz[k] = zc[k] - (z0 + double(k) * h);
    const int ncol = N * N;
    int it = 0, st = 0;
    bool conv = false;
    for (; it < kEsgIters; ++it) {
        __syncthreads();                  // the previous pass's readers are done
        // node tables, one thread per (axis, node)
        if (tid < 3 * N) {
            const int d = tid / N, k = tid - (tid / N) * N;
            const double x = zc[k];
            const double w = exp(q[1 + d] * x + q[4 + d] * x * x);
            if (d == 0) {
                Ax[k] = w;
            } else if (d == 1) {
                Ay[k] = w;
            } else {
                double wp = w;
#pragma unroll
                for (int p = 0; p < 5; ++p) {
                    BD[k][p] = wp;
                    BD[k][5 + p] = wp * dz[k];
                    wp *= x;
                }
            }
        }
        __syncthreads();
        double v[kEsgMono];
#pragma unroll
        for (int m = 0; m < kEsgMono; ++m) v[m] = 0.0;
        for (int col0 = tid * CB; col0 < ncol; col0 += NT * CB) {
            double beta[CB], rho[CB], H[CB][10];
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                const int col = min(col0 + b, ncol - 1);
                const double x = zc[col / N], y = zc[col % N];
                beta[b] = q[8] * y + q[9] * x;
                rho[b] = exp(beta[b] * h);
#pragma unroll
                for (int r = 0; r < 10; ++r) H[b][r] = 0.0;
            }
#pragma unroll 2
            for (int k = N - 1; k >= 0; --k) {
                double row[10];
#pragma unroll
                for (int r = 0; r < 10; ++r) row[r] = BD[k][r];
#pragma unroll
                for (int b = 0; b < CB; ++b)
#pragma unroll
                    for (int r = 0; r < 10; ++r) H[b][r] = fma(H[b][r], rho[b], row[r]);
            }
            // v[x^a y^b z^c] += x^a y^b W_c, W_c = sum_k z_k^c G the column's moment
#pragma unroll
            for (int b = 0; b < CB; ++b) {
                const int col = col0 + b;
                if (col >= ncol) break;
                const int i = col / N, j = col - (col / N) * N;
                const double x = zc[i], y = zc[j];
                const double pre = q[0] * Ax[i] * Ay[j] * exp(q[7] * x * y + beta[b] * z0);
                double W[5];
#pragma unroll
                for (int p = 0; p < 5; ++p) W[p] = pre * fma(beta[b], H[b][5 + p], H[b][p]);
                const double xyz[3] = {x, y, 1.0};
                double g[kEsgMono];
                sfor<0, kEsgMono>([&](auto I) {
                    constexpr int m = I;
                    constexpr int var = kMonoVar[m];
                    if constexpr (m == 0) {
                        g[0] = 1.0;
                    } else if constexpr (var == 2) {
                        g[m] = g[kMonoParent[m]];            // z^c enters through W_c
                    } else {
                        g[m] = g[kMonoParent[m]] * xyz[var];
                    }
                });
                sfor<0, kEsgMono>([&](auto I) {
                    constexpr int m = I;
                    v[m] = fma(g[m], W[kMonoZ[m]], v[m]);
                });
            }
        }
        // block sums through an LDS transpose, in a fixed order: thread (m, g),
        // m < 35, g < kEsgRedG, adds value m of threads g, g + kEsgRedG, ...;
        // then thread m adds the kEsgRedG parts
#pragma unroll
        for (int m = 0; m < kEsgMono; ++m) red[m][tid] = v[m];
        __syncthreads();
        if (tid < kEsgMono * kEsgRedG) {
            const int m = tid / kEsgRedG, g = tid - m * kEsgRedG;
            double a0 = 0.0, a1 = 0.0;
            int e = g;
            for (; e + kEsgRedG < NT; e += 2 * kEsgRedG) {
                a0 += red[m][e];
                a1 += red[m][e + kEsgRedG];
            }
            if (e < NT) a0 += red[m][e];
            part[tid] = a0 + a1;
        }
        __syncthreads();
        if (tid < kEsgMono) {
            double a0 = 0.0;
#pragma unroll
            for (int g = 0; g < kEsgRedG; ++g) a0 += part[tid * kEsgRedG + g];
            red[tid][0] = a0;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kEsgMono; ++i) v[i] = red[i][0];
        const int r = esg_update(v, tk, cw, q);
        if (r < 0) {
            st |= kStNonFinite;
            break;
        }
        if (r > 0) {
            conv = true;
            break;
        }
    }
    if (tid == 0) esg_finish(q, it, conv, st, c, eesg + size_t(t) * kEsgCst, status + t);
}

// Q[gslice[g]] += sum_{t in group g} prefac rho T^{1-omega} (G(t) - f_t)
template <typename T>
__global__ void __launch_bounds__(kRB)
k_esg_apply(int K, int n, const int* __restrict__ nterm, const int* __restrict__ slice,
            const T* __restrict__ coef, const int* __restrict__ gstart,
            const int* __restrict__ gslice, const T* __restrict__ u, const T* __restrict__ cx,
            const T* __restrict__ cy, const T* __restrict__ cz,
            const double* __restrict__ eesg, T* __restrict__ Q) {
    const int g = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    double ph[10];
    esg_phi(cx[v], cy[v], cz[v], ph);
    double acc = 0.0;
    for (int t = gstart[g]; t < gstart[g + 1]; ++t) {
        const double* e = eesg + size_t(t) * kEsgCst;
        double x = 0.0;
        for (int l = 1; l < 10; ++l) x += e[l] * ph[l];
        const double f = shk_task_value(nterm, slice, coef, K, n, u, t, v);
        acc += e[10] * (e[0] * exp(x) - f);
    }
    const size_t o = size_t(gslice[g]) * n + v;
    Q[o] = T(double(Q[o]) + acc);
}

}  // namespace
}  // namespace dgfs

using namespace dgfs;

extern "C" size_t dgfs_projection_work_bytes(int nslice, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return size_t(nslice) * (nb * 5 + 8) * sizeof(double);
}

namespace {

// Launches of the single-species projection on a caller-provided workspace
// (work: dgfs_projection_work_bytes(nslice, Nv) bytes of device memory).
int project_conservative_on(int dtype, int nslice, int Nv, const void* cx, const void* cy,
                            const void* cz, const double* gram_inv, const void* Q, void* out,
                            int accumulate, double* work, hipStream_t stream) {
    Gram5 G;
    for (int i = 0; i < 25; ++i) G.g[i] = gram_inv[i];
    const int nb = (Nv + kRB - 1) / kRB;
    const dim3 grid(nb, nslice);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_proj_partial<double>, grid, dim3(kRB), 0, stream, Nv,
                           (const double*)Q, (const double*)cx, (const double*)cy,
                           (const double*)cz, work);
    else
        hipLaunchKernelGGL(k_proj_partial<float>, grid, dim3(kRB), 0, stream, Nv,
                           (const float*)Q, (const float*)cx, (const float*)cy,
                           (const float*)cz, work);
    DGFS_HIP_CHECK(hipGetLastError());
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_proj_apply_fused<double>, grid, dim3(kRB), 0, stream, Nv,
                           (const double*)Q, (const double*)cx, (const double*)cy,
                           (const double*)cz, nb, G, (const double*)work, (double*)out,
                           accumulate);
    else
        hipLaunchKernelGGL(k_proj_apply_fused<float>, grid, dim3(kRB), 0, stream, Nv,
                           (const float*)Q, (const float*)cx, (const float*)cy,
                           (const float*)cz, nb, G, (const double*)work, (float*)out,
                           accumulate);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

// Workspace of dgfs_project_conservative_noalloc: one array per GPU, held
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is only a stand-in for the symbol): its device
// address is looked up once per device with hipGetSymbolAddress and kept in
// scratch_dev[].
constexpr size_t kScratchBytes = DGFS_PROJECTION_SCRATCH_BYTES;
__device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];

constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_dev[kMaxDevices] = {};

// Device address of the current device's scratch array (looked up on the
// device's first call).
int scratch_address(double** out) {
    int dev = -1;
    DGFS_HIP_CHECK(hipGetDevice(&dev));
    DGFS_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (!scratch_dev[dev]) {
        void* p = nullptr;
        DGFS_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        DGFS_REQUIRE(p != nullptr, "no device address for the projection scratch array");
        scratch_dev[dev] = static_cast<double*>(p);
    }
    *out = scratch_dev[dev];
    return 0;
}

int project_check(int dtype, int nslice, int Nv, const void* cx, const void* cy, const void* cz,
                  const double* gram_inv, const void* Q, void* out) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(nslice >= 0 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(gram_inv && Q && out && cx && cy && cz, "null pointer");
    DGFS_REQUIRE(nslice <= 65535, "too many slices");
    return 0;
}

}  // namespace

extern "C" int dgfs_project_conservative(int dtype, int nslice, int Nv, const void* cx,
                                         const void* cy, const void* cz, const double* gram_inv,
                                         const void* Q, void* out, int accumulate, double* work,
                                         hipStream_t stream) {
    if (int st = project_check(dtype, nslice, Nv, cx, cy, cz, gram_inv, Q, out)) return st;
    DGFS_REQUIRE(work, "null pointer");
    if (nslice == 0) return 0;
    return project_conservative_on(dtype, nslice, Nv, cx, cy, cz, gram_inv, Q, out, accumulate,
                                   work, stream);
}

extern "C" int dgfs_project_conservative_noalloc(int dtype, int nslice, int Nv, const void* cx,
                                                 const void* cy, const void* cz,
                                                 const double* gram_inv, const void* Q,
                                                 void* out, int accumulate, hipStream_t stream) {
    if (int st = project_check(dtype, nslice, Nv, cx, cy, cz, gram_inv, Q, out)) return st;
    if (nslice == 0) return 0;
    const size_t need = dgfs_projection_work_bytes(nslice, Nv);
    DGFS_REQUIRE(need <= kScratchBytes,
                 "nslice=%d, Nv=%d needs %zu bytes of workspace; the no-allocation projection "
                 "holds %zu (use dgfs_project_conservative with a caller workspace)",
                 nslice, Nv, need, kScratchBytes);
    double* work = nullptr;
    if (int st = scratch_address(&work)) return st;
    return project_conservative_on(dtype, nslice, Nv, cx, cy, cz, gram_inv, Q, out, accumulate,
                                   work, stream);
}

extern "C" size_t dgfs_projection_mixture_work_bytes(int nslice, int Nv) {
    return 2 * dgfs_projection_work_bytes(nslice, Nv);
}

extern "C" int dgfs_project_conservative_mixture(int dtype, int nslice, int Nv, const void* cx,
                                                 const void* cy, const void* cz, double m0,
                                                 double m1, const double* gram_inv,
                                                 const void* Q0, const void* Q1, void* out0,
                                                 void* out1, int accumulate, double* work,
                                                 hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(nslice >= 0 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(m0 > 0 && m1 > 0, "masses must be positive");
    DGFS_REQUIRE(gram_inv && Q0 && Q1 && out0 && out1 && work && cx && cy && cz, "null pointer");
    if (nslice == 0) return 0;
    DGFS_REQUIRE(nslice <= 65535, "too many slices");
    Gram6 G;
    for (int i = 0; i < 36; ++i) G.g[i] = gram_inv[i];
    const int nb = (Nv + kRB - 1) / kRB;
    const dim3 grid(nb, nslice);
    double* part0 = work;
    double* part1 = work + size_t(nslice) * nb * 5;
    double* lam0 = part1 + size_t(nslice) * nb * 5;
    double* lam1 = lam0 + size_t(nslice) * 8;
    const void* Qs[2] = {Q0, Q1};
    double* parts[2] = {part0, part1};
    for (int p = 0; p < 2; ++p) {
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_proj_partial<double>, grid, dim3(kRB), 0, stream, Nv,
                               (const double*)Qs[p], (const double*)cx, (const double*)cy,
                               (const double*)cz, parts[p]);
        else
            hipLaunchKernelGGL(k_proj_partial<float>, grid, dim3(kRB), 0, stream, Nv,
                               (const float*)Qs[p], (const float*)cx, (const float*)cy,
                               (const float*)cz, parts[p]);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    hipLaunchKernelGGL(k_proj_final_mix, dim3(nslice), dim3(kRB), 0, stream, nb, G, m0, m1,
                       (const double*)part0, (const double*)part1, lam0, lam1);
    DGFS_HIP_CHECK(hipGetLastError());
    void* outs[2] = {out0, out1};
    const double* lams[2] = {lam0, lam1};
    for (int p = 0; p < 2; ++p) {
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_proj_apply<double>, grid, dim3(kRB), 0, stream, Nv,
                               (const double*)Qs[p], (const double*)cx, (const double*)cy,
                               (const double*)cz, lams[p], (double*)outs[p], accumulate);
        else
            hipLaunchKernelGGL(k_proj_apply<float>, grid, dim3(kRB), 0, stream, Nv,
                               (const float*)Qs[p], (const float*)cx, (const float*)cy,
                               (const float*)cz, lams[p], (float*)outs[p], accumulate);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

extern "C" size_t dgfs_shakov_work_bytes(int ntask, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return size_t(ntask) * (nb * kShkMom + kShkCst) * sizeof(double);
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
        double* cst = work + size_t(ntask) * nb * kShkMom;
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_shk_partial<double>, dim3(nb, ntask), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const double*)task_coef,
                               (const double*)u, (const double*)cx, (const double*)cy,
                               (const double*)cz, work);
        else
            hipLaunchKernelGGL(k_shk_partial<float>, dim3(nb, ntask), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const float*)task_coef,
                               (const float*)u, (const float*)cx, (const float*)cy,
                               (const float*)cz, work);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_shk_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)work, cst);
        DGFS_HIP_CHECK(hipGetLastError());
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_shk_apply<double>, dim3(nb, ngroup), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const double*)task_coef, group_start,
                               group_slice, (const double*)u, (const double*)cx,
                               (const double*)cy, (const double*)cz, (const double*)cst,
                               (double*)Q);
        else
            hipLaunchKernelGGL(k_shk_apply<float>, dim3(nb, ngroup), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const float*)task_coef, group_start,
                               group_slice, (const float*)u, (const float*)cx,
                               (const float*)cy, (const float*)cz, (const double*)cst,
                               (float*)Q);
        DGFS_HIP_CHECK(hipGetLastError());
        (void)Ne;
        return 0;
    }
}

extern "C" size_t dgfs_esbgk_work_bytes(int ntask, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return size_t(ntask) * (nb * (kShkMom + kEsMom) + kShkCst + kEsCst) * sizeof(double);
}

extern "C" int dgfs_esbgk(int dtype, int K, int Ne, int Nv, int ntask, const int* task_nterm,
                          const int* task_slice, const void* task_coef, int ngroup,
                          const int* group_start, const int* group_slice, const void* u,
                          const void* cx, const void* cy, const void* cz, double cw,
                          double prefac, double omega, double Pr, void* Q, double* work,
                          hipStream_t stream) {
    {
        DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
        DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0 && ntask >= 0 && ngroup >= 0, "bad sizes");
        DGFS_REQUIRE(Pr > 0.0, "esbgk: Pr must be positive");
        if (ntask == 0 || ngroup == 0) return 0;
        DGFS_REQUIRE(ntask <= 65535 && ngroup <= 65535, "too many tasks");
        DGFS_REQUIRE(task_nterm && task_slice && task_coef && group_start && group_slice && u &&
                         cx && cy && cz && Q && work,
                     "null pointer");
        const int nb = (Nv + kRB - 1) / kRB;
        const ShkParams P{prefac, omega, Pr, cw};
        double* part1 = work;
        double* part2 = part1 + size_t(ntask) * nb * kShkMom;
        double* cst = part2 + size_t(ntask) * nb * kEsMom;
        double* est = cst + size_t(ntask) * kShkCst;
        if (dtype == DGFS_F64) {
            hipLaunchKernelGGL(k_shk_partial<double>, dim3(nb, ntask), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const double*)task_coef,
                               (const double*)u, (const double*)cx, (const double*)cy,
                               (const double*)cz, part1);
        } else {
            hipLaunchKernelGGL(k_shk_partial<float>, dim3(nb, ntask), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const float*)task_coef,
                               (const float*)u, (const float*)cx, (const float*)cy,
                               (const float*)cz, part1);
        }
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_shk_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)part1, cst);
        DGFS_HIP_CHECK(hipGetLastError());
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_es_partial<double>, dim3(nb, ntask), dim3(kRB), 0, stream, Nv,
                               (const double*)cx, (const double*)cy, (const double*)cz,
                               (const double*)cst, part2);
        else
            hipLaunchKernelGGL(k_es_partial<float>, dim3(nb, ntask), dim3(kRB), 0, stream, Nv,
                               (const float*)cx, (const float*)cy, (const float*)cz,
                               (const double*)cst, part2);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_es_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)part1, (const double*)part2, (const double*)cst, est);
        DGFS_HIP_CHECK(hipGetLastError());
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_es_apply<double>, dim3(nb, ngroup), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const double*)task_coef, group_start,
                               group_slice, (const double*)u, (const double*)cx,
                               (const double*)cy, (const double*)cz, (const double*)est,
                               (double*)Q);
        else
            hipLaunchKernelGGL(k_es_apply<float>, dim3(nb, ngroup), dim3(kRB), 0, stream, K,
                               Nv, task_nterm, task_slice, (const float*)task_coef, group_start,
                               group_slice, (const float*)u, (const float*)cx,
                               (const float*)cy, (const float*)cz, (const double*)est,
                               (float*)Q);
        DGFS_HIP_CHECK(hipGetLastError());
        (void)Ne;
        return 0;
    }
}

namespace dgfs {
namespace {

// work-groups of kRB threads of `kernel` that can be resident at once on the
// current device (cached per kernel and device)
int coresident_groups(const void* kernel) {
    struct Entry { const void* k; int dev, groups; };
    static Entry cache[32];
    static int used = 0;
    static std::mutex mu;
    std::lock_guard<std::mutex> lock(mu);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    for (int i = 0; i < used; ++i)
        if (cache[i].k == kernel && cache[i].dev == dev) return cache[i].groups;
    int ncu = 0, per = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, kRB, 0) != hipSuccess)
        return 0;
    if (used < 32) cache[used++] = Entry{kernel, dev, per * ncu};
    return per * ncu;
}

// Newton work-group layout: nch work-groups per task (enough to keep every
// velocity point in registers, within kNchMax and the co-resident capacity),
// tasks in launches of `per` so that each launch's grid is resident at once
struct NewtonGrid { int nch, per; };

NewtonGrid newton_grid(const void* kernel, int ntask, int n) {
    const int cap = coresident_groups(kernel);
    if (cap <= 0) return NewtonGrid{0, 0};
    const int want = (n + kNewtonPts * kRB - 1) / (kNewtonPts * kRB);
    const int nch = std::max(1, std::min(std::min(kNchMax, want), cap / ntask));
    return NewtonGrid{nch, std::max(1, cap / nch)};
}

// partials and flags behind the sums of the Newton solves: per task two
// buffers of kNchMax partial vectors, then two sets of kNchMax flags (the
// BGK solve's and the ES-BGK solve's) in kNchMax doubles
size_t newton_work_doubles(int ntask, int nsums) {
    return size_t(ntask) * 2 * kNchMax * nsums + size_t(ntask) * kNchMax;
}

// The BGK solves: with 0 < grid_n <= kTensorMaxN the tensor-grid kernel (one
// wave per task; the nodes c0 are cz[0 .. grid_n), as cz[k] = c0[k] at
// i = j = 0), otherwise the point-sum kernel.
template <typename T, bool ES>
int bgk_newton(int Nv, int grid_n, int ntask, const void* cx, const void* cy, const void* cz,
               double cw, const double* cst, double* npart, unsigned* flags, double* est,
               int* status, hipStream_t stream) {
    if (grid_n > 0 && grid_n <= kTensorMaxN) {
        for (int t0 = 0; t0 < ntask; t0 += 65535) {
            const int nt = std::min(65535, ntask - t0);
            hipLaunchKernelGGL((k_bgk_newton_tensor<T, ES>), dim3(nt), dim3(64), 0, stream,
                               grid_n, t0, (const T*)cz, cw, cst, est, status);
            DGFS_HIP_CHECK(hipGetLastError());
        }
        return 0;
    }
    const void* k = reinterpret_cast<const void*>(&k_bgk_newton<T, ES>);
    const NewtonGrid g = newton_grid(k, ntask, Nv);
    DGFS_REQUIRE(g.nch > 0, "conservative BGK: cannot size the Newton launch");
    for (int t0 = 0; t0 < ntask; t0 += g.per) {
        const int nt = std::min(g.per, ntask - t0);
        hipLaunchKernelGGL((k_bgk_newton<T, ES>), dim3(g.nch, nt), dim3(kRB), 0, stream, Nv, t0,
                           (const T*)cx, (const T*)cy, (const T*)cz, cw, cst, npart, flags, est,
                           status);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

// The ES-BGK solves: the tensor-grid kernel as for the BGK solves, otherwise
// the point-sum kernel.
constexpr int kEsgTensorNT = 256, kEsgTensorCB = 4;
template <typename T>
int esg_newton(int Nv, int grid_n, int ntask, const void* cx, const void* cy, const void* cz,
               double cw, double Pr, const double* cst, const double* ebgk, double* npart,
               unsigned* flags, double* eesg, int* status, hipStream_t stream) {
    if (grid_n > 0 && grid_n <= kEsgTensorMaxN) {
        for (int t0 = 0; t0 < ntask; t0 += 65535) {
            const int nt = std::min(65535, ntask - t0);
            hipLaunchKernelGGL((k_esg_newton_tensor<T, kEsgTensorNT, kEsgTensorCB>), dim3(nt),
                               dim3(kEsgTensorNT), 0, stream, grid_n, t0, (const T*)cz, cw, Pr,
                               cst, ebgk, eesg, status);
            DGFS_HIP_CHECK(hipGetLastError());
        }
        return 0;
    }
    const void* k = reinterpret_cast<const void*>(&k_esg_newton<T>);
    const NewtonGrid g = newton_grid(k, ntask, Nv);
    DGFS_REQUIRE(g.nch > 0, "conservative ES-BGK: cannot size the Newton launch");
    for (int t0 = 0; t0 < ntask; t0 += g.per) {
        const int nt = std::min(g.per, ntask - t0);
        hipLaunchKernelGGL(k_esg_newton<T>, dim3(g.nch, nt), dim3(kRB), 0, stream, Nv, t0,
                           (const T*)cx, (const T*)cy, (const T*)cz, cw, Pr, cst, ebgk, npart,
                           flags, eesg, status);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

template <typename T>
void bgk_apply(int K, int Nv, int nb, int ngroup, const int* nterm, const int* slice,
               const void* coef, const int* gstart, const int* gslice, const void* u,
               const void* cx, const void* cy, const void* cz, const double* est, void* Q,
               hipStream_t stream) {
    hipLaunchKernelGGL(k_bgk_apply<T>, dim3(nb, ngroup), dim3(kRB), 0, stream, K, Nv, nterm,
                       slice, (const T*)coef, gstart, gslice, (const T*)u, (const T*)cx,
                       (const T*)cy, (const T*)cz, est, (T*)Q);
}

template <typename T>
void esg_apply(int K, int Nv, int nb, int ngroup, const int* nterm, const int* slice,
               const void* coef, const int* gstart, const int* gslice, const void* u,
               const void* cx, const void* cy, const void* cz, const double* eesg, void* Q,
               hipStream_t stream) {
    hipLaunchKernelGGL(k_esg_apply<T>, dim3(nb, ngroup), dim3(kRB), 0, stream, K, Nv, nterm,
                       slice, (const T*)coef, gstart, gslice, (const T*)u, (const T*)cx,
                       (const T*)cy, (const T*)cz, eesg, (T*)Q);
}

template <typename T>
void shk_partial(int K, int Nv, int nb, int ntask, const int* nterm, const int* slice,
                 const void* coef, const void* u, const void* cx, const void* cy, const void* cz,
                 double* part, hipStream_t stream) {
    hipLaunchKernelGGL(k_shk_partial<T>, dim3(nb, ntask), dim3(kRB), 0, stream, K, Nv, nterm,
                       slice, (const T*)coef, (const T*)u, (const T*)cx, (const T*)cy,
                       (const T*)cz, part);
}

}  // namespace
}  // namespace dgfs

extern "C" size_t dgfs_bgk_work_bytes(int ntask, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return (size_t(ntask) * (nb * kShkMom + kShkCst + kEsCst) +
            newton_work_doubles(ntask, kBgkSums)) * sizeof(double);
}

extern "C" int dgfs_bgk_conservative(int dtype, int K, int Ne, int Nv, int grid_n, int ntask,
                                     const int* task_nterm, const int* task_slice,
                                     const void* task_coef, int ngroup, const int* group_start,
                                     const int* group_slice, const void* u, const void* cx,
                                     const void* cy, const void* cz, double cw, double prefac,
                                     double omega, void* Q, int* status, double* work,
                                     hipStream_t stream) {
    {
        DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
        DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0 && ntask >= 0 && ngroup >= 0, "bad sizes");
        DGFS_REQUIRE(grid_n == 0 || (grid_n > 0 && long(grid_n) * grid_n * grid_n == Nv),
                     "grid_n must be 0 or the cube root of Nv");
        if (ntask == 0 || ngroup == 0) return 0;
        DGFS_REQUIRE(ntask <= 65535 && ngroup <= 65535, "too many tasks");
        DGFS_REQUIRE(task_nterm && task_slice && task_coef && group_start && group_slice && u &&
                         cx && cy && cz && Q && status && work,
                     "null pointer");
        const int nb = (Nv + kRB - 1) / kRB;
        const ShkParams P{prefac, omega, 1.0, cw};
        double* part = work;
        double* cst = part + size_t(ntask) * nb * kShkMom;
        double* est = cst + size_t(ntask) * kShkCst;
        double* npart = est + size_t(ntask) * kEsCst;
        unsigned* flags = reinterpret_cast<unsigned*>(npart + size_t(ntask) * 2 * kNchMax * kBgkSums);
        const bool f64 = dtype == DGFS_F64;
        (f64 ? shk_partial<double> : shk_partial<float>)(K, Nv, nb, ntask, task_nterm, task_slice,
                                                         task_coef, u, cx, cy, cz, part, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_shk_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)part, cst, flags, 1, status);
        DGFS_HIP_CHECK(hipGetLastError());
        const int rc = (f64 ? bgk_newton<double, false> : bgk_newton<float, false>)(
            Nv, grid_n, ntask, cx, cy, cz, cw, cst, npart, flags, est, status, stream);
        if (rc) return rc;
        (f64 ? bgk_apply<double> : bgk_apply<float>)(K, Nv, nb, ngroup, task_nterm, task_slice,
                                                     task_coef, group_start, group_slice, u, cx,
                                                     cy, cz, est, Q, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        return 0;
    }
}

extern "C" size_t dgfs_esbgk_conservative_work_bytes(int ntask, int Nv) {
    const int nb = (Nv + kRB - 1) / kRB;
    return (size_t(ntask) * (nb * kShkMom + kShkCst + kEsCst + kEsgCst) +
            newton_work_doubles(ntask, kEsNewtonSums)) * sizeof(double);
}

extern "C" int dgfs_esbgk_conservative(int dtype, int K, int Ne, int Nv, int grid_n, int ntask,
                                       const int* task_nterm, const int* task_slice,
                                       const void* task_coef, int ngroup, const int* group_start,
                                       const int* group_slice, const void* u, const void* cx,
                                       const void* cy, const void* cz, double cw, double prefac,
                                       double omega, double Pr, void* Q, int* status,
                                       double* work, hipStream_t stream) {
    {
        DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
        DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0 && ntask >= 0 && ngroup >= 0, "bad sizes");
        DGFS_REQUIRE(Pr > 0.0, "esbgk_conservative: Pr must be positive");
        DGFS_REQUIRE(grid_n == 0 || (grid_n > 0 && long(grid_n) * grid_n * grid_n == Nv),
                     "grid_n must be 0 or the cube root of Nv");
        if (ntask == 0 || ngroup == 0) return 0;
        DGFS_REQUIRE(ntask <= 65535 && ngroup <= 65535, "too many tasks");
        DGFS_REQUIRE(task_nterm && task_slice && task_coef && group_start && group_slice && u &&
                         cx && cy && cz && Q && status && work,
                     "null pointer");
        const int nb = (Nv + kRB - 1) / kRB;
        const ShkParams P{prefac, omega, 1.0, cw};
        double* part = work;
        double* cst = part + size_t(ntask) * nb * kShkMom;
        double* ebgk = cst + size_t(ntask) * kShkCst;
        double* eesg = ebgk + size_t(ntask) * kEsCst;
        double* npart = eesg + size_t(ntask) * kEsgCst;
        unsigned* flags = reinterpret_cast<unsigned*>(
            npart + size_t(ntask) * 2 * kNchMax * kEsNewtonSums);
        const bool f64 = dtype == DGFS_F64;
        (f64 ? shk_partial<double> : shk_partial<float>)(K, Nv, nb, ntask, task_nterm, task_slice,
                                                         task_coef, u, cx, cy, cz, part, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_shk_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)part, cst, flags, 2, status);
        DGFS_HIP_CHECK(hipGetLastError());
        int rc = (f64 ? bgk_newton<double, true> : bgk_newton<float, true>)(
            Nv, grid_n, ntask, cx, cy, cz, cw, cst, npart, flags, ebgk, status, stream);
        if (rc) return rc;
        rc = (f64 ? esg_newton<double> : esg_newton<float>)(
            Nv, grid_n, ntask, cx, cy, cz, cw, Pr, cst, ebgk, npart,
            flags + size_t(ntask) * kNchMax, eesg, status, stream);
        if (rc) return rc;
        (f64 ? esg_apply<double> : esg_apply<float>)(K, Nv, nb, ngroup, task_nterm, task_slice,
                                                     task_coef, group_start, group_slice, u, cx,
                                                     cy, cz, eesg, Q, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        return 0;
    }
}
