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

// multipliers lam = G (Phi Q) of slice s -> lam[s][8] (first 5 used)
__global__ void __launch_bounds__(kRB)
k_proj_final(int nb, Gram5 G, const double* __restrict__ part, double* __restrict__ lam) {
    __shared__ double sh[5 * 4];
    const int s = blockIdx.x;
    double m[5];
    reduce_partials<5>(part + size_t(s) * nb * 5, nb, m, sh);
    if (threadIdx.x < 5) {
        double a = 0.0;
        for (int j = 0; j < 5; ++j) a += G.g[threadIdx.x * 5 + j] * m[j];
        lam[size_t(s) * 8 + threadIdx.x] = a;
    }
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

template <typename T>
__global__ void __launch_bounds__(kRB)
k_proj_apply(int n, const T* __restrict__ Q, const T* __restrict__ cx,
             const T* __restrict__ cy, const T* __restrict__ cz,
             const double* __restrict__ lam, T* __restrict__ out, int accumulate) {
    const int s = blockIdx.y;
    const int v = blockIdx.x * kRB + threadIdx.x;
    if (v >= n) return;
    const double* L = lam + size_t(s) * 8;
    const double x = cx[v], y = cy[v], z = cz[v];
    const size_t o = size_t(s) * n + v;
    const double qc = double(Q[o]) -
                      (L[0] + L[1] * x + L[2] * y + L[3] * z + L[4] * (x * x + y * y + z * z));
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
// (1-Pr)/5 * 4/(rho T^2), qx, qy, qz, prefac rho T^(1-omega)}  (mako:50-141)
constexpr int kShkCst = 16;

// With cnt non-null it also zeroes the task's arrival counters cnt[i][t],
// i < ncnt, and its status word (read by the Newton solves that follow on the stream).
__global__ void __launch_bounds__(kRB)
k_shk_final(int nb, ShkParams P, const double* __restrict__ part, double* __restrict__ cst,
            unsigned* __restrict__ cnt = nullptr, int ncnt = 0, int* __restrict__ status = nullptr) {
    __shared__ double sh[kShkMom * 4];
    const int t = blockIdx.x;
    double m[kShkMom];
    reduce_partials<kShkMom>(part + size_t(t) * nb * kShkMom, nb, m, sh);
    if (threadIdx.x != 0) return;
    if (cnt) {
        for (int i = 0; i < ncnt; ++i) cnt[size_t(i) * gridDim.x + t] = 0u;
        status[t] = 0;
    }
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
// sums over nch work-groups: grid (nch, ntask), launched cooperatively, so
// every work-group of the grid is resident at once.  Pass k writes one
// partial vector per work-group into npart[t][k & 1][ch][NV]; the task's
// work-groups then meet at the task's arrival counter (each adds 1 per pass;
// the counter is zeroed by k_shk_final before the launch and never reset
// inside it), and every one of them adds the task's partials in the same
// order and solves the same small system, so all hold identical iterates and
// leave the loop together.  The partials are double-buffered: buffer k & 1 is
// written again only at pass k + 2, which a work-group reaches after the
// meeting of pass k + 1, and every work-group arrives there only after it has
// read pass k's partials.  The wait is bounded: a meeting that does not fill
// within kMeetSpins polls (possible only if the grid were not resident) sets
// kStMeetTimeout and ends the solve, so every wave exits.
constexpr int kNchMax = 32;
constexpr unsigned kMeetSpins = 1u << 22;

// per-task status bits written by the Newton kernels
constexpr int kStNonFinite = 1;      // a residual or iterate became NaN or infinite
constexpr int kStNotConverged = 2;   // the iteration limit was reached above tolerance
constexpr int kStMeetTimeout = 4;    // a work-group meeting gave up waiting

DGFS_DEV bool task_meet(unsigned* cnt, unsigned target) {
    __shared__ int ok;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();                  // this work-group's partial is visible
        atomicAdd(cnt, 1u);
        unsigned seen = __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        for (unsigned s = 0; seen < target && s < kMeetSpins; ++s) {
            __builtin_amdgcn_s_sleep(1);
            seen = __hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
        }
        ok = seen >= target;
    }
    __syncthreads();
    __threadfence();                      // the other work-groups' partials are visible
    return ok != 0;
}

// Sum v over the task's work-groups (pass `pass`); every thread receives the
// task totals.  base: the task's npart block [2][nch][NV]; tot: NV doubles of LDS.
template <int NV>
DGFS_DEV bool task_sum(double (&v)[NV], double* sh, double* tot, double* base, int nch,
                       int pass, unsigned* cnt) {
    block_sum_all<NV>(v, sh);
    double* buf = base + size_t(pass & 1) * nch * NV;
    if (threadIdx.x == 0)
        for (int i = 0; i < NV; ++i) buf[size_t(blockIdx.x) * NV + i] = v[i];
    const bool ok = task_meet(cnt, unsigned(pass + 1) * unsigned(nch));
    if (threadIdx.x < NV) {
        double s = 0.0;
        for (int b = 0; b < nch; ++b) s += buf[size_t(b) * NV + threadIdx.x];
        tot[threadIdx.x] = s;
    }
    __syncthreads();
    for (int i = 0; i < NV; ++i) v[i] = tot[i];
    __syncthreads();
    return ok;
}

// residual max that keeps a NaN (fmax would drop it)
DGFS_DEV double res_max(double res, double r) { return (r > res || r != r) ? r : res; }

// --------------------------------------------------------- conservative BGK
// 'bgk-gll' (std/scattering.py:445-706, bgk-gll.mako:83-229): per task the
// discrete Maxwellian fe = a0 exp(-a1 |C|^2 + a2 Cx + a3 Cy + a4 Cz), C = c - u,
// whose discrete mass, momentum and energy (sums times cw) equal f's, found
// by Newton's method from a = (rho/(pi T)^1.5, 1/T, 0, 0, 0); then
// Q += prefac rho T^{1-omega} (fe - f).  At most kBgkIters iterations, as the
// reference's loop (scattering.py:647-695), which raises on a NaN residual.
constexpr int kBgkIters = 20;
constexpr double kBgkTol = 1e-12;
constexpr int kBgkSums = 30;

template <typename T>
__global__ void __launch_bounds__(kRB)
k_bgk_newton(int n, int t0, const T* __restrict__ cx, const T* __restrict__ cy,
             const T* __restrict__ cz, double cw, const double* __restrict__ cst,
             double* __restrict__ npart /*[task][2][nch][30]*/, unsigned* __restrict__ cnt,
             double* __restrict__ est /*[task][16]: ux uy uz a0..a4 rate iters*/,
             int* __restrict__ status) {
    __shared__ double sh[kBgkSums * 4];
    __shared__ double tot[kBgkSums];
    const int nch = gridDim.x, ch = blockIdx.x, t = t0 + blockIdx.y;
    const double* c = cst + size_t(t) * kShkCst;
    const double ux = c[0], uy = c[1], uz = c[2], Tn = 1.0 / c[3];
    const double rho = c[4] * pow(kPi * Tn, 1.5);
    const double m[5] = {rho, rho * ux, rho * uy, rho * uz,
                         1.5 * rho * Tn + rho * (ux * ux + uy * uy + uz * uz)};
    // residual scales: |rho| (a task's f may enter with a negative coefficient,
    // e.g. the basis' U/V terms, so that rho, a0 and the rate are negative)
    const double ar = fabs(rho);
    const double sc[5] = {ar, ar * sqrt(Tn), ar * sqrt(Tn), ar * sqrt(Tn), ar * Tn};
    double a[5] = {c[4], c[3], 0.0, 0.0, 0.0};
    double* base = npart + size_t(t) * 2 * nch * kBgkSums;
    int it = 0, st = 0;
    bool conv = false;
    for (; it < kBgkIters; ++it) {
        double v[kBgkSums];
        for (int i = 0; i < kBgkSums; ++i) v[i] = 0.0;
        for (int p = ch * kRB + threadIdx.x; p < n; p += nch * kRB) {
            const double x = cx[p], y = cy[p], z = cz[p];
            const double Cx = x - ux, Cy = y - uy, Cz = z - uz;
            const double C2 = Cx * Cx + Cy * Cy + Cz * Cz;
            const double fe = a[0] * exp(-a[1] * C2 + a[2] * Cx + a[3] * Cy + a[4] * Cz);
            const double phi[5] = {1.0, x, y, z, x * x + y * y + z * z};
            const double d[5] = {fe / a[0], -fe * C2, fe * Cx, fe * Cy, fe * Cz};
            for (int i = 0; i < 5; ++i) {
                v[i] += phi[i] * fe;
                for (int j = 0; j < 5; ++j) v[5 + 5 * i + j] += phi[i] * d[j];
            }
        }
        if (!task_sum<kBgkSums>(v, sh, tot, base, nch, it, cnt + t)) {
            st |= kStMeetTimeout;
            break;
        }
        double F[5], J[5][6], res = 0.0;
        for (int i = 0; i < 5; ++i) {
            F[i] = cw * v[i] - m[i];
            res = res_max(res, fabs(F[i]) / sc[i]);
            for (int j = 0; j < 5; ++j) J[i][j] = cw * v[5 + 5 * i + j];
            J[i][5] = -F[i];
        }
        if (!isfinite(res)) {
            st |= kStNonFinite;
            break;
        }
        if (res < kBgkTol) {
            conv = true;
            break;
        }
        // Gaussian elimination with partial pivoting on [J | -F]
        for (int k = 0; k < 5; ++k) {
            int piv = k;
            for (int i = k + 1; i < 5; ++i)
                if (fabs(J[i][k]) > fabs(J[piv][k])) piv = i;
            if (piv != k)
                for (int j = 0; j < 6; ++j) {
                    const double tmp = J[k][j];
                    J[k][j] = J[piv][j];
                    J[piv][j] = tmp;
                }
            for (int i = k + 1; i < 5; ++i) {
                const double f = J[i][k] / J[k][k];
                for (int j = k; j < 6; ++j) J[i][j] -= f * J[k][j];
            }
        }
        double dlt[5];
        for (int i = 4; i >= 0; --i) {
            double acc = J[i][5];
            for (int j = i + 1; j < 5; ++j) acc -= J[i][j] * dlt[j];
            dlt[i] = acc / J[i][i];
        }
        for (int i = 0; i < 5; ++i) a[i] += dlt[i];
    }
    if (!conv && st == 0) st = kStNotConverged;
    for (int i = 0; i < 5; ++i)
        if (!isfinite(a[i])) st |= kStNonFinite;
    if (ch == 0 && threadIdx.x == 0) {
        double* e = est + size_t(t) * kEsCst;
        e[0] = ux;
        e[1] = uy;
        e[2] = uz;
        for (int i = 0; i < 5; ++i) e[3 + i] = a[i];
        e[8] = c[9];
        e[9] = double(it);
        status[t] |= st;
    }
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
// The Jacobian cw sum phi_k phi_l G (phi_0 = 1) is symmetric: 55 sums.
// The task's work-groups share the sums as k_bgk_newton's do (pass 0: the
// tensor, pass k >= 1: Newton iteration k - 1); in each work-group thread 0
// solves the 10 x 10 system and the others read the result from LDS.
constexpr int kEsgIters = 40;
constexpr int kEsgCst = 16;
constexpr int kEsgSums = 55;

DGFS_DEV void esg_phi(double x, double y, double z, double (&ph)[10]) {
    ph[0] = 1.0; ph[1] = x; ph[2] = y; ph[3] = z;
    ph[4] = x * x; ph[5] = y * y; ph[6] = z * z;
    ph[7] = x * y; ph[8] = y * z; ph[9] = z * x;
}

template <typename T>
__global__ void __launch_bounds__(kRB)
k_esg_newton(int K, int n, int t0, const int* __restrict__ nterm, const int* __restrict__ slice,
             const T* __restrict__ coef, const T* __restrict__ u, const T* __restrict__ cx,
             const T* __restrict__ cy, const T* __restrict__ cz, double cw, double Pr,
             const double* __restrict__ cst, const double* __restrict__ ebgk,
             double* __restrict__ npart /*[task][2][nch][55]*/, unsigned* __restrict__ cnt,
             double* __restrict__ eesg /*[task][16]: p0..p9 rate iters*/,
             int* __restrict__ status) {
    __shared__ double sh[kEsgSums * 4];
    __shared__ double tot[kEsgSums];
    __shared__ double sp[11];
    const int nch = gridDim.x, ch = blockIdx.x, t = t0 + blockIdx.y;
    const double* c = cst + size_t(t) * kShkCst;
    const double* b = ebgk + size_t(t) * kEsCst;
    const double ux = c[0], uy = c[1], uz = c[2], Tn = 1.0 / c[3];
    const double rho = c[4] * pow(kPi * Tn, 1.5);
    double* base = npart + size_t(t) * 2 * nch * kEsgSums;
    int st = 0;
    // pass 0: tensor moments of (1 - 1/Pr) f + (1/Pr) fe
    double S[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int p = ch * kRB + threadIdx.x; p < n; p += nch * kRB) {
        const double Cx = double(cx[p]) - ux, Cy = double(cy[p]) - uy, Cz = double(cz[p]) - uz;
        const double fe = b[3] * exp(-b[4] * (Cx * Cx + Cy * Cy + Cz * Cz) + b[5] * Cx +
                                     b[6] * Cy + b[7] * Cz);
        const double w = (1.0 - 1.0 / Pr) * shk_task_value(nterm, slice, coef, K, n, u, t, p) +
                         (1.0 / Pr) * fe;
        S[0] += w * Cx * Cx; S[1] += w * Cy * Cy; S[2] += w * Cz * Cz;
        S[3] += w * Cx * Cy; S[4] += w * Cy * Cz; S[5] += w * Cz * Cx;
    }
    const bool met = task_sum<6>(S, sh, tot, base, nch, 0, cnt + t);
    const double m[10] = {rho, rho * ux, rho * uy, rho * uz,
                          rho * ux * ux + cw * S[0], rho * uy * uy + cw * S[1],
                          rho * uz * uz + cw * S[2], rho * ux * uy + cw * S[3],
                          rho * uy * uz + cw * S[4], rho * uz * ux + cw * S[5]};
    const double ar = fabs(rho), sT = sqrt(Tn);
    const double sc[10] = {ar, ar * sT, ar * sT, ar * sT, ar * Tn, ar * Tn, ar * Tn,
                           ar * Tn, ar * Tn, ar * Tn};
    // start: the Maxwellian of (rho, u, T) written in the monomials
    double q[10] = {c[4] * exp(-(ux * ux + uy * uy + uz * uz) / Tn), 2.0 * ux / Tn, 2.0 * uy / Tn,
                    2.0 * uz / Tn, -1.0 / Tn, -1.0 / Tn, -1.0 / Tn, 0.0, 0.0, 0.0};
    int it = 0;
    bool conv = false;
    if (!met) st |= kStMeetTimeout;
    for (; met && it < kEsgIters; ++it) {
        double v[kEsgSums];
        for (int i = 0; i < kEsgSums; ++i) v[i] = 0.0;
        for (int p = ch * kRB + threadIdx.x; p < n; p += nch * kRB) {
            double ph[10];
            esg_phi(cx[p], cy[p], cz[p], ph);
            double e = 0.0;
            for (int l = 1; l < 10; ++l) e += q[l] * ph[l];
            const double G = q[0] * exp(e);
            int k = 0;
            for (int i = 0; i < 10; ++i)
                for (int j = i; j < 10; ++j) v[k++] += ph[i] * ph[j] * G;
        }
        if (!task_sum<kEsgSums>(v, sh, tot, base, nch, it + 1, cnt + t)) {
            st |= kStMeetTimeout;
            break;
        }
        if (threadIdx.x == 0) {
            double J[10][11], res = 0.0;
            int k = 0;
            for (int i = 0; i < 10; ++i)
                for (int j = i; j < 10; ++j) {
                    J[i][j] = J[j][i] = cw * v[k];
                    ++k;
                }
            for (int i = 0; i < 10; ++i) {
                const double F = J[i][0] - m[i];          // cw sum phi_i G
                res = res_max(res, fabs(F) / sc[i]);
                J[i][10] = -F;
                J[i][0] /= q[0];                          // d G / d p0 = G / p0
            }
            sp[10] = res;
            if (isfinite(res) && res >= kBgkTol) {
                for (int kk = 0; kk < 10; ++kk) {
                    int piv = kk;
                    for (int i = kk + 1; i < 10; ++i)
                        if (fabs(J[i][kk]) > fabs(J[piv][kk])) piv = i;
                    if (piv != kk)
                        for (int j = 0; j < 11; ++j) {
                            const double tmp = J[kk][j];
                            J[kk][j] = J[piv][j];
                            J[piv][j] = tmp;
                        }
                    for (int i = kk + 1; i < 10; ++i) {
                        const double f = J[i][kk] / J[kk][kk];
                        for (int j = kk; j < 11; ++j) J[i][j] -= f * J[kk][j];
                    }
                }
                double d[10];
                for (int i = 9; i >= 0; --i) {
                    double acc = J[i][10];
                    for (int j = i + 1; j < 10; ++j) acc -= J[i][j] * d[j];
                    d[i] = acc / J[i][i];
                }
                for (int i = 0; i < 10; ++i) sp[i] = q[i] + d[i];
            } else {
                for (int i = 0; i < 10; ++i) sp[i] = q[i];
            }
        }
        __syncthreads();
        const double res = sp[10];
        for (int i = 0; i < 10; ++i) q[i] = sp[i];
        __syncthreads();
        if (!isfinite(res)) {
            st |= kStNonFinite;
            break;
        }
        if (res < kBgkTol) {
            conv = true;
            break;
        }
    }
    if (!conv && st == 0) st = kStNotConverged;
    for (int i = 0; i < 10; ++i)
        if (!isfinite(q[i])) st |= kStNonFinite;
    if (ch == 0 && threadIdx.x == 0) {
        double* e = eesg + size_t(t) * kEsgCst;
        for (int i = 0; i < 10; ++i) e[i] = q[i];
        e[10] = c[9];
        e[11] = double(it);
        status[t] |= st;
    }
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
    double* lam = work + size_t(nslice) * nb * 5;
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_proj_partial<double>, grid, dim3(kRB), 0, stream, Nv,
                           (const double*)Q, (const double*)cx, (const double*)cy,
                           (const double*)cz, work);
    else
        hipLaunchKernelGGL(k_proj_partial<float>, grid, dim3(kRB), 0, stream, Nv,
                           (const float*)Q, (const float*)cx, (const float*)cy,
                           (const float*)cz, work);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_proj_final, dim3(nslice), dim3(kRB), 0, stream, nb, G,
                       (const double*)work, lam);
    DGFS_HIP_CHECK(hipGetLastError());
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_proj_apply<double>, grid, dim3(kRB), 0, stream, Nv,
                           (const double*)Q, (const double*)cx, (const double*)cy,
                           (const double*)cz, (const double*)lam, (double*)out, accumulate);
    else
        hipLaunchKernelGGL(k_proj_apply<float>, grid, dim3(kRB), 0, stream, Nv,
                           (const float*)Q, (const float*)cx, (const float*)cy,
                           (const float*)cz, (const double*)lam, (float*)out, accumulate);
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

template <typename... A>
hipError_t launch_coop(const void* kernel, dim3 grid, hipStream_t st, A... args) {
    void* p[] = {static_cast<void*>(&args)...};
    return hipLaunchCooperativeKernel(kernel, grid, dim3(kRB), p, 0, st);
}

// Newton work-group layout: nch work-groups per task, tasks in launches of
// `per` so that each launch's grid is resident
struct NewtonGrid { int nch, per; };

NewtonGrid newton_grid(const void* kernel, int ntask, int nb) {
    const int cap = coresident_groups(kernel);
    if (cap <= 0) return NewtonGrid{0, 0};
    const int nch = std::max(1, std::min(std::min(kNchMax, nb), cap / ntask));
    return NewtonGrid{nch, std::max(1, cap / nch)};
}

// status and arrival counters behind the sums of the Newton solves
size_t newton_work_doubles(int ntask, int nsums) {
    return size_t(ntask) * 2 * kNchMax * nsums + size_t(ntask);   // + 2 uint counters per task
}

template <typename T>
int bgk_newton(int Nv, int ntask, const void* cx, const void* cy, const void* cz, double cw,
               const double* cst, double* npart, unsigned* cnt, double* est, int* status,
               hipStream_t stream) {
    const void* k = reinterpret_cast<const void*>(&k_bgk_newton<T>);
    const NewtonGrid g = newton_grid(k, ntask, (Nv + kRB - 1) / kRB);
    DGFS_REQUIRE(g.nch > 0, "conservative BGK: cannot size the cooperative Newton launch");
    for (int t0 = 0; t0 < ntask; t0 += g.per) {
        const int nt = std::min(g.per, ntask - t0);
        DGFS_HIP_CHECK(launch_coop(k, dim3(g.nch, nt), stream, Nv, t0, (const T*)cx,
                                   (const T*)cy, (const T*)cz, cw, cst, npart, cnt, est, status));
    }
    return 0;
}

template <typename T>
int esg_newton(int K, int Nv, int ntask, const int* nterm, const int* slice, const void* coef,
               const void* u, const void* cx, const void* cy, const void* cz, double cw,
               double Pr, const double* cst, const double* ebgk, double* npart, unsigned* cnt,
               double* eesg, int* status, hipStream_t stream) {
    const void* k = reinterpret_cast<const void*>(&k_esg_newton<T>);
    const NewtonGrid g = newton_grid(k, ntask, (Nv + kRB - 1) / kRB);
    DGFS_REQUIRE(g.nch > 0, "conservative ES-BGK: cannot size the cooperative Newton launch");
    for (int t0 = 0; t0 < ntask; t0 += g.per) {
        const int nt = std::min(g.per, ntask - t0);
        DGFS_HIP_CHECK(launch_coop(k, dim3(g.nch, nt), stream, K, Nv, t0, nterm, slice,
                                   (const T*)coef, (const T*)u, (const T*)cx, (const T*)cy,
                                   (const T*)cz, cw, Pr, cst, ebgk, npart, cnt, eesg, status));
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

extern "C" int dgfs_bgk_conservative(int dtype, int K, int Ne, int Nv, int ntask,
                                     const int* task_nterm, const int* task_slice,
                                     const void* task_coef, int ngroup, const int* group_start,
                                     const int* group_slice, const void* u, const void* cx,
                                     const void* cy, const void* cz, double cw, double prefac,
                                     double omega, void* Q, int* status, double* work,
                                     hipStream_t stream) {
    {
        DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
        DGFS_REQUIRE(K > 0 && Ne > 0 && Nv > 0 && ntask >= 0 && ngroup >= 0, "bad sizes");
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
        unsigned* cnt = reinterpret_cast<unsigned*>(npart + size_t(ntask) * 2 * kNchMax * kBgkSums);
        const bool f64 = dtype == DGFS_F64;
        (f64 ? shk_partial<double> : shk_partial<float>)(K, Nv, nb, ntask, task_nterm, task_slice,
                                                         task_coef, u, cx, cy, cz, part, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_shk_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)part, cst, cnt, 1, status);
        DGFS_HIP_CHECK(hipGetLastError());
        const int rc = (f64 ? bgk_newton<double> : bgk_newton<float>)(
            Nv, ntask, cx, cy, cz, cw, cst, npart, cnt, est, status, stream);
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
            newton_work_doubles(ntask, kEsgSums)) * sizeof(double);
}

extern "C" int dgfs_esbgk_conservative(int dtype, int K, int Ne, int Nv, int ntask,
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
        unsigned* cnt = reinterpret_cast<unsigned*>(npart + size_t(ntask) * 2 * kNchMax * kEsgSums);
        const bool f64 = dtype == DGFS_F64;
        (f64 ? shk_partial<double> : shk_partial<float>)(K, Nv, nb, ntask, task_nterm, task_slice,
                                                         task_coef, u, cx, cy, cz, part, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_shk_final, dim3(ntask), dim3(kRB), 0, stream, nb, P,
                           (const double*)part, cst, cnt, 2, status);
        DGFS_HIP_CHECK(hipGetLastError());
        int rc = (f64 ? bgk_newton<double> : bgk_newton<float>)(
            Nv, ntask, cx, cy, cz, cw, cst, npart, cnt, ebgk, status, stream);
        if (rc) return rc;
        rc = (f64 ? esg_newton<double> : esg_newton<float>)(
            K, Nv, ntask, task_nterm, task_slice, task_coef, u, cx, cy, cz, cw, Pr, cst, ebgk,
            npart, cnt + ntask, eesg, status, stream);
        if (rc) return rc;
        (f64 ? esg_apply<double> : esg_apply<float>)(K, Nv, nb, ngroup, task_nterm, task_slice,
                                                     task_coef, group_start, group_slice, u, cx,
                                                     cy, cz, eesg, Q, stream);
        DGFS_HIP_CHECK(hipGetLastError());
        return 0;
    }
}
