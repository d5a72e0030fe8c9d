// DG transport, boundary conditions, matrix operators, integrator stages,
// moments and residual for the DGFS 1D3V single-species path (gfx950).
//
// Reference behaviour (jaisw7/dgfs1D_gpu):
//   rhs() transport part          dgfs1D/std/std.py:283-348
//   face / flux / lift kernels    dgfs1D/std/kernels/std.mako:4-190
//   boundary conditions           dgfs1D/std/bc.py:25-324, std/kernels/bcs/*.mako
//   GiMMiK matrix operators       dgfs1D/util.py:20-25 (C = A B, A constant)
//   axnpby                        dgfs1D/axnpby.py:13-43
//   moments                       dgfs1D/std/moments.py:11-70 (host FP64 there)
//   residual                      dgfs1D/std/residual.py:15-29
//
// All reductions accumulate in FP64 with a fixed, deterministic order
// (block tree + ordered sum of block partials).
#include <algorithm>
#include <vector>

#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {

constexpr int kBlk = 256;
constexpr int kMaxK = 16;

// deterministic block reduction of NV doubles; result valid in thread 0
template <int NV>
DGFS_DEV void block_reduce(double (&v)[NV], double* sh /* [NV][kBlk] */) {
    const int tid = threadIdx.x;
    for (int i = 0; i < NV; ++i) sh[i * kBlk + tid] = v[i];
    __syncthreads();
    for (int s = kBlk / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int i = 0; i < NV; ++i) sh[i * kBlk + tid] += sh[i * kBlk + tid + s];
        __syncthreads();
    }
    for (int i = 0; i < NV; ++i) v[i] = sh[i * kBlk];
}

// --------------------------------------------------------- matrix operator
// out[k][j] (+)= sum_m A[k][m] in[m][j], j < ncol   (GiMMiK gimmik_mm)
template <typename T>
__global__ void __launch_bounds__(kBlk)
k_matop(int Kout, int Kin, long ncol, const double* __restrict__ A, const T* __restrict__ in,
        T* __restrict__ out, int accumulate) {
    const long j = long(blockIdx.x) * kBlk + threadIdx.x;
    if (j >= ncol) return;
    T x[kMaxK];
    for (int m = 0; m < Kin; ++m) x[m] = in[m * ncol + j];
    for (int k = 0; k < Kout; ++k) {
        T acc = T(0);
        for (int m = 0; m < Kin; ++m) acc += T(A[k * Kin + m]) * x[m];
        if (accumulate) out[k * ncol + j] += acc;
        else out[k * ncol + j] = acc;
    }
}

// ---------------------------------------------------------------- axnpby
// x0 = a0 x0 + sum_{i=1}^{nv-1} a_i x_i  (axnpby.py:13-43; a0 == 0 does not read x0)
template <typename T>
__global__ void __launch_bounds__(kBlk)
k_axnpby(long n, int nv, T* __restrict__ x0, const T* __restrict__ x1, const T* __restrict__ x2,
         const T* __restrict__ x3, T a0, T a1, T a2, T a3) {
    const long i = long(blockIdx.x) * kBlk + threadIdx.x;
    if (i >= n) return;
    T acc = (a0 == T(0)) ? T(0) : a0 * x0[i];
    if (nv > 1) acc += a1 * x1[i];
    if (nv > 2) acc += a2 * x2[i];
    if (nv > 3) acc += a3 * x3[i];
    x0[i] = acc;
}

// ------------------------------------------- compensated FP32 axnpby
// Double-float (hi + lo) arithmetic from FP32 error-free transformations:
// two_sum (Knuth) and the fma residual of a product.  HIP contracts a*b + c
// into one fma across statements, which would fuse the rounded product into
// the following two_sum and break it; an empty asm statement makes the
// rounded product opaque, so it stays a separate v_mul_f32.
struct df32 { float hi, lo; };

DGFS_DEV df32 two_sum(float a, float b) {
    const float s = __fadd_rn(a, b);
    const float bb = __fsub_rn(s, a);
    const float e = __fadd_rn(__fsub_rn(a, __fsub_rn(s, bb)), __fsub_rn(b, bb));
    return {s, e};
}

// (ahi + alo) * (xh + xl) + acc, |result error| ~ 2^-44 |result|
DGFS_DEV df32 df_fma(float ahi, float alo, float xh, float xl, df32 acc) {
    float ph = __fmul_rn(ahi, xh);
    asm volatile("" : "+v"(ph));
    const float pe = __fmaf_rn(ahi, xh, -ph);                   // exact low part of ahi*xh
    const float pl = __fadd_rn(pe, __fadd_rn(__fmul_rn(ahi, xl), __fmul_rn(alo, xh)));
    const df32 s = two_sum(acc.hi, ph);
    const float lo = __fadd_rn(__fadd_rn(s.lo, pl), acc.lo);
    const float hi = __fadd_rn(s.hi, lo);                       // fast two-sum renormalisation
    return {hi, __fsub_rn(lo, __fsub_rn(hi, s.hi))};
}

struct Coef2 { float hi, lo; };
inline Coef2 split_coef(double a) {
    const float h = float(a);
    return {h, float(a - double(h))};
}

// x0 = a0 x0 + sum a_i x_i with every x_i = hi_i + lo_i (lo_i may be null);
// x0_lo receives the low word (or null: x0 keeps the rounded value).
__global__ void __launch_bounds__(kBlk)
k_axnpby_df(long n, int nv, float* __restrict__ x0, float* __restrict__ x0l,
            const float* __restrict__ x1, const float* __restrict__ x1l,
            const float* __restrict__ x2, const float* __restrict__ x2l,
            const float* __restrict__ x3, const float* __restrict__ x3l, Coef2 a0, Coef2 a1,
            Coef2 a2, Coef2 a3) {
    const long i = long(blockIdx.x) * kBlk + threadIdx.x;
    if (i >= n) return;
    df32 acc{0.f, 0.f};
    if (a0.hi != 0.f) acc = df_fma(a0.hi, a0.lo, x0[i], x0l ? x0l[i] : 0.f, acc);
    if (nv > 1) acc = df_fma(a1.hi, a1.lo, x1[i], x1l ? x1l[i] : 0.f, acc);
    if (nv > 2) acc = df_fma(a2.hi, a2.lo, x2[i], x2l ? x2l[i] : 0.f, acc);
    if (nv > 3) acc = df_fma(a3.hi, a3.lo, x3[i], x3l ? x3l[i] : 0.f, acc);
    x0[i] = acc.hi;
    if (x0l) x0l[i] = acc.lo;
}

// ------------------------------------------------------------- wall BCs
struct WallDev {
    int kind;         // DGFS_BC_*
    double nl;        // outward normal (-1 left, +1 right)
    double ushift;    // x-velocity entering fac = (c_x - ushift) nl
    const void* f0;   // Maxwellian / inlet state (n), kinds DIFFUSE, INLET
    double ux, uy, uz, Tw;   // kind DIFFUSE_EXPR (non-dimensional)
};

template <typename T>
DGFS_DEV T wall_inflow(const WallDev& w, int v, T cx, T cy, T cz) {
    if (w.kind == DGFS_BC_DIFFUSE_EXPR) {
        const T dx = cx - T(w.ux), dy = cy - T(w.uy), dz = cz - T(w.uz);
        return exp(-(dx * dx + dy * dy + dz * dz) / T(w.Tw));   // bcs/dgfs-wall-expr-diffuse.mako:14-18
    }
    return static_cast<const T*>(w.f0)[v];
}

struct AdvDev {
    int K, Ne, n;
    const double* D;      // K x K : ux = D u  (derivative operator as GiMMiK gets it)
    const double* BL;     // K : left-face reconstruction  (Bqf[:, 0])
    const double* BR;     // K : right-face reconstruction (Bqf[:, Nqf-1])
    const double* gLD;    // K
    const double* gRD;    // K
    const double* invjac; // Ne
    double cw;
    const double* Dm;     // K x K or null: upwind split form (D used for c < 0, Dm for c > 0)
    const double* De;     // K x K or null: entropy flux-differencing volume term
};

// logarithmic mean (a - b)/(ln a - ln b), evaluated as the entropy driver's
// aveL (astd/kernels/limiters.mako, entropy section): with xi = |a/(b+1e-20)|,
// f = (xi-1)/(xi+1), u = f^2: (a+b)/2 / (1 + u/3 + u^2/5 + u^3/7) for u < 1e-2,
// else (a+b)/2 / (ln(xi+1e-20) / 2 / (f+1e-20)); in the dtype
template <typename T>
DGFS_DEV T log_mean(T a, T b) {
    const T tiny = T(1e-20);
    const T xi = fabs(a / (b + tiny));
    const T f = (xi - T(1)) / (xi + T(1));
    const T u = f * f;
    const T den = u < T(1e-2) ? (T(1) + u / T(3) + u * u / T(5) + u * u * u / T(7))
                              : log(xi + tiny) / T(2) / (f + tiny);
    return T(0.5) * (a + b) / den;
}

template <typename T>
DGFS_DEV T face_value(const AdvDev& a, const T* u, const double* B, int e, int v) {
    T acc = T(0);
    for (int k = 0; k < a.K; ++k) acc += T(B[k]) * u[(size_t(k) * a.Ne + e) * a.n + v];
    return acc;
}

// partial sums of the diffuse-wall number density (bc.py:67-76, bcs/dgfs-wall-diffuse.mako:24-37)
//   part[blk] = {num_L, den_L, num_R, den_R}
template <typename T>
__global__ void __launch_bounds__(kBlk)
k_bc_partial(AdvDev a, const T* __restrict__ u, const T* __restrict__ cx, const T* __restrict__ cy,
             const T* __restrict__ cz, WallDev wl, WallDev wr, double* __restrict__ part) {
    __shared__ double sh[4 * kBlk];
    const int v = blockIdx.x * kBlk + threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < a.n) {
        const T c = cx[v];
        for (int s = 0; s < 2; ++s) {
            const WallDev& w = s == 0 ? wl : wr;
            if (w.kind != DGFS_BC_DIFFUSE && w.kind != DGFS_BC_DIFFUSE_EXPR) continue;
            const T ul = s == 0 ? face_value(a, u, a.BL, 0, v) : face_value(a, u, a.BR, a.Ne - 1, v);
            const double fac = (double(c) - w.ushift) * w.nl;
            if (fac >= 0.0) acc[2 * s] = fac * double(ul) * a.cw;
            else acc[2 * s + 1] = fac * double(wall_inflow<T>(w, v, c, cy[v], cz[v])) * a.cw;
        }
    }
    block_reduce<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; ++i) part[blockIdx.x * 4 + i] = acc[i];
}

template <typename T>
DGFS_DEV T ghost_value(const WallDev& w, double nw, T ul, int v, T cx, T cy, T cz) {
    const double fac = (double(cx) - w.ushift) * w.nl;
    switch (w.kind) {
        case DGFS_BC_COPY: return ul;                                            // periodic/cyclic
        case DGFS_BC_DIFFUSE:
        case DGFS_BC_DIFFUSE_EXPR:
            return fac < 0.0 ? T(double(wall_inflow<T>(w, v, cx, cy, cz)) * nw) : ul;
        case DGFS_BC_INLET: return fac < 0.0 ? wall_inflow<T>(w, v, cx, cy, cz) : ul;
        case DGFS_BC_VANISH: return fac >= 0.0 ? T(0) : ul;                     // dgfs-vanish.mako:14
        default: return ul;
    }
}

// Fused transport (std.py:283-348): for element e and velocity v
//   ux[l][e] = invjac[e] * ( -c (D u)[l] - jL[e+1] gRD[l] + jR[e] gLD[l] )
template <typename T>
DGFS_DEV void advect_point(const AdvDev& a, const T* __restrict__ u, const T* __restrict__ cx,
                           const T* __restrict__ cy, const T* __restrict__ cz, const WallDev& wl,
                           const WallDev& wr, double nw0, double nw1, T* __restrict__ ux, int e,
                           int v) {
    const int K = a.K, Ne = a.Ne, n = a.n;
    const T c = cx[v];
    T uk[kMaxK];
    for (int k = 0; k < K; ++k) uk[k] = u[(size_t(k) * Ne + e) * n + v];
    // face e: left state uL[e], right state uR[e] (std.py:62-66 face maps)
    T uLe, uRe, uLp, uRp;
    {
        T bl = T(0), br = T(0);
        for (int k = 0; k < K; ++k) { bl += T(a.BL[k]) * uk[k]; br += T(a.BR[k]) * uk[k]; }
        uRe = bl;   // element e's left-face value is the right state of face e
        uLp = br;   // element e's right-face value is the left state of face e+1
    }
    if (e > 0) uLe = face_value(a, u, a.BR, e - 1, v);
    else uLe = ghost_value<T>(wl, nw0, uRe, v, c, cy[v], cz[v]);
    if (e < Ne - 1) uRp = face_value(a, u, a.BL, e + 1, v);
    else uRp = ghost_value<T>(wr, nw1, uLp, v, c, cy[v], cz[v]);
    const T ac = fabs(c);
    const T ij = T(a.invjac[e]);
    if (a.De != nullptr) {
        // entropy-stable form (abstde2.py:214-217, eDeriv): volume term
        // -2 sum_m De[l][m] c L(u_l, u_m), standard upwind flux and lift
        const T jR_e = T(0.5) * ac * (uLe - uRe) + T(0.5) * c * (uLe - uRe);
        const T jL_p = T(0.5) * ac * (uLp - uRp) + T(0.5) * c * (uRp - uLp);
        for (int l = 0; l < K; ++l) {
            T r = T(0);
            for (int m = 0; m < K; ++m) r += T(-2.0 * a.De[l * K + m]) * c * log_mean(uk[l], uk[m]);
            r += -jL_p * T(a.gRD[l]) + jR_e * T(a.gLD[l]);
            ux[(size_t(l) * Ne + e) * n + v] = r * ij;
        }
        return;
    }
    if (a.Dm == nullptr) {
        // flux kernel (std.mako:115-152)
        const T jR_e = T(0.5) * ac * (uLe - uRe) + T(0.5) * c * (uLe - uRe);
        const T jL_p = T(0.5) * ac * (uLp - uRp) + T(0.5) * c * (uRp - uLp);
        for (int l = 0; l < K; ++l) {
            T d = T(0);
            for (int m = 0; m < K; ++m) d += T(a.D[l * K + m]) * uk[m];
            T r = -c * d;                                           // mul_by_adv
            r += -jL_p * T(a.gRD[l]) + jR_e * T(a.gLD[l]);          // totalFlux
            ux[(size_t(l) * Ne + e) * n + v] = r * ij;              // mul_by_invjac
        }
        return;
    }
    // upwind summation-by-parts form (abstdu.py:360-378, upwind.mako): flux
    // splitting f = c+ u, g = c- u, ux = -(Dm f + Dp g), face jumps
    // jL = c- (uR - uL), jR = -c+ (uR - uL)
    const T cp = c > T(0) ? c : T(0), cm = c < T(0) ? c : T(0);
    const T jR_e = (uRe - uLe) * -cp;
    const T jL_p = (uRp - uLp) * cm;
    const double* Dsel = c > T(0) ? a.Dm : a.D;
    const T cs = c > T(0) ? cp : cm;
    for (int l = 0; l < K; ++l) {
        T d = T(0);
        for (int m = 0; m < K; ++m) d += T(Dsel[l * K + m]) * (cs * uk[m]);
        T r = -d;
        r += -jL_p * T(a.gRD[l]) + jR_e * T(a.gLD[l]);              // totalFlux
        ux[(size_t(l) * Ne + e) * n + v] = r * ij;
    }
}

// nbw > 0 (a diffuse wall): every work-group forms the wall densities
// nw[s] = -num_s / den_s from the nbw block partials of k_bc_partial (an
// ordered sum, so the same values in every work-group), and work-group 0
// stores them in nw
template <typename T>
__global__ void __launch_bounds__(kBlk)
k_advect(AdvDev a, const T* __restrict__ u, const T* __restrict__ cx, const T* __restrict__ cy,
         const T* __restrict__ cz, WallDev wl, WallDev wr, int nbw,
         const double* __restrict__ part, double* __restrict__ nw, T* __restrict__ ux) {
    double n0 = 0.0, n1 = 0.0;
    if (nbw > 0) {
        // the first wave sums the partials (a fixed shuffle tree) and hands
        // the two densities to the work-group through LDS
        __shared__ double nws[2];
        if (threadIdx.x < 64) {
            double acc[4] = {0.0, 0.0, 0.0, 0.0};
            for (int b = threadIdx.x; b < nbw; b += 64)
                for (int i = 0; i < 4; ++i) acc[i] += part[b * 4 + i];
            for (int o = 32; o > 0; o >>= 1)
                for (int i = 0; i < 4; ++i) acc[i] += __shfl_xor(acc[i], o);
            if (threadIdx.x == 0) {
                nws[0] = acc[1] != 0.0 ? -acc[0] / acc[1] : 0.0;
                nws[1] = acc[3] != 0.0 ? -acc[2] / acc[3] : 0.0;
                if (blockIdx.x == 0) {
                    nw[0] = nws[0];
                    nw[1] = nws[1];
                }
            }
        }
        __syncthreads();
        n0 = nws[0];
        n1 = nws[1];
    }
    const long gid = long(blockIdx.x) * kBlk + threadIdx.x;
    if (gid >= long(a.Ne) * a.n) return;
    const int e = int(gid / a.n);
    const int v = int(gid - long(e) * a.n);
    advect_point<T>(a, u, cx, cy, cz, wl, wr, n0, n1, ux, e, v);
}

// ------------------------------------------------------- batched transport
// Independent problems of one dtype (no reference counterpart): problem p
// reads u + off and writes ux + off; its walls' parameters are fixed at
// batch creation.  Three launches for the batch: the diffuse-wall partial
// sums (grid: blocks x problems), the wall densities (one block per
// problem), and the transport over the concatenated (problem, element,
// velocity) range.
struct ProbDev {
    AdvDev a;
    WallDev wl, wr;
    const void* cx;
    const void* cy;
    const void* cz;
    long off;       // offset of the problem's (K, Ne, Nv) state, in elements
    long start;     // first index of the problem in the concatenated Ne*Nv range
    int nb;         // wall partial blocks (0: no diffuse wall)
};

template <typename T>
__global__ void __launch_bounds__(kBlk)
k_bc_partial_batch(const ProbDev* __restrict__ P, int nbmax, const T* __restrict__ u,
                   double* __restrict__ part) {
    __shared__ double sh[4 * kBlk];
    const ProbDev& q = P[blockIdx.y];
    if (int(blockIdx.x) >= q.nb) return;              // uniform across the block
    const AdvDev& a = q.a;
    const T* up = u + q.off;
    const T* cx = static_cast<const T*>(q.cx);
    const int v = blockIdx.x * kBlk + threadIdx.x;
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    if (v < a.n) {
        const T c = cx[v];
        for (int s = 0; s < 2; ++s) {
            const WallDev& w = s == 0 ? q.wl : q.wr;
            if (w.kind != DGFS_BC_DIFFUSE && w.kind != DGFS_BC_DIFFUSE_EXPR) continue;
            const T ul = s == 0 ? face_value(a, up, a.BL, 0, v) : face_value(a, up, a.BR, a.Ne - 1, v);
            const double fac = (double(c) - w.ushift) * w.nl;
            if (fac >= 0.0) acc[2 * s] = fac * double(ul) * a.cw;
            else acc[2 * s + 1] = fac * double(wall_inflow<T>(w, v, c, static_cast<const T*>(q.cy)[v],
                                                               static_cast<const T*>(q.cz)[v])) * a.cw;
        }
    }
    block_reduce<4>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 4; ++i) part[(size_t(blockIdx.y) * nbmax + blockIdx.x) * 4 + i] = acc[i];
}

__global__ void __launch_bounds__(kBlk)
k_bc_final_batch(const ProbDev* __restrict__ P, int nbmax, const double* __restrict__ part,
                 double* __restrict__ nw) {
    const int p = blockIdx.x;
    if (P[p].nb == 0) return;                          // uniform across the block
    __shared__ double sh[4 * kBlk];
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double* pp = part + size_t(p) * nbmax * 4;
    for (int b = threadIdx.x; b < P[p].nb; b += kBlk)
        for (int i = 0; i < 4; ++i) acc[i] += pp[b * 4 + i];
    block_reduce<4>(acc, sh);
    if (threadIdx.x == 0) {
        nw[2 * p] = acc[1] != 0.0 ? -acc[0] / acc[1] : 0.0;
        nw[2 * p + 1] = acc[3] != 0.0 ? -acc[2] / acc[3] : 0.0;
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlk)
k_advect_batch(const ProbDev* __restrict__ P, int nprob, long total, const T* __restrict__ u,
               const double* __restrict__ nw, T* __restrict__ ux) {
    const long gid = long(blockIdx.x) * kBlk + threadIdx.x;
    if (gid >= total) return;
    int lo = 0, hi = nprob - 1;                        // last problem with start <= gid
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (P[mid].start <= gid) lo = mid;
        else hi = mid - 1;
    }
    const ProbDev& q = P[lo];
    const long loc = gid - q.start;
    const int e = int(loc / q.a.n);
    const int v = int(loc - long(e) * q.a.n);
    advect_point<T>(q.a, u + q.off, static_cast<const T*>(q.cx), static_cast<const T*>(q.cy),
                    static_cast<const T*>(q.cz), q.wl, q.wr, nw[2 * lo], nw[2 * lo + 1],
                    ux + q.off, e, v);
}

// ---------------------------------------------------------------- moments
// Two-pass FP64 moments at the reconstruction points (moments.py:11-70).
// Node = (r, e): s(v) = sum_m im[m][r] u[m][e][v].
// pass 1 partials: {sum s, sum s cx, sum s cy}
// pass 2 partials: {sum s C^2, sum s C^2 Cx, sum s C^2 Cy, sum s Cx Cx,
//                   sum s Cy Cy, sum s Cx Cy},  C = (cx-ux, cy-uy, cz)
template <typename T>
__global__ void __launch_bounds__(kBlk)
k_mom_partial(int pass, int K, int Ne, int Nqr, int n, const double* __restrict__ im,
              const T* __restrict__ u, const T* __restrict__ cx, const T* __restrict__ cy,
              const T* __restrict__ cz, const double* __restrict__ bulk /*[node][3]*/,
              double* __restrict__ part /*[node][nblk][6]*/) {
    __shared__ double sh[6 * kBlk];
    const int nblk = gridDim.x;
    const int node = blockIdx.y;        // node = r*Ne + e
    const int r = node / Ne, e = node - (node / Ne) * Ne;
    const int v = blockIdx.x * kBlk + threadIdx.x;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    if (v < n) {
        double s = 0.0;
        for (int m = 0; m < K; ++m) s += im[m * Nqr + r] * double(u[(size_t(m) * Ne + e) * n + v]);
        const double x = cx[v], y = cy[v], z = cz[v];
        if (pass == 1) {
            acc[0] = s;
            acc[1] = s * x;
            acc[2] = s * y;
        } else {
            const double Cx = x - bulk[node * 3 + 1], Cy = y - bulk[node * 3 + 2], Cz = z;
            const double C2 = Cx * Cx + Cy * Cy + Cz * Cz;
            acc[0] = s * C2;
            acc[1] = s * C2 * Cx;
            acc[2] = s * C2 * Cy;
            acc[3] = s * Cx * Cx;
            acc[4] = s * Cy * Cy;
            acc[5] = s * Cx * Cy;
        }
    }
    block_reduce<6>(acc, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < 6; ++i) part[(size_t(node) * nblk + blockIdx.x) * 6 + i] = acc[i];
}

// ordered sum of partials per node; pass 1 -> bulk {rho, ux, uy}; pass 2 -> out
__global__ void __launch_bounds__(kBlk)
k_mom_final(int pass, int nnode, int nblk, double cw, const double* __restrict__ part,
            double* __restrict__ bulk, double* __restrict__ out /*[node][9]*/) {
    const int node = blockIdx.x * kBlk + threadIdx.x;
    if (node >= nnode) return;
    double S[6] = {0, 0, 0, 0, 0, 0};
    for (int b = 0; b < nblk; ++b)
        for (int i = 0; i < 6; ++i) S[i] += part[(size_t(node) * nblk + b) * 6 + i];
    if (pass == 1) {
        const double rho = S[0] * cw;
        bulk[node * 3 + 0] = rho;
        bulk[node * 3 + 1] = S[1] * cw / rho;
        bulk[node * 3 + 2] = S[2] * cw / rho;
    } else {
        const double rho = bulk[node * 3 + 0];
        double* o = out + size_t(node) * 9;
        o[0] = rho;
        o[1] = bulk[node * 3 + 1];
        o[2] = bulk[node * 3 + 2];
        o[3] = S[0] * (2.0 / 3.0 * cw) / rho;   // T
        o[4] = S[1] * cw;                       // q_x
        o[5] = S[2] * cw;                       // q_y
        o[6] = 2.0 * S[3] * cw;                 // P_xx
        o[7] = 2.0 * S[4] * cw;                 // P_yy
        o[8] = 2.0 * S[5] * cw;                 // P_xy
    }
}

// --------------------------------------------------------------- residual
// part[blk] = {sum (a-b)^2, sum b^2}   (residual.py:21-22)
template <typename T>
__global__ void __launch_bounds__(kBlk)
k_resid_partial(long n, const T* __restrict__ cur, const T* __restrict__ prev,
                double* __restrict__ part) {
    __shared__ double sh[2 * kBlk];
    double acc[2] = {0.0, 0.0};
    for (long i = long(blockIdx.x) * kBlk + threadIdx.x; i < n; i += long(gridDim.x) * kBlk) {
        const double d = double(cur[i]) - double(prev[i]);
        acc[0] += d * d;
        acc[1] += double(prev[i]) * double(prev[i]);
    }
    block_reduce<2>(acc, sh);
    if (threadIdx.x == 0) {
        part[blockIdx.x * 2] = acc[0];
        part[blockIdx.x * 2 + 1] = acc[1];
    }
}

__global__ void __launch_bounds__(kBlk)
k_resid_final(int nblk, const double* __restrict__ part, double* __restrict__ out) {
    __shared__ double sh[2 * kBlk];
    double acc[2] = {0.0, 0.0};
    for (int b = threadIdx.x; b < nblk; b += kBlk) {
        acc[0] += part[b * 2];
        acc[1] += part[b * 2 + 1];
    }
    block_reduce<2>(acc, sh);
    if (threadIdx.x == 0) {
        out[0] = acc[0];
        out[1] = acc[1];
    }
}

// ------------------------------------------------------------ type dispatch
template <typename T>
static int transport_rhs_t(const ::dgfs_transport_t* tp, const void* u, void* ux, double t,
                           hipStream_t st) {
    (void)t;
    AdvDev a{tp->K, tp->Ne, tp->Nv, tp->D, tp->BL, tp->BR, tp->gLD, tp->gRD, tp->invjac, tp->cw,
             tp->Dm, tp->De};
    WallDev w[2];
    for (int s = 0; s < 2; ++s) {
        const ::dgfs_wall_t& x = s == 0 ? tp->left : tp->right;
        w[s] = WallDev{x.kind, x.nl, x.ushift, x.f0, x.ux, x.uy, x.uz, x.T};
    }
    const bool need = w[0].kind == DGFS_BC_DIFFUSE || w[0].kind == DGFS_BC_DIFFUSE_EXPR ||
                      w[1].kind == DGFS_BC_DIFFUSE || w[1].kind == DGFS_BC_DIFFUSE_EXPR;
    const int nb = (tp->Nv + kBlk - 1) / kBlk;
    if (need) {
        hipLaunchKernelGGL((k_bc_partial<T>), dim3(nb), dim3(kBlk), 0, st, a, (const T*)u,
                           (const T*)tp->cx, (const T*)tp->cy, (const T*)tp->cz, w[0], w[1],
                           tp->work);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    // the wall densities are formed inside k_advect from the partials
    const long tot = long(tp->Ne) * tp->Nv;
    hipLaunchKernelGGL((k_advect<T>), dim3((tot + kBlk - 1) / kBlk), dim3(kBlk), 0, st, a,
                       (const T*)u, (const T*)tp->cx, (const T*)tp->cy, (const T*)tp->cz, w[0],
                       w[1], need ? nb : 0, (const double*)tp->work, tp->nw, (T*)ux);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

}  // namespace dgfs

using namespace dgfs;

extern "C" size_t dgfs_transport_work_bytes(int Nv) {
    return size_t((Nv + kBlk - 1) / kBlk) * 4 * sizeof(double);
}

extern "C" int dgfs_transport_rhs(const dgfs_transport_t* tp, const void* u, void* ux, double t,
                                  hipStream_t stream) {
    DGFS_REQUIRE(tp && u && ux, "transport_rhs: null argument");
    DGFS_REQUIRE(tp->K > 0 && tp->K <= kMaxK && tp->Ne > 0 && tp->Nv > 0,
                 "transport_rhs: bad sizes K=%d Ne=%d Nv=%d", tp->K, tp->Ne, tp->Nv);
    DGFS_REQUIRE(tp->D && tp->BL && tp->BR && tp->gLD && tp->gRD && tp->invjac && tp->cx &&
                 tp->cy && tp->cz && tp->work && tp->nw, "transport_rhs: null table");
    for (const dgfs_wall_t* w : {&tp->left, &tp->right}) {
        DGFS_REQUIRE(w->kind >= DGFS_BC_COPY && w->kind <= DGFS_BC_VANISH,
                     "transport_rhs: bad wall kind %d", w->kind);
        DGFS_REQUIRE(!(w->kind == DGFS_BC_DIFFUSE || w->kind == DGFS_BC_INLET) || w->f0,
                     "transport_rhs: wall state f0 missing");
    }
    if (tp->dtype == DGFS_F64) return transport_rhs_t<double>(tp, u, ux, t, stream);
    DGFS_REQUIRE(tp->dtype == DGFS_F32, "transport_rhs: bad dtype");
    return transport_rhs_t<float>(tp, u, ux, t, stream);
}

namespace {

struct TransportBatch {
    int dtype = 0, nprob = 0, nbmax = 0;
    long total = 0;
    ProbDev* P = nullptr;       // device [nprob]
    double* part = nullptr;     // device [nprob][nbmax][4]
    double* nw = nullptr;       // device [nprob][2]
};

int check_transport(const dgfs_transport_t* tp) {
    DGFS_REQUIRE(tp->K > 0 && tp->K <= kMaxK && tp->Ne > 0 && tp->Nv > 0,
                 "transport: bad sizes K=%d Ne=%d Nv=%d", tp->K, tp->Ne, tp->Nv);
    DGFS_REQUIRE(tp->D && tp->BL && tp->BR && tp->gLD && tp->gRD && tp->invjac && tp->cx &&
                 tp->cy && tp->cz, "transport: null table");
    for (const dgfs_wall_t* w : {&tp->left, &tp->right}) {
        DGFS_REQUIRE(w->kind >= DGFS_BC_COPY && w->kind <= DGFS_BC_VANISH,
                     "transport: bad wall kind %d", w->kind);
        DGFS_REQUIRE(!(w->kind == DGFS_BC_DIFFUSE || w->kind == DGFS_BC_INLET) || w->f0,
                     "transport: wall state f0 missing");
    }
    return 0;
}

void free_batch(TransportBatch* b) {
    if (!b) return;
    for (void* p : {(void*)b->P, (void*)b->part, (void*)b->nw})
        if (p) (void)hipFree(p);
    delete b;
}

}  // namespace

extern "C" int dgfs_transport_batch_create(const dgfs_transport_t* tps, int nprob,
                                           const long* offset, dgfs_transport_batch_t* out) {
    DGFS_REQUIRE(out, "transport_batch_create: null output handle");
    *out = nullptr;
    DGFS_REQUIRE(tps && offset && nprob > 0, "transport_batch_create: bad arguments");
    const int dtype = tps[0].dtype;
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "transport_batch_create: bad dtype");
    std::vector<ProbDev> host(nprob);
    long total = 0;
    int nbmax = 1;
    for (int p = 0; p < nprob; ++p) {
        const dgfs_transport_t* tp = tps + p;
        if (int st = check_transport(tp)) return st;
        DGFS_REQUIRE(tp->dtype == dtype, "transport_batch_create: problem %d has another dtype", p);
        DGFS_REQUIRE(offset[p] >= 0, "transport_batch_create: negative offset");
        DGFS_REQUIRE(tp->left.kind != DGFS_BC_DIFFUSE_EXPR && tp->right.kind != DGFS_BC_DIFFUSE_EXPR,
                     "transport_batch_create: problem %d has an expression wall (its parameters "
                     "change with t; use dgfs_transport_rhs)", p);
        ProbDev& q = host[p];
        q.a = AdvDev{tp->K, tp->Ne, tp->Nv, tp->D, tp->BL, tp->BR, tp->gLD, tp->gRD, tp->invjac,
                     tp->cw, tp->Dm, tp->De};
        const dgfs_wall_t* ws[2] = {&tp->left, &tp->right};
        WallDev* wd[2] = {&q.wl, &q.wr};
        bool diffuse = false;
        for (int s = 0; s < 2; ++s) {
            *wd[s] = WallDev{ws[s]->kind, ws[s]->nl, ws[s]->ushift, ws[s]->f0, ws[s]->ux, ws[s]->uy,
                             ws[s]->uz, ws[s]->T};
            diffuse = diffuse || ws[s]->kind == DGFS_BC_DIFFUSE;
        }
        q.cx = tp->cx;
        q.cy = tp->cy;
        q.cz = tp->cz;
        q.off = offset[p];
        q.start = total;
        q.nb = diffuse ? (tp->Nv + kBlk - 1) / kBlk : 0;
        nbmax = std::max(nbmax, q.nb);
        total += long(tp->Ne) * tp->Nv;
    }
    auto* b = new TransportBatch();
    b->dtype = dtype;
    b->nprob = nprob;
    b->nbmax = nbmax;
    b->total = total;
    if (hipMalloc(&b->P, sizeof(ProbDev) * nprob) != hipSuccess ||
        hipMalloc(&b->part, sizeof(double) * 4 * size_t(nprob) * nbmax) != hipSuccess ||
        hipMalloc(&b->nw, sizeof(double) * 2 * nprob) != hipSuccess) {
        free_batch(b);
        set_error("transport_batch_create: allocation failed");
        return 1;
    }
    if (hipMemcpy(b->P, host.data(), sizeof(ProbDev) * nprob, hipMemcpyHostToDevice) !=
            hipSuccess ||
        hipMemset(b->nw, 0, sizeof(double) * 2 * nprob) != hipSuccess) {
        free_batch(b);
        set_error("transport_batch_create: upload failed");
        return 1;
    }
    *out = b;
    return 0;
}

extern "C" int dgfs_transport_batch_destroy(dgfs_transport_batch_t batch) {
    free_batch(static_cast<TransportBatch*>(batch));
    return 0;
}

template <typename T>
static int transport_batch_t(TransportBatch* b, const void* u, void* ux, hipStream_t st) {
    hipLaunchKernelGGL((k_bc_partial_batch<T>), dim3(b->nbmax, b->nprob), dim3(kBlk), 0, st,
                       b->P, b->nbmax, (const T*)u, b->part);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_bc_final_batch, dim3(b->nprob), dim3(kBlk), 0, st, b->P, b->nbmax,
                       b->part, b->nw);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL((k_advect_batch<T>), dim3((b->total + kBlk - 1) / kBlk), dim3(kBlk), 0, st,
                       b->P, b->nprob, b->total, (const T*)u, b->nw, (T*)ux);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_transport_batch_rhs(dgfs_transport_batch_t batch, const void* u, void* ux,
                                        hipStream_t stream) {
    auto* b = static_cast<TransportBatch*>(batch);
    DGFS_REQUIRE(b && u && ux, "transport_batch_rhs: null argument");
    if (b->dtype == DGFS_F64) return transport_batch_t<double>(b, u, ux, stream);
    return transport_batch_t<float>(b, u, ux, stream);
}

extern "C" int dgfs_transport_batch_nw(dgfs_transport_batch_t batch, double* nw_host) {
    auto* b = static_cast<TransportBatch*>(batch);
    DGFS_REQUIRE(b && nw_host, "transport_batch_nw: null argument");
    DGFS_HIP_CHECK(hipMemcpy(nw_host, b->nw, sizeof(double) * 2 * b->nprob, hipMemcpyDeviceToHost));
    return 0;
}

extern "C" int dgfs_matop(int dtype, int Kout, int Kin, long ncol, const double* A,
                          const void* in, void* out, int accumulate, hipStream_t stream) {
    DGFS_REQUIRE(Kout > 0 && Kin > 0 && Kin <= kMaxK && ncol >= 0, "matop: bad sizes");
    DGFS_REQUIRE(A && in && out, "matop: null argument");
    if (ncol == 0) return 0;
    const dim3 g((ncol + kBlk - 1) / kBlk);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL((k_matop<double>), g, dim3(kBlk), 0, stream, Kout, Kin, ncol, A,
                           (const double*)in, (double*)out, accumulate);
    else
        hipLaunchKernelGGL((k_matop<float>), g, dim3(kBlk), 0, stream, Kout, Kin, ncol, A,
                           (const float*)in, (float*)out, accumulate);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_axnpby(int dtype, long n, int nv, void* x0, const void* x1, const void* x2,
                           const void* x3, double a0, double a1, double a2, double a3,
                           hipStream_t stream) {
    DGFS_REQUIRE(nv >= 1 && nv <= 4 && n >= 0 && x0, "axnpby: bad arguments");
    DGFS_REQUIRE((nv < 2 || x1) && (nv < 3 || x2) && (nv < 4 || x3), "axnpby: null operand");
    if (n == 0) return 0;
    const dim3 g((n + kBlk - 1) / kBlk);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL((k_axnpby<double>), g, dim3(kBlk), 0, stream, n, nv, (double*)x0,
                           (const double*)x1, (const double*)x2, (const double*)x3, a0, a1, a2,
                           a3);
    else
        hipLaunchKernelGGL((k_axnpby<float>), g, dim3(kBlk), 0, stream, n, nv, (float*)x0,
                           (const float*)x1, (const float*)x2, (const float*)x3, float(a0),
                           float(a1), float(a2), float(a3));
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_axnpby_compensated(long n, int nv, float* x0, float* x0_lo, const float* x1,
                                       const float* x1_lo, const float* x2, const float* x2_lo,
                                       const float* x3, const float* x3_lo, double a0, double a1,
                                       double a2, double a3, hipStream_t stream) {
    DGFS_REQUIRE(nv >= 1 && nv <= 4 && n >= 0 && x0, "axnpby_compensated: bad arguments");
    DGFS_REQUIRE((nv < 2 || x1) && (nv < 3 || x2) && (nv < 4 || x3),
                 "axnpby_compensated: null operand");
    if (n == 0) return 0;
    const dim3 g((n + kBlk - 1) / kBlk);
    hipLaunchKernelGGL(k_axnpby_df, g, dim3(kBlk), 0, stream, n, nv, x0, x0_lo, x1, x1_lo, x2,
                       x2_lo, x3, x3_lo, split_coef(a0), split_coef(a1), split_coef(a2),
                       split_coef(a3));
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" size_t dgfs_moments_work_bytes(int Nv, int nnode) {
    const int nb = (Nv + kBlk - 1) / kBlk;
    return (size_t(nnode) * nb * 6 + size_t(nnode) * 3) * sizeof(double);
}

extern "C" int dgfs_moments(int dtype, int K, int Ne, int Nqr, int Nv, const double* im,
                            const void* u, const void* cx, const void* cy, const void* cz,
                            double cw, double* work, double* out, hipStream_t stream) {
    DGFS_REQUIRE(K > 0 && Ne > 0 && Nqr > 0 && Nv > 0, "moments: bad sizes");
    DGFS_REQUIRE(im && u && cx && cy && cz && work && out, "moments: null argument");
    const int nnode = Nqr * Ne;
    const int nb = (Nv + kBlk - 1) / kBlk;
    double* part = work;
    double* bulk = work + size_t(nnode) * nb * 6;
    for (int pass = 1; pass <= 2; ++pass) {
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL((k_mom_partial<double>), dim3(nb, nnode), dim3(kBlk), 0, stream,
                               pass, K, Ne, Nqr, Nv, im, (const double*)u, (const double*)cx,
                               (const double*)cy, (const double*)cz, bulk, part);
        else
            hipLaunchKernelGGL((k_mom_partial<float>), dim3(nb, nnode), dim3(kBlk), 0, stream,
                               pass, K, Ne, Nqr, Nv, im, (const float*)u, (const float*)cx,
                               (const float*)cy, (const float*)cz, bulk, part);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL(k_mom_final, dim3((nnode + kBlk - 1) / kBlk), dim3(kBlk), 0, stream,
                           pass, nnode, nb, cw, part, bulk, out);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    return 0;
}

extern "C" size_t dgfs_residual_work_bytes(void) { return (1024 * 2 + 2) * sizeof(double); }

extern "C" int dgfs_residual(int dtype, long n, const void* cur, const void* prev, double* work,
                             double* out2, hipStream_t stream) {
    DGFS_REQUIRE(n > 0 && cur && prev && work && out2, "residual: bad arguments");
    long nb = (n + kBlk - 1) / kBlk;
    if (nb > 1024) nb = 1024;
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL((k_resid_partial<double>), dim3(nb), dim3(kBlk), 0, stream, n,
                           (const double*)cur, (const double*)prev, work);
    else
        hipLaunchKernelGGL((k_resid_partial<float>), dim3(nb), dim3(kBlk), 0, stream, n,
                           (const float*)cur, (const float*)prev, work);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(kBlk), 0, stream, int(nb), work, out2);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}
