// Binary-mixture VSS fast-spectral collision operator, one species pair
// (p, q), batched over collision tasks.
//
// Reference behaviour (jaisw7/dgfs1D_gpu): DGFSBiVSSScatteringModel.fs
// (bi/scattering.py:261-321) with kernels bi/kernels/scattering/vss.mako:
// 96-285.  Per node the reference runs 10 launches and batched M*N_rho cuFFT
// transforms, nodes one after another.  Here every node of a pair shares one
// launch of each of five kernels:
//
//   A  fwd   : F = FFT3(f_p)/n, G = FFT3(f_q)/n per task (the forward kernels
//              of the single-species operator, collision_kernels.h)
//   B  gain  : per (task, radius x, direction y, z-residue r):
//                u1 = IFFT3(exp(+i m_q/(m_p+m_q) a_xy) F),
//                u2 = IFFT3(exp(-i m_p/(m_p+m_q) a_xy) G)  (Fft3Engine, complex),
//              and the forward transform along x of u1 u2 in registers
//   C  plane : per (task, k1-plane, xy group): partial sums
//              G^_g = sum_{xy in g} W_xy FFT2_{y,z}(acc_xy), summed in fixed
//              order by a second kernel, which also does the inverses:
//              W_xy = qw_x sw / n * b_pq[x,y] (complex, per direction: the VSS
//              kernel depends on the direction, so the sum over y cannot be
//              taken in physical space as the single-species gain does);
//              then IFFT2 of G^ and of c_pq G
//   D  line  : inverse along k1 and Q[out] += prefac (Re gain - f_p Re loss)
//              (k_line_output of the single-species operator)
//
// Where the engine's LDS cube needs a residue split (FP64 N >= 24, FP32
// N = 32), B is done by two streaming kernels instead (k_bi_yz: phase and the
// l3, l2 inverses per l1 plane in LDS; k_bi_xprod: the l1 inverses, the
// product and the forward x transform per pencil), so no work-group reloads F.
//
// Unlike the single-species operator the two inverse transforms are not
// packed into one: with unequal mass fractions they carry different phases,
// and both products are kept complex as in the reference.
#pragma once
#include "collision_kernels.h"

namespace dgfs {

template <typename T, int N>
constexpr int bi_split() { return pick_split<T>(N); }

// grids whose gain cube needs a residue split take the two-kernel slab path
template <typename T, int N>
constexpr bool bi_slabs() { return bi_split<T, N>() > 1; }

// B: one work-group per (task, xy, residue r).  acc layout
// [t][xy][k1][r][y][j] with z = S*j + r (the layout k_bi_plane_part reads).
template <typename T, int N, int S>
__global__ void __launch_bounds__(round64(N * (N / S)))
k_bi_gain(const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ phase, int NXY,
          cplx<T>* __restrict__ acc_out) {
    using E = Fft3Engine<T, N, S, round64(N * (N / S)), +1>;
    constexpr int n = N * N * N, NS = N / S;
    __shared__ cplx<T> V[E::LDS_ELEMS];
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int r = lb % S;
    const int xy = (lb / S) % NXY;
    const int t = lb / (S * NXY);
    typename E::Split sp;
    sp.init(r);
    const cplx<T>* ph = phase + size_t(xy) * 2 * 3 * N;     // [side][axis][l]
    auto map1 = [](int pp, int& i1, int& i2) { i1 = pp / N; i2 = pp - (pp / N) * N; };
    auto map3 = [](int P, int& o2, int& j) { o2 = P / NS; j = P - (P / NS) * NS; };
    cplx<T> R1[N];
    sfor<0, 2>([&](auto Si) {
        constexpr int side = Si;
        const cplx<T>* F = Fhat + (size_t(t) * 2 + side) * n;     // [l3][l1][l2]
        const cplx<T>* e = ph + side * 3 * N;
        // the separable phase exp(i th_1 l1) exp(i th_2 l2) exp(i th_3 l3) is
        // applied per axis just before that axis' transform: l3 in the load,
        // l2 in pass 2 and l1 in pass 3 (these touch only this residue's data)
        E::run(V, sp, map1,
               [&](int i1, int i2, int i3) { return e[2 * N + i3] * F[(size_t(i3) * N + i1) * N + i2]; },
               [&](auto I, cplx<T> v) { return e[N + int(I)] * v; }, map3,
               [&](auto I, cplx<T> v) { return e[int(I)] * v; },
               [&](int o2, int j, const cplx<T>* o) {
                   if constexpr (side == 0) {
                       sfor<0, N>([&](auto I) { R1[I] = o[I]; });
                   } else {
                       cplx<T> pr[N];
                       sfor<0, N>([&](auto I) { pr[I] = R1[I] * o[I]; });
                       RegDFT<T, N, -1>::template run<1, 0>(pr, R1);
                       cplx<T>* dst = acc_out + (size_t(t) * NXY + xy) * n;
                       sfor<0, N>([&](auto I) {
                           dst[((size_t(I) * S + r) * N + o2) * NS + j] = R1[I];
                       });
                   }
               });
    });
}

// B', for grids whose engine cube does not fit LDS whole (residue split
// S > 1): two streaming kernels instead of k_bi_gain, so that no work-group
// reloads F.
// B'1: one N-thread slot per plane (t, xy, side, l1): P[l2][l3] =
//      e1(l1) e2(l2) e3(l3) F[l3][l1][l2], inverse along l3 then l2 in LDS ->
//      Y[t][xy][side][l1][y][z].
template <typename T, int N>
__global__ void __launch_bounds__(planes_per_wg(N) * N)
k_bi_yz(const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ phase, int NXY,
        long nplanes, cplx<T>* __restrict__ Y) {
    constexpr int PL = planes_per_wg(N), NT = PL * N, NN = N * N, n = N * NN, PS = N * (N + 1);
    __shared__ cplx<T> L[PL * PS];
    const long plane0 = long(blockIdx.x) * PL;
    const int pl = threadIdx.x / N, c = threadIdx.x - pl * N;
    const long pid = plane0 + pl;
    const bool live = pid < nplanes;
    cplx<T>* P = L + pl * PS;
    if (live) {
        const int l1 = int(pid % N);
        const long tws = pid / N;                      // (t*NXY + xy)*2 + side
        const int side = int(tws % 2);
        const long txy = tws / 2;
        const int xy = int(txy % NXY);
        const long t = txy / NXY;
        const cplx<T>* F = Fhat + (size_t(t) * 2 + side) * n + size_t(l1) * N + c;
        const cplx<T>* e = phase + (size_t(xy) * 2 + side) * 3 * N;
        const cplx<T> e12 = e[l1] * e[N + c];          // thread c = l2
        sfor<0, N>([&](auto J) {
            P[c * (N + 1) + J] = (e12 * e[2 * N + J]) * F[size_t(J) * NN];
        });
    }
    __syncthreads();
    if (live) lds_fft_row<T, N, +1>(P, c);     // row l2: l3 -> z
    __syncthreads();
    if (live) lds_fft_col<T, N, +1>(P, c);     // column z: l2 -> y
    __syncthreads();
    for (int k = threadIdx.x; k < PL * NN; k += NT) {
        const int q = k / NN, r = k - q * NN;
        const long pq = plane0 + q;
        if (pq >= nplanes) continue;
        Y[pq * NN + r] = L[q * PS + (r / N) * (N + 1) + r % N];     // [y][z]
    }
}

// B'2: one thread per (t, xy, y, z) pencil: u1, u2 = inverse along l1 of the
// two sides, forward along x of u1 u2, written over side 0's pencil
// (acc[t][xy][k1][y][z] with stride 2n per (t, xy)).
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_bi_xprod(long npen, cplx<T>* __restrict__ Y) {
    constexpr int NN = N * N, n = N * NN;
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= npen) return;
    const long txy = gid / NN;
    const int yz = int(gid - txy * NN);
    cplx<T>* Y0 = Y + txy * 2 * n + yz;
    const cplx<T>* Y1 = Y0 + n;
    cplx<T> a[N], R1[N], o[N];
    sfor<0, N>([&](auto I) { a[I] = Y0[size_t(I) * NN]; });
    RegDFT<T, N, +1>::template run<1, 0>(a, R1);
    sfor<0, N>([&](auto I) { a[I] = Y1[size_t(I) * NN]; });
    RegDFT<T, N, +1>::template run<1, 0>(a, o);
    sfor<0, N>([&](auto I) { a[I] = R1[I] * o[I]; });
    RegDFT<T, N, -1>::template run<1, 0>(a, R1);
    sfor<0, N>([&](auto I) { Y0[size_t(I) * NN] = R1[I]; });
}

// C1: one work-group per (t, k1, xy-group g): slot s of PL takes the group's
// xy = xy0 + s, xy0 + s + PL, ...; the group's partial
// G^_g(k1, :, :) = sum_xy W_xy FFT2_{y,z}(acc_xy(k1)) goes to
// Gp[t][g][k1][k2][k3].  acc of (t, xy) starts at (t*NXY + xy)*accstride.
template <typename T, int N>
__global__ void __launch_bounds__(planes_per_wg(N) * N)
k_bi_plane_part(const cplx<T>* __restrict__ acc_in, long accstride,
                const cplx<T>* __restrict__ wtab, int NXY, int SB, int XG,
                cplx<T>* __restrict__ Gp) {
    constexpr int PL = planes_per_wg(N), NN = N * N, n = N * NN, PS = N * (N + 1);
    __shared__ cplx<T> L[PL * PS];
    const int g = blockIdx.x % XG;
    const int k1 = (blockIdx.x / XG) % N, t = blockIdx.x / (XG * N);
    const int slot = threadIdx.x / N, c = threadIdx.x - slot * N;
    cplx<T>* P = L + slot * PS;
    const int per = (NXY + XG - 1) / XG;
    const int xy0 = g * per < NXY ? g * per : NXY;
    const int xy1 = xy0 + per < NXY ? xy0 + per : NXY;
    const int NSB = N / SB;
    cplx<T> G[N];
    sfor<0, N>([&](auto I) { G[I] = mk(T(0), T(0)); });
    const int nround = (xy1 - xy0 + PL - 1) / PL;
    for (int rd = 0; rd < nround; ++rd) {
        const int xy = xy0 + rd * PL + slot;
        const bool live = xy < xy1;
        __syncthreads();
        if (live) {
            const cplx<T>* src = acc_in + (size_t(t) * NXY + xy) * accstride + size_t(k1) * NN;
            for (int pos = c; pos < NN; pos += N) {
                const int r = pos / (N * NSB), rem = pos - r * N * NSB;
                const int y = rem / NSB, z = (rem - y * NSB) * SB + r;
                P[y * (N + 1) + z] = src[pos];
            }
        }
        __syncthreads();
        if (live) lds_fft_row<T, N, -1>(P, c);     // z -> k3
        __syncthreads();
        if (live) {
            lds_fft_col<T, N, -1>(P, c);            // y -> k2 (column c = k3)
            const cplx<T>* w = wtab + size_t(xy) * n + size_t(k1) * NN + c;   // [k1][k2][k3]
            sfor<0, N>([&](auto K2) { G[K2] = G[K2] + w[size_t(K2) * N] * P[K2 * (N + 1) + c]; });
        }
    }
    __syncthreads();
    sfor<0, N>([&](auto K2) { P[K2 * (N + 1) + c] = G[K2]; });
    __syncthreads();
    if (slot == 0) {
        for (int o = 1; o < PL; ++o)
            sfor<0, N>([&](auto K2) { G[K2] = G[K2] + L[o * PS + K2 * (N + 1) + c]; });
        cplx<T>* dst = Gp + ((size_t(t) * XG + g) * N + k1) * NN + c;
        sfor<0, N>([&](auto K2) { dst[size_t(K2) * N] = G[K2]; });
    }
}

// C2: one work-group per plane (t, k1): G^ = sum_g Gp[t][g][k1] (fixed
// order), IFFT2 of G^ (slot 0) and of c_pq G_q (slot 1) ->
// H[t][0 = loss | 1 = gain][k1][y][z], as k_line_output reads it.
template <typename T, int N>
__global__ void __launch_bounds__(planes_per_wg(N) * N)
k_bi_plane_fin(const cplx<T>* __restrict__ Gp, int XG, const cplx<T>* __restrict__ Fhat,
               const cplx<T>* __restrict__ ctab, cplx<T>* __restrict__ H) {
    constexpr int PL = planes_per_wg(N), NT = PL * N, NN = N * N, n = N * NN, PS = N * (N + 1);
    static_assert(PL >= 2, "plane kernel needs two slots");
    __shared__ cplx<T> L[PL * PS];
    const int k1 = blockIdx.x % N, t = blockIdx.x / N;
    const int slot = threadIdx.x / N, c = threadIdx.x - slot * N;
    cplx<T>* P = L + slot * PS;
    if (slot == 0) {
        cplx<T> G[N];
        sfor<0, N>([&](auto K2) { G[K2] = mk(T(0), T(0)); });
        for (int g = 0; g < XG; ++g) {
            const cplx<T>* src = Gp + ((size_t(t) * XG + g) * N + k1) * NN + c;
            sfor<0, N>([&](auto K2) { G[K2] = G[K2] + src[size_t(K2) * N]; });
        }
        sfor<0, N>([&](auto K2) { P[K2 * (N + 1) + c] = G[K2]; });
    } else if (slot == 1) {
        // loss plane: c_pq * G_q at l1 = k1, column c = l3
        const cplx<T>* Fg = Fhat + (size_t(t) * 2 + 1) * n;
        sfor<0, N>([&](auto L2) {
            P[L2 * (N + 1) + c] = ctab[size_t(k1) * NN + L2 * N + c] *
                                  Fg[(size_t(c) * N + k1) * N + L2];
        });
    }
    __syncthreads();
    if (slot < 2) lds_fft_col<T, N, +1>(P, c);      // k2 -> y
    __syncthreads();
    if (slot < 2) lds_fft_row<T, N, +1>(P, c);      // k3 -> z
    __syncthreads();
    for (int i = threadIdx.x; i < 2 * NN; i += NT) {
        const int which = i / NN, rr = i - which * NN;    // slot 0 = gain, slot 1 = loss
        H[((size_t(t) * 2 + (which == 0 ? 1 : 0)) * N + k1) * NN + rr] =
            L[which * PS + (rr / N) * (N + 1) + rr % N];
    }
}

// xy groups of the plane stage: enough work-groups to fill the chip, at
// least one slot round per group
template <int N>
inline int bi_xgroups(int nt, int NXY) {
    constexpr int PL = planes_per_wg(N);
    const int maxg = (NXY + PL - 1) / PL;
    int g = (2048 + nt * N - 1) / (nt * N);
    return g < 1 ? 1 : (g > maxg ? maxg : g);
}
template <int N>
constexpr int bi_xgroups_max(int NXY) { return (NXY + planes_per_wg(N) - 1) / planes_per_wg(N); }

// ================================================================ the plan
struct BiPlan {
    int N, M, Nrho, Mpre, dtype, ntask_max, chunk;
    double prefac, mp, mq;
    size_t n;
    void* phase = nullptr;  // [Nrho*M][2][3][N] cplx<T>
    void* wtab = nullptr;   // [Nrho*M][k1][k2][k3] cplx<T>  qw sw / n * b_pq
    void* ctab = nullptr;   // [l1][l2][l3] cplx<T>  c_pq
    void* Fhat = nullptr;   // [chunk][2][n] cplx<T>
    void* Iw = nullptr;     // [chunk][2][n] cplx<T>
    void* acc = nullptr;    // [chunk][Nrho*M][n] cplx<T>, or [chunk][Nrho*M][2][n] on the slab path
    void* H = nullptr;      // [ntask_max][2][n] cplx<T>
    void* Gp = nullptr;     // [chunk][xgroups_max][n] cplx<T>  partial G^ per xy group
};

template <typename T, int N>
int bi_launch(BiPlan* pl, const void* up, const void* uq, void* Q, TaskView tv, int ntask,
              int ngroup, const int* gstart, const int* gslice, hipStream_t st) {
    constexpr int S = bi_split<T, N>();
    constexpr int PL = planes_per_wg(N);
    constexpr long NN = long(N) * N;
    const int NXY = pl->Nrho * pl->M;
    const size_t n = pl->n;
    for (int c0 = 0; c0 < ntask; c0 += pl->chunk) {
        const int nt = ntask - c0 < pl->chunk ? ntask - c0 : pl->chunk;
        TaskView tc = tv;
        tc.nterm = tv.nterm + 2 * size_t(c0);
        tc.slice = tv.slice + size_t(c0) * 2 * tv.K;
        tc.coef = static_cast<const T*>(tv.coef) + size_t(c0) * 2 * tv.K;
        const int XG = bi_xgroups<N>(nt, NXY);
        const long npl = long(nt) * 2 * N;
        constexpr int FP = fwd_planes_per_wg(N);
        hipLaunchKernelGGL((k_fwd_plane<T, N, false>), dim3((npl + FP - 1) / FP), dim3(FP * N),
                           0, st, (const T*)up, (const T*)uq, tc, 2, npl, (cplx<T>*)pl->Iw);
        DGFS_HIP_CHECK(hipGetLastError());
        const long nli = long(nt) * 2 * NN;
        hipLaunchKernelGGL((k_fwd_line<T, N>), dim3((nli + 255) / 256), dim3(256), 0, st, nli, 2,
                           (const cplx<T>*)pl->Iw, (cplx<T>*)pl->Fhat);
        DGFS_HIP_CHECK(hipGetLastError());
        if constexpr (bi_slabs<T, N>()) {
            const long npl2 = long(nt) * NXY * 2 * N;
            hipLaunchKernelGGL((k_bi_yz<T, N>), dim3((npl2 + PL - 1) / PL), dim3(PL * N), 0, st,
                               (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase, NXY, npl2,
                               (cplx<T>*)pl->acc);
            DGFS_HIP_CHECK(hipGetLastError());
            const long npen = long(nt) * NXY * NN;
            hipLaunchKernelGGL((k_bi_xprod<T, N>), dim3((npen + 255) / 256), dim3(256), 0, st,
                               npen, (cplx<T>*)pl->acc);
            DGFS_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL((k_bi_plane_part<T, N>), dim3(nt * N * XG), dim3(PL * N), 0, st,
                               (const cplx<T>*)pl->acc, long(2 * n), (const cplx<T>*)pl->wtab,
                               NXY, 1, XG, (cplx<T>*)pl->Gp);
            DGFS_HIP_CHECK(hipGetLastError());
        } else {
            hipLaunchKernelGGL((k_bi_gain<T, N, S>), dim3(nt * NXY * S),
                               dim3(round64(N * (N / S))), 0, st, (const cplx<T>*)pl->Fhat,
                               (const cplx<T>*)pl->phase, NXY, (cplx<T>*)pl->acc);
            DGFS_HIP_CHECK(hipGetLastError());
            hipLaunchKernelGGL((k_bi_plane_part<T, N>), dim3(nt * N * XG), dim3(PL * N), 0, st,
                               (const cplx<T>*)pl->acc, long(n), (const cplx<T>*)pl->wtab, NXY,
                               S, XG, (cplx<T>*)pl->Gp);
            DGFS_HIP_CHECK(hipGetLastError());
        }
        hipLaunchKernelGGL((k_bi_plane_fin<T, N>), dim3(nt * N), dim3(PL * N), 0, st,
                           (const cplx<T>*)pl->Gp, XG, (const cplx<T>*)pl->Fhat,
                           (const cplx<T>*)pl->ctab, (cplx<T>*)pl->H + size_t(c0) * 2 * n);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    const long nlo = long(ngroup) * NN;
    hipLaunchKernelGGL((k_line_output<T, N, false>),
                       dim3((nlo * line_lanes(N) + kLineWG - 1) / kLineWG), dim3(kLineWG), 0, st,
                       (const T*)up, (const T*)nullptr, (const double*)nullptr, tv,
                       (const cplx<T>*)pl->H, nlo, gstart, gslice, T(pl->prefac), (T*)Q);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

#define DGFS_BI_EXTERN(T, N)                                                              \
    extern template int bi_launch<T, N>(BiPlan*, const void*, const void*, void*, TaskView, \
                                        int, int, const int*, const int*, hipStream_t);
#define DGFS_BI_INST(T, N)                                                                \
    template int bi_launch<T, N>(BiPlan*, const void*, const void*, void*, TaskView, int, int, \
                                 const int*, const int*, hipStream_t);

DGFS_COLLIDE_SIZES(DGFS_BI_EXTERN, float)
DGFS_BI_EXTERN(float, 48)
DGFS_COLLIDE_SIZES(DGFS_BI_EXTERN, double)

}  // namespace dgfs
