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
//   C  plane : per (task, k1-plane): G^ = sum_xy W_xy FFT2_{y,z}(acc_xy),
//              W_xy = qw_x sw / n * b_pq[x,y] (complex, per direction: the VSS
//              kernel depends on the direction, so the sum over y cannot be
//              taken in physical space as the single-species gain does);
//              then IFFT2 of G^ and of c_pq G
//   D  line  : inverse along k1 and Q[out] += prefac (Re gain - f_p Re loss)
//              (k_line_output of the single-species operator)
//
// Unlike the single-species operator the two inverse transforms are not
// packed into one: with unequal mass fractions they carry different phases,
// and both products are kept complex as in the reference.
#pragma once
#include "collision_kernels.h"

namespace dgfs {

template <typename T, int N>
constexpr int bi_split() { return pick_split<T>(N); }

// B: one work-group per (task, xy, residue r).  acc layout
// [t][xy][k1][r][y][j] with z = S*j + r (the layout k_bi_plane reads).
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

// C: one work-group per plane (t, k1); slot s of PL takes xy = s, s+PL, ...
// H[t][0 = loss | 1 = gain][k1][y][z], as k_line_output reads it.
template <typename T, int N>
__global__ void __launch_bounds__(planes_per_wg(N) * N)
k_bi_plane(const cplx<T>* __restrict__ acc_in, const cplx<T>* __restrict__ wtab,
           const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ ctab, int NXY, int SB,
           cplx<T>* __restrict__ H) {
    constexpr int PL = planes_per_wg(N), NT = PL * N, NN = N * N, n = N * NN, PS = N * (N + 1);
    static_assert(PL >= 2, "plane kernel needs two slots");
    __shared__ cplx<T> L[PL * PS];
    const int k1 = blockIdx.x % N, t = blockIdx.x / N;
    const int slot = threadIdx.x / N, c = threadIdx.x - slot * N;
    cplx<T>* P = L + slot * PS;
    const int NSB = N / SB;
    cplx<T> G[N];
    sfor<0, N>([&](auto I) { G[I] = mk(T(0), T(0)); });
    const int nround = (NXY + PL - 1) / PL;
    for (int rd = 0; rd < nround; ++rd) {
        const int xy = rd * PL + slot;
        const bool live = xy < NXY;
        __syncthreads();
        if (live) {
            const cplx<T>* src = acc_in + (size_t(t) * NXY + xy) * n + size_t(k1) * NN;
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
    if (slot == 0)
        for (int o = 1; o < PL; ++o)
            sfor<0, N>([&](auto K2) { G[K2] = G[K2] + L[o * PS + K2 * (N + 1) + c]; });
    __syncthreads();
    if (slot == 0) {
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
    void* acc = nullptr;    // [chunk][Nrho*M][n] cplx<T>
    void* H = nullptr;      // [ntask_max][2][n] cplx<T>
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
        const long npl = long(nt) * 2 * N;
        hipLaunchKernelGGL((k_fwd_plane<T, N>), dim3((npl + PL - 1) / PL), dim3(PL * N), 0, st,
                           (const T*)up, (const T*)uq, tc, 2, npl, (cplx<T>*)pl->Iw);
        DGFS_HIP_CHECK(hipGetLastError());
        const long nli = long(nt) * 2 * NN;
        hipLaunchKernelGGL((k_fwd_line<T, N>), dim3((nli + 255) / 256), dim3(256), 0, st, nli, 2,
                           (const cplx<T>*)pl->Iw, (cplx<T>*)pl->Fhat);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL((k_bi_gain<T, N, S>), dim3(nt * NXY * S), dim3(round64(N * (N / S))),
                           0, st, (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase, NXY,
                           (cplx<T>*)pl->acc);
        DGFS_HIP_CHECK(hipGetLastError());
        hipLaunchKernelGGL((k_bi_plane<T, N>), dim3(nt * N), dim3(PL * N), 0, st,
                           (const cplx<T>*)pl->acc, (const cplx<T>*)pl->wtab,
                           (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->ctab, NXY, S,
                           (cplx<T>*)pl->H + size_t(c0) * 2 * n);
        DGFS_HIP_CHECK(hipGetLastError());
    }
    const long nlo = long(ngroup) * NN;
    hipLaunchKernelGGL((k_line_output<T, N>), dim3((nlo + 255) / 256), dim3(256), 0, st,
                       (const T*)up, tv, (const cplx<T>*)pl->H, nlo, gstart, gslice,
                       T(pl->prefac), (T*)Q);
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
DGFS_COLLIDE_SIZES(DGFS_BI_EXTERN, double)

}  // namespace dgfs
