// Fused fast-spectral VHS collision operator, batched over collision tasks.
//
// Reference behaviour (jaisw7/dgfs1D_gpu): DGFSVHSScatteringModelStd.fs
// (std/scattering.py:346-429) and DGFSVHSGLLScatteringModelStd.fs
// (:251-313) with kernels vhs-gll.mako:43-433.  The reference runs 17 launches
// and 5 batched B*N^3 cuFFT transforms per node, nodes one after another.
// Here one launch of each of four kernels covers every node of an RHS:
//
//   A  fwd   : F^ = FFT3(f)/n, G^ = FFT3(g)/n for every task (f, g are the
//              U/V contractions of the coefficient array, fused into the load):
//              k_fwd_plane (2D over y,z per x-plane) + k_fwd_line (1D over x)
//   B  gain  : for every (task, radius p, residue r, q-group):
//                acc_p(x) = sum_q 1/2 (u+f u-g + u-f u+g)(x),
//                u+-h = IFFT3(exp(+-i a_pq) H^)   (Fft3Engine: 3 LDS passes)
//              cos/sin products of the reference are the same quantity:
//              t1 t2 + t3 t4 = 1/2 (u+f u-g + u-f u+g).  The q-sum is taken
//              in physical space before the forward transform (linearity);
//              the forward transform of acc_p along x is done in registers.
//   C  plane : per (task, k1-plane): G^ = sum_p (2 w_p sw/n) bb1_p FFT2(acc_p),
//              then the inverse plane transforms of G^ and of bb2 G^_g
//   D  line  : inverse along k1 per (y, z) line and
//              Q[out] += prefac (Re IFFT3(sum G^) - f Re IFFT3(bb2 G^_g))
//
// Reference-state form (FP32 rounding error scaled by |f - f0|): with the
// reference state f0 set once (dgfs_collision_reference_set: F0^ and
// u+-_f0 = IFFT3(exp(+-i a_pq) F0^) for every (task, p, q) kept in the plan),
// each call transforms h = f - f0 only and
//   gain  = sum_q u+h (u-h + u-f0) + u-h u+f0  (= Gain(f) - Gain(f0)),
//   loss  = h L[f] + f0 L[h],  Q += Q(f0, f0) + prefac (gain - loss)
// with Q(f0, f0) supplied by the caller in FP64.
//
// Tables (phases, weights, bb2) are built once in FP64 on the host at plan
// creation.
#pragma once
#include "fft3.h"

#include "dgfs_hip.h"

namespace dgfs {

// ------------------------------------------------------------------ sizing
constexpr int round64(int x) { return (x + 63) / 64 * 64; }

template <typename T>
constexpr int pick_split(int N) {
    // smallest S dividing N with N*N*(N/S+1)*sizeof(cplx<T>) <= 144 KiB
    for (int S = 1; S <= N; ++S) {
        if (N % S) continue;
        long bytes = long(N) * N * (N / S + 1) * 2 * long(sizeof(T));
        if (bytes <= 144 * 1024) return S;
    }
    return N;
}

// planes per work-group of the 2D plane kernels (PL * N threads)
constexpr int planes_per_wg(int N) { return N >= 128 ? 2 : (256 / N); }

// task description (device arrays, see dgfs_hip.h)
struct TaskView {
    int K, Ne, n;
    const int* nterm;   // [ntask][2]
    const int* slice;   // [ntask][2][K]
    const void* coef;   // [ntask][2][K]
};

template <typename T>
DGFS_DEV T task_value(const TaskView& tv, const T* u, int t, int which, int v) {
    const int* nt = tv.nterm + 2 * t + which;
    const int* sl = tv.slice + (size_t(t) * 2 + which) * tv.K;
    const T* cf = static_cast<const T*>(tv.coef) + (size_t(t) * 2 + which) * tv.K;
    T acc = T(0);
    const int nterm = *nt;
    for (int a = 0; a < nterm; ++a) acc += cf[a] * u[size_t(sl[a]) * tv.n + v];
    return acc;
}

// ============================================================ plane helpers
// P points at PL planes of N x (N+1) complex in LDS; thread (pl, i) owns row
// or column i of plane pl.
template <typename T, int N, int DIR>
DGFS_DEV void lds_fft_row(cplx<T>* P, int i) {
    cplx<T> a[N], o[N];
    sfor<0, N>([&](auto J) { a[J] = P[i * (N + 1) + J]; });
    RegDFT<T, N, DIR>::template run<1, 0>(a, o);
    sfor<0, N>([&](auto J) { P[i * (N + 1) + J] = o[J]; });
}
template <typename T, int N, int DIR>
DGFS_DEV void lds_fft_col(cplx<T>* P, int i) {
    cplx<T> a[N], o[N];
    sfor<0, N>([&](auto J) { a[J] = P[J * (N + 1) + i]; });
    RegDFT<T, N, DIR>::template run<1, 0>(a, o);
    sfor<0, N>([&](auto J) { P[J * (N + 1) + i] = o[J]; });
}

// ================================================================ kernel A
// A1: plane x of task t's f (or g): 2D forward transform over (y, z);
//     Iw[t][which][x][l3][l2]
//     DEV (reference-state form): ug is the reference state u0; which = 0
//     loads h = f - f0, which = 1 loads s = f + f0 (both with the f terms)
// planes per work-group of k_fwd_plane: two at N >= 32, so that a task set's
// N planes per task spread over many CUs
constexpr int fwd_planes_per_wg(int N) { return N >= 32 ? 2 : planes_per_wg(N); }

template <typename T, int N, bool DEV>
__global__ void __launch_bounds__(fwd_planes_per_wg(N) * N)
k_fwd_plane(const T* __restrict__ u, const T* __restrict__ ug, TaskView tv, int nfg,
            long nplanes, cplx<T>* __restrict__ Iw) {
    constexpr int PL = fwd_planes_per_wg(N), NN = N * N;
    __shared__ cplx<T> L[PL * N * (N + 1)];
    const long plane0 = long(blockIdx.x) * PL;
    const int pl = threadIdx.x / N, c = threadIdx.x - pl * N;
    cplx<T>* P = L + pl * N * (N + 1);
    {
        // thread c of plane pl gathers the points (y, z = c), y < N: the task's
        // terms in the outer loop, the N loads of a term issued together
        T v[N];
        sfor<0, N>([&](auto I) { v[I] = T(0); });
        const long pid = plane0 + pl;
        if (pid < nplanes) {
            const int x = int(pid % N);
            const long tw = pid / N;
            const int which = int(tw % nfg), t = int(tw / nfg);
            auto gather = [&](const T* src, int wh, T sign) {
                const int nt = tv.nterm[2 * t + wh];
                const int* sl = tv.slice + (size_t(t) * 2 + wh) * tv.K;
                const T* cf = static_cast<const T*>(tv.coef) + (size_t(t) * 2 + wh) * tv.K;
                for (int a = 0; a < nt; ++a) {
                    const T ca = sign * cf[a];
                    const T* b = src + size_t(sl[a]) * tv.n + size_t(x) * NN + c;
                    sfor<0, N>([&](auto I) { v[I] += ca * b[int(I) * N]; });
                }
            };
            if constexpr (DEV) {
                // which = 0: h = f - f0; which = 1: s = f + f0 (f terms of u, u0)
                gather(u, 0, T(1));
                gather(ug, 0, which ? T(1) : T(-1));
            } else {
                gather(which ? ug : u, which, T(1));
            }
        }
        sfor<0, N>([&](auto I) { P[int(I) * (N + 1) + c] = mk(v[I], T(0)); });
    }
    __syncthreads();
    lds_fft_row<T, N, -1>(P, c);     // z -> l3
    __syncthreads();
    lds_fft_col<T, N, -1>(P, c);     // y -> l2
    __syncthreads();
    constexpr int NT = PL * N;
    for (int k = threadIdx.x; k < PL * NN; k += NT) {
        const int q = k / NN, r = k - q * NN;
        const long pid = plane0 + q;
        if (pid >= nplanes) continue;
        const int l3 = r / N, l2 = r - l3 * N;
        Iw[pid * NN + r] = L[q * N * (N + 1) + l2 * (N + 1) + l3];   // [l3][l2]
    }
}

// A2: line (l3, l2) of task-which tw: transform over x, scale 1/n;
//     F^[t][which][l3][l1][l2]
template <typename T, int N>
__global__ void __launch_bounds__(256)
k_fwd_line(long nlines, int nfg, const cplx<T>* __restrict__ Iw, cplx<T>* __restrict__ Fhat) {
    constexpr int NN = N * N;
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= nlines) return;
    const long tw = gid / NN;
    const int r = int(gid - tw * NN);             // r = l3*N + l2
    const int t = int(tw / nfg), which = int(tw % nfg);
    const cplx<T>* src = Iw + tw * size_t(N) * NN + r;
    cplx<T> a[N], o[N];
    sfor<0, N>([&](auto X) { a[X] = src[size_t(X) * NN]; });
    RegDFT<T, N, -1>::template run<1, 0>(a, o);
    const T inv_n = T(1) / T(N * NN);
    const int l3 = r / N, l2 = r - l3 * N;
    cplx<T>* dst = Fhat + (size_t(t) * 2 + which) * size_t(N) * NN + size_t(l3) * NN + l2;
    sfor<0, N>([&](auto K1) { dst[size_t(K1) * N] = inv_n * o[K1]; });
}

// ================================================================ kernel B
// Nyquist planes.  f is real, so F^(-l) = conj F^(l) and a(-l) = -a(l) for
// every l off the Nyquist planes (no index equal to N/2): there
// u+-_h = IFFT3(e^{+-ia} F^) is real, and one packed transform
// w = IFFT3((1+i)(cos a + sin a) F^) = u+ + i u- carries both.  On the set
// {l1 = N/2} U {l2 = N/2} U {l3 = N/2} the symmetry fails; only the imaginary
// parts of that set's contribution are needed to unpack w exactly:
//   beta  = Im u+_nyq,  delta = Im u-_nyq,
//   u+ = (Re w + delta) + i beta,   u- = (Im w - beta) + i delta.
// The set splits into three disjoint planes, each a 2D transform times a
// (-1)^coordinate factor:
//   A: l1 = N/2                 -> (-1)^x A(y, z)
//   B: l2 = N/2, l1 != N/2      -> (-1)^y B(x, z)
//   C: l3 = N/2, l1, l2 != N/2  -> (-1)^z C(x, y)
// k_nyq writes Im of A+-, B+-, C+- for every (task, which, p, q):
// nyq[t][w][p][q][6][N*N] with planes A+ A- [y][z], B+ B- [z][x], C+ C- [y][x].
template <typename T, int N>
__global__ void __launch_bounds__(6 * N)
k_nyq(const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ phase, int M, int Nrho,
      int NW, T* __restrict__ nyq) {
    // one 2D plane per N-thread slot: slot = 2 type + sign; thread c walks k
    constexpr int NN = N * N, n = N * NN, PS = N * (N + 1), H = N / 2;
    __shared__ cplx<T> L[6 * PS];
    const int bid = blockIdx.x;
    const int q = bid % M, p = (bid / M) % Nrho, w = (bid / (M * Nrho)) % NW;
    const int t = bid / (M * Nrho * NW);
    const cplx<T>* F = Fhat + (size_t(t) * 2 + w) * n;
    const cplx<T>* ph = phase + (size_t(p) * M + q) * 3 * N;
    const int slot = threadIdx.x / N, c = threadIdx.x - slot * N;
    const int type = slot >> 1;
    const bool minus = (slot & 1) != 0;
    cplx<T>* P = L + slot * PS;
    // loads: rows a, columns b of the plane (A: a = l2, b = l3; B: a = l1,
    // b = l3; C: a = l1, b = l2); consecutive threads read consecutive l2
    // where the plane has an l2 axis (A, C)
    if (type == 0) {                       // thread c = l2, k = l3
        const cplx<T> e0 = ph[H] * ph[N + c];
        sfor<0, N>([&](auto K) {
            constexpr int k = K;
            const cplx<T> e = e0 * ph[2 * N + k];
            P[c * (N + 1) + k] = cmul_conj_if(F[(size_t(k) * N + H) * N + c], e, minus);
        });
    } else if (type == 1) {                // thread c = l1, k = l3
        const cplx<T> e0 = ph[c] * ph[N + H];
        const bool zero = c == H;
        sfor<0, N>([&](auto K) {
            constexpr int k = K;
            const cplx<T> e = e0 * ph[2 * N + k];
            const cplx<T> v = cmul_conj_if(F[(size_t(k) * N + c) * N + H], e, minus);
            P[c * (N + 1) + k] = zero ? mk(T(0), T(0)) : v;
        });
    } else {                               // thread c = l2, k = l1
        const cplx<T> e0 = ph[N + c] * ph[2 * N + H];
        sfor<0, N>([&](auto K) {
            constexpr int k = K;
            const cplx<T> e = ph[k] * e0;
            const cplx<T> v = cmul_conj_if(F[(size_t(H) * N + k) * N + c], e, minus);
            P[k * (N + 1) + c] = (k == H || c == H) ? mk(T(0), T(0)) : v;
        });
    }
    __syncthreads();
    lds_fft_row<T, N, +1>(P, c);          // over b for row c
    __syncthreads();
    lds_fft_col<T, N, +1>(P, c);          // over a for column c
    __syncthreads();
    // stores (coalesced): A[y][z] = P[y][z]; B[z][x] = P[x][z]; C[y][x] = P[x][y]
    T* dst = nyq + ((((size_t(t) * NW + w) * Nrho + p) * M + q) * 6 + slot) * NN;
    sfor<0, N>([&](auto K) {
        constexpr int k = K;
        dst[k * N + c] = (type == 0) ? P[k * (N + 1) + c].y : P[c * (N + 1) + k].y;
    });
}

// LDS image of one (which)'s Nyquist data for residue r, row pitch BP
template <int N, int S>
struct NyLayout {
    static constexpr int NS = N / S, BP = N + 4;
    static constexpr int A = 0;                   // [sgn][y][j]
    static constexpr int B = 2 * N * NS;          // [sgn][j][x]  (pitch BP)
    static constexpr int C = B + 2 * NS * BP;     // [sgn][y][x]  (pitch BP)
    static constexpr int SIZE = C + 2 * N * BP;
};

template <typename T, int N, int S>
constexpr int gain_lds_bytes(int NW) {
    return N * N * (N / S + 1) * 2 * int(sizeof(T)) + NW * NyLayout<N, S>::SIZE * int(sizeof(T));
}
// smallest residue split S (a divisor of N) whose LDS image fits a work-group
template <typename T, int N, int NW, int S = 1>
constexpr int pick_split_gain() {
    if constexpr (S > N) {
        static_assert(S <= N, "no residue split fits the gain kernel in LDS");
        return N;
    } else if constexpr (N % S == 0 && gain_lds_bytes<T, N, S>(NW) <= 160 * 1024) {
        return S;
    } else {
        return pick_split_gain<T, N, NW, S + 1>();
    }
}

// u+ and u- of one point (the reference state's cache entry)
template <typename T>
struct alignas(4 * sizeof(T)) UPair {
    cplx<T> p, m;
};

// the cache is streamed once per call: non-temporal accesses keep it from
// evicting the transforms the other work-groups of a task re-read from L2
template <typename T>
DGFS_DEV UPair<T> load_pair_nt(const UPair<T>* p) {
    const T* q = reinterpret_cast<const T*>(p);
    UPair<T> v;
    v.p.x = __builtin_nontemporal_load(q);
    v.p.y = __builtin_nontemporal_load(q + 1);
    v.m.x = __builtin_nontemporal_load(q + 2);
    v.m.y = __builtin_nontemporal_load(q + 3);
    return v;
}
template <typename T>
DGFS_DEV void store_pair_nt(UPair<T>* p, const UPair<T>& v) {
    T* q = reinterpret_cast<T*>(p);
    __builtin_nontemporal_store(v.p.x, q);
    __builtin_nontemporal_store(v.p.y, q + 1);
    __builtin_nontemporal_store(v.m.x, q + 2);
    __builtin_nontemporal_store(v.m.y, q + 3);
}

// Raw buffer views of device arrays.  A buffer access adds a per-lane VGPR
// byte offset and a uniform SGPR byte offset to the descriptor's base, so a
// stream with compile-time strides needs no 64-bit address arithmetic per
// access.  Word 3 of the gfx9 descriptor selects the 32-bit data format;
// accesses at or beyond `bytes` from the base read 0 and drop stores.
DGFS_DEV __amdgpu_buffer_rsrc_t buffer_view(const void* base, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, int(bytes),
                                             0x00020000);
}
constexpr int kBufNT = 2;   // cache-policy operand of a non-temporal access
// (the loaded vectors are reinterpreted whole: extracting their elements one
// by one has been seen to lose the upper dwords)
DGFS_DEV cplx<float> buf_load_c(__amdgpu_buffer_rsrc_t r, int voff, int soff, float) {
    return __builtin_bit_cast(cplx<float>, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}
DGFS_DEV cplx<double> buf_load_c(__amdgpu_buffer_rsrc_t r, int voff, int soff, double) {
    return __builtin_bit_cast(cplx<double>,
                              __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
DGFS_DEV UPair<float> buf_load_pair_nt(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
    return __builtin_bit_cast(UPair<float>,
                              __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, kBufNT));
}
// A pair is one 16-byte store with its whole offset in the VGPR and a zero
// SGPR offset.  The compiler's wait-state check for the store-data hazard of
// stores wider than 8 bytes skips buffer stores whose SGPR offset is a
// register: with one, the next VALU instruction overwrote the store's data
// VGPRs at once and the stored values differed from run to run.
DGFS_DEV void buf_store_pair_nt(__amdgpu_buffer_rsrc_t r, int voff, const UPair<float>& u) {
    using V4 = decltype(__builtin_amdgcn_raw_buffer_load_b128(r, 0, 0, 0));
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(V4, u), r, voff, 0, kBufNT);
}

enum GainMode { kGainAccumulate = 0, kGainStore = 1, kGainDelta = 2 };

// Gain: for (task, radius p, z-residue r, q-group), acc = sum_q t1 t2 + t3 t4
// with t1 t2 + t3 t4 = u+f u-f (NW = 1, f = g) or 1/2 (u+f u-g + u-f u+g).
// MODE kGainStore (NW = 1, F^ = F0^): write u+-_f0 of every (q, point) to
//   ucache[t][p][q][x][r][P] (P = this thread's pencil) instead.
// MODE kGainDelta (NW = 1, F^ = h^): acc = sum_q u+h (u-h + u-f0) + u-h u+f0
//   with u+-_f0 read from ucache.
template <typename T, int N, int S, int NW, int MODE = kGainAccumulate, int PAD = 1>
__global__ void __launch_bounds__(round64(N * (N / S)))
k_gain(const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ phase,
       const T* __restrict__ nyq, int M, int Nrho, int QG, cplx<T>* __restrict__ acc_out,
       UPair<T>* __restrict__ ucache, int remap) {
    static_assert(MODE == kGainAccumulate || NW == 1, "reference-state modes transform one function");
    using E = Fft3Engine<T, N, S, round64(N * (N / S)), +1, PAD>;
    using NL = NyLayout<N, S>;
    constexpr int n = N * N * N, NN = N * N;
    constexpr int NS = N / S, BP = NL::BP;
    __shared__ cplx<T> V[E::LDS_ELEMS];
    __shared__ T Ny[NW * NL::SIZE];
    const int nblk = gridDim.x;
    const int lb = remap ? xcd_remap(blockIdx.x, nblk) : int(blockIdx.x);
    const int qg = lb % QG;
    const int r = (lb / QG) % S;
    const int p = (lb / (QG * S)) % Nrho;
    const int t = lb / (QG * S * Nrho);
    typename E::Split sp;
    sp.init(r);

    const cplx<T>* Ff = Fhat + size_t(t) * 2 * n;
    const int qchunk = (M + QG - 1) / QG;
    const int q0 = qg * qchunk;
    const int q1 = (q0 + qchunk < M) ? q0 + qchunk : M;

    cplx<T> acc[N];
    [[maybe_unused]] cplx<T> R1[NW == 2 ? N : 1];
    sfor<0, N>([&](auto I) { acc[I] = mk(T(0), T(0)); });

    auto map1 = [](int pp, int& i1, int& i2) { i1 = pp / N; i2 = pp - (pp / N) * N; };  // (l1, l2)
    auto map3 = [](int P, int& o2, int& j) { o2 = P / NS; j = P - (P / NS) * NS; };      // (y, j)

    for (int q = q0; q < q1; ++q) {
        const cplx<T>* ph = phase + (size_t(p) * M + q) * 3 * N;   // [axis][l]
        sfor<0, NW>([&](auto Wi) {
            constexpr int wv = Wi;
            const cplx<T>* F = Ff + size_t(wv) * n;
            E::run(V, sp,
                [&] {
                    if constexpr (wv != 0) return;
                    // stage this q's Nyquist imaginary parts for residue r
                    for (int i = threadIdx.x; i < NW * 6 * N * NS + NW * 2 * NN;
                         i += E::NT) {
                        if (i < NW * 6 * N * NS) {
                            // A+- [y][j] and B+- [j][x] (both N*NS per sign)
                            const int w = i / (6 * N * NS), k = i - w * 6 * N * NS;
                            const int seg = k / (N * NS), e = k - seg * (N * NS);  // seg 0..5
                            if (seg >= 4) continue;
                            const T* src = nyq + ((((size_t(t) * NW + w) * Nrho + p) * M + q) * 6 +
                                                  seg) * NN;
                            T* dstw = Ny + w * NL::SIZE;
                            if (seg < 2) {       // A: e = y*NS + j
                                const int y = e / NS, j = e - y * NS;
                                dstw[NL::A + seg * N * NS + e] = src[y * N + S * j + r];
                            } else {             // B: e = j*N + x
                                const int j = e / N, x = e - j * N;
                                dstw[NL::B + (seg - 2) * NS * BP + j * BP + x] =
                                    src[(S * j + r) * N + x];
                            }
                        } else {
                            // C+- [y][x]
                            const int k0 = i - NW * 6 * N * NS;
                            const int w = k0 / (2 * NN), k = k0 - w * 2 * NN;
                            const int sg = k / NN, e = k - sg * NN;
                            const T* src = nyq + ((((size_t(t) * NW + w) * Nrho + p) * M + q) * 6 +
                                                  4 + sg) * NN;
                            Ny[w * NL::SIZE + NL::C + sg * N * BP + (e / N) * BP + e % N] = src[e];
                        }
                    }
                },
                map1,
                [&](int i1, int i2, int i3) {
                    const cplx<T> e = ph[i1] * ph[N + i2] * ph[2 * N + i3];
                    return (e.x + e.y) * F[(size_t(i3) * N + i1) * N + i2];
                },
                IdentityPre{}, map3, IdentityPre{},
                [&](int o2, int j, const cplx<T>* o) {
                    const int z = S * j + r;
                    const T sy = (o2 & 1) ? T(-1) : T(1), sz = (z & 1) ? T(-1) : T(1);
                    auto unpack = [&](int w, auto I, cplx<T> v, cplx<T>& up, cplx<T>& um) {
                        const T* nw = Ny + w * NL::SIZE;
                        constexpr T sx = (int(I) & 1) ? T(-1) : T(1);
                        const T beta = sx * nw[NL::A + o2 * NS + j] +
                                       sy * nw[NL::B + j * BP + int(I)] +
                                       sz * nw[NL::C + o2 * BP + int(I)];
                        const T delta = sx * nw[NL::A + N * NS + o2 * NS + j] +
                                        sy * nw[NL::B + NS * BP + j * BP + int(I)] +
                                        sz * nw[NL::C + N * BP + o2 * BP + int(I)];
                        // w = (1 + i) v = u+ + i u-
                        const T wr = v.x - v.y, wi = v.x + v.y;
                        up = mk(wr + delta, beta);
                        um = mk(wi - beta, delta);
                    };
                    if constexpr (MODE == kGainStore) {
                        UPair<T>* dst = ucache + ((((size_t(t) * Nrho + p) * M + q) * N) * S + r) *
                                                     size_t(E::NPEN) + threadIdx.x;
                        sfor<0, N>([&](auto I) {
                            UPair<T> v;
                            unpack(0, I, o[I], v.p, v.m);
                            store_pair_nt(dst + size_t(int(I)) * S * E::NPEN, v);
                        });
                    } else if constexpr (MODE == kGainDelta) {
                        const UPair<T>* src = ucache +
                            ((((size_t(t) * Nrho + p) * M + q) * N) * S + r) * size_t(E::NPEN) +
                            threadIdx.x;
                        // the stored pairs are read in chunks of CH points, all
                        // CH loads issued before the first is used
                        constexpr int CH = (N % 4 == 0) ? 4 : 2;
                        sfor<0, N / CH>([&](auto C) {
                            UPair<T> c0[CH];
                            sfor<0, CH>([&](auto k) {
                                c0[k] = load_pair_nt(src + size_t(int(C) * CH + int(k)) * S *
                                                               E::NPEN);
                            });
                            sfor<0, CH>([&](auto k) {
                                constexpr int i = int(C) * CH + int(k);
                                cplx<T> up, um;
                                unpack(0, std::integral_constant<int, i>{}, o[i], up, um);
                                acc[i] = acc[i] + up * (um + c0[k].m) + um * c0[k].p;
                            });
                        });
                    } else if constexpr (NW == 1) {
                        sfor<0, N>([&](auto I) {
                            cplx<T> up, um;
                            unpack(0, I, o[I], up, um);
                            acc[I] = acc[I] + up * um;
                        });
                    } else if constexpr (wv == 0) {
                        sfor<0, N>([&](auto I) { R1[I] = o[I]; });
                    } else {
                        sfor<0, N>([&](auto I) {
                            cplx<T> upf, umf, upg, umg;
                            unpack(0, I, R1[I], upf, umf);
                            unpack(1, I, o[I], upg, umg);
                            acc[I] = acc[I] + upf * umg + umf * upg;
                        });
                    }
                });
        });
    }
    if constexpr (MODE == kGainStore) return;
    if (threadIdx.x < E::NPEN) {
        int o2, j;
        map3(threadIdx.x, o2, j);
        const T scale = NW == 1 ? T(1) : T(0.5);
        // forward transform along x of Re acc for this thread's pencil (first
        // axis of FFT3(Re acc_p); only Re acc reaches the output, and its
        // transform is Hermitian, so the planes k1 <= N/2 carry it); the
        // (y, z) plane transforms follow in k_plane_weight
        cplx<T> tmp[N];
        sfor<0, N>([&](auto I) { tmp[I] = mk(scale * acc[I].x, T(0)); });
        RegDFT<T, N, -1>::template run<1, 0>(tmp, acc);
        // acc layout [t][p][qg][k1][r][y][j] with z = S*j + r, k1 <= N/2
        cplx<T>* dst = acc_out + ((size_t(t) * Nrho + p) * QG + qg) * n;
        sfor<0, N / 2 + 1>([&](auto I) {
            dst[((size_t(I) * S + r) * N + o2) * NS + j] = acc[I];
        });
    }
}

// k_gain on the lane-pair engine (Fft3Pair): the same work-group grid,
// products, acc layout and stored-pair layout as k_gain with one transform
// (f = g, or the reference-state modes), with 2 N (N / S) threads.  FP32 at
// N = 32 (S = 2, padded cube): 1024 threads of at most 128 VGPRs, four waves
// per SIMD instead of two.  FP64 at N = 32 (S = 4, unpadded 128 KiB cube):
// 512 threads of at most 256 VGPRs, two waves per SIMD instead of half a
// wave.  Lane h of pencil pair P holds the a1-outputs (and acc entries)
// o1 = (N/2) h + k.
constexpr int kPairDeltaChunk = 8;   // stored pairs in flight per lane (register budget)

template <typename T, int N, int S, int MODE = kGainAccumulate, int PAD = 1>
__global__ void __launch_bounds__(2 * N * (N / S))
k_gain_pair(const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ phase,
            const cplx<T>* __restrict__ phz, const T* __restrict__ nyq, int M, int Nrho, int QG,
            cplx<T>* __restrict__ acc_out, UPair<T>* __restrict__ ucache, int remap) {
    static_assert(MODE == kGainAccumulate || std::is_same<T, float>::value,
                  "the stored reference transforms are FP32");
    using E = Fft3Pair<T, N, S, +1, PAD>;
    using NL = NyLayout<N, S>;
    constexpr int n = N * N * N, NN = N * N;
    constexpr int NS = N / S, BP = NL::BP, H = E::H;
    __shared__ cplx<T> V[E::LDS_ELEMS];
    __shared__ T Ny[NL::SIZE];
    const int nblk = gridDim.x;
    const int lb = remap ? xcd_remap(blockIdx.x, nblk) : int(blockIdx.x);
    const int qg = lb % QG;
    const int r = (lb / QG) % S;
    const int p = (lb / (QG * S)) % Nrho;
    const int t = lb / (QG * S * Nrho);
    const int h = threadIdx.x & 1, P = threadIdx.x >> 1;

    const cplx<T>* Ff = Fhat + size_t(t) * 2 * n;
    const int qchunk = (M + QG - 1) / QG;
    const int q0 = qg * qchunk;
    const int q1 = (q0 + qchunk < M) ? q0 + qchunk : M;

    // Only Re acc reaches the output: the plane weights are real and even in
    // k (functions of |l|, the Nyquist index its own negative), so the
    // transform of i Im acc is Hermitian times i, weighted and inverted to a
    // purely imaginary field, and the line stage keeps the real part.
    T acc[H];
    sfor<0, H>([&](auto I) { acc[I] = T(0); });

    auto map1 = [](int pp, int& i1, int& i2) { i1 = pp / N; i2 = pp - (pp / N) * N; };  // (l1, l2)
    auto map3 = [](int Q, int& o2, int& j) { o2 = Q / NS; j = Q - (Q / NS) * NS; };      // (y, j)

    for (int q = q0; q < q1; ++q) {
        const cplx<T>* ph = phase + (size_t(p) * M + q) * 3 * N;   // [axis][l]
        const cplx<T>* pzq = phz + (size_t(p) * M + q) * N;          // [l3]
        // pass 1 reads F[l3][l1][l2] at the pencil's lane offset plus a
        // compile-time plane offset
        const auto Fb = buffer_view(Ff, unsigned(n * sizeof(cplx<T>)));
        auto prologue = [&] {
                // stage this q's Nyquist imaginary parts for residue r
                for (int i = threadIdx.x; i < 6 * N * NS + 2 * NN; i += E::NT) {
                    if (i < 6 * N * NS) {
                        // A+- [y][j] and B+- [j][x] (both N*NS per sign)
                        const int seg = i / (N * NS), e = i - seg * (N * NS);   // seg 0..5
                        if (seg >= 4) continue;
                        const T* src = nyq + (((size_t(t) * Nrho + p) * M + q) * 6 + seg) * NN;
                        if (seg < 2) {       // A: e = y*NS + j
                            const int y = e / NS, j = e - y * NS;
                            Ny[NL::A + seg * N * NS + e] = src[y * N + S * j + r];
                        } else {             // B: e = j*N + x
                            const int j = e / N, x = e - j * N;
                            Ny[NL::B + (seg - 2) * NS * BP + j * BP + x] =
                                src[(S * j + r) * N + x];
                        }
                    } else {
                        // C+- [y][x]
                        const int k = i - 6 * N * NS;
                        const int sg = k / NN, e = k - sg * NN;
                        const T* src = nyq + (((size_t(t) * Nrho + p) * M + q) * 6 + 4 + sg) * NN;
                        Ny[NL::C + sg * N * BP + (e / N) * BP + e % N] = src[e];
                    }
                }
            };
        auto load = [&](int i1, int i2, int i3) {
                // Re e + Im e of e = ph_x ph_y ph_z from the z factors' pair
                const cplx<T> a = ph[i1] * ph[N + i2], z = pzq[i3];
                const cplx<T> f = buf_load_c(Fb, (i1 * N + i2) * int(sizeof(cplx<T>)),
                                             i3 * NN * int(sizeof(cplx<T>)), T(0));
                return (a.x * z.x + a.y * z.y) * f;
            };
        auto epi = [&](int o2, int j, const cplx<T>* o) {
                const int z = S * j + r;
                const T sy = (o2 & 1) ? T(-1) : T(1), sz = (z & 1) ? T(-1) : T(1);
                // o[k] is the output at x = H h + k; H is even, so x and k
                // have the same parity
                auto unpack = [&](auto K, cplx<T> v, cplx<T>& up, cplx<T>& um) {
                    const int x = H * h + int(K);
                    constexpr T sx = (int(K) & 1) ? T(-1) : T(1);
                    const T beta = sx * Ny[NL::A + o2 * NS + j] + sy * Ny[NL::B + j * BP + x] +
                                   sz * Ny[NL::C + o2 * BP + x];
                    const T delta = sx * Ny[NL::A + N * NS + o2 * NS + j] +
                                    sy * Ny[NL::B + NS * BP + j * BP + x] +
                                    sz * Ny[NL::C + N * BP + o2 * BP + x];
                    // w = (1 + i) v = u+ + i u-
                    const T wr = v.x - v.y, wi = v.x + v.y;
                    up = mk(wr + delta, beta);
                    um = mk(wi - beta, delta);
                };
                // stored pairs [t][p][q][k][r][P][h] (x = H h + k): one
                // contiguous run of 2 NPEN pairs per (k, r), read by the
                // work-group's lanes in order; point k sits k S NT pairs
                // after the view's base
                constexpr int PB = int(sizeof(UPair<T>));
                const auto Ub = buffer_view(
                    ucache + ((((size_t(t) * Nrho + p) * M + q) * H) * S + r) * size_t(E::NT),
                    unsigned(H * S * E::NT * PB));
                const int lane_off = int(threadIdx.x) * PB;
                if constexpr (MODE == kGainStore) {
                    sfor<0, H>([&](auto K) {
                        UPair<T> v;
                        unpack(K, o[K], v.p, v.m);
                        buf_store_pair_nt(Ub, lane_off + int(K) * S * E::NT * PB, v);
                    });
                } else if constexpr (MODE == kGainDelta) {
                    constexpr int CH = kPairDeltaChunk;
                    sfor<0, H / CH>([&](auto C) {
                        UPair<T> c0[CH];
                        sfor<0, CH>([&](auto k) {
                            c0[k] = buf_load_pair_nt(Ub, lane_off,
                                                     (int(C) * CH + int(k)) * S * E::NT * PB);
                        });
                        sfor<0, CH>([&](auto k) {
                            constexpr int i = int(C) * CH + int(k);
                            cplx<T> up, um;
                            unpack(std::integral_constant<int, i>{}, o[i], up, um);
                            // Re[up (um + u0-) + um u0+]
                            acc[i] += up.x * (um.x + c0[k].m.x) - up.y * (um.y + c0[k].m.y) +
                                      um.x * c0[k].p.x - um.y * c0[k].p.y;
                        });
                        asm volatile("" ::: "memory");   // one chunk of loads in flight
                    });
                } else {
                    sfor<0, H>([&](auto K) {
                        cplx<T> up, um;
                        unpack(K, o[K], up, um);
                        acc[K] += up.x * um.x - up.y * um.y;             // Re[up um]
                    });
                }
            };
        // the residue r is uniform per work-group: one instantiation per r, so
        // that the engine's split twiddles are compile-time constants
        sfor<0, S>([&](auto Ri) {
            if (r == int(Ri)) E::template run<int(Ri)>(V, prologue, map1, load, map3, epi);
        });
    }
    if constexpr (MODE == kGainStore) return;
    int o2, j;
    map3(P, o2, j);
    // forward transform along x of the pair's pencil (first axis of
    // FFT3(acc_p)); lane h receives k1 = 2 m + h
    cplx<T> in[H], out[H];
    sfor<0, H>([&](auto K) { in[K] = mk(acc[K], T(0)); });
    pair_dft_dif<T, N, -1>(in, out, h);
    // acc layout [t][p][qg][k1][r][y][j] with z = S*j + r; the transform of
    // the real acc is Hermitian in k1, so only k1 <= N/2 is stored
    cplx<T>* dst = acc_out + ((size_t(t) * Nrho + p) * QG + qg) * n;
    sfor<0, H>([&](auto K) {
        const int k1 = 2 * int(K) + h;
        if (k1 <= N / 2) dst[((size_t(k1) * S + r) * N + o2) * NS + j] = out[K];
    });
}

// ================================================================ kernel C
// One work-group per plane (t, k1), k1 <= N/2: everything this kernel
// transforms is the transform of a real field (Re acc_p, f, h, f0), so
// planes N - k1 are the conjugates of planes k1 and are written from them.
// Its PL slots of N threads take radii
// p = slot, slot+PL, ...:  G^(k1,:,:) = sum_p wtab_p(k1,:,:) FFT2_{y,z}(sum_qg acc[t][p][qg][k1]),
// reduced over the slots through LDS; then Hg = IFFT2(G^(k1)) (slot 0) and
// Hl = IFFT2(bb2 G^_g(k1)) (slot 1).  H[t][0 = loss | 1 = gain][k1][y][z].
// DEV (reference-state form, F^ = h^, F0^ the reference state's): slot 1
// inverts bb2 f^ with f^ = h^ + F0^ and slot 2 inverts bb2 h^:
// H[t][0 = L[f] | 1 = gain | 2 = L[h]].
template <typename T, int N, bool DEV>
__global__ void __launch_bounds__(planes_per_wg(N) * N)
k_plane_weight(const cplx<T>* __restrict__ acc_in, const T* __restrict__ wtab,
               const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ Fhat0,
               const T* __restrict__ bb2, int fg_same, int Nrho, int QG, int SB,
               cplx<T>* __restrict__ H) {
    constexpr int PL = planes_per_wg(N), NT = PL * N, NN = N * N, n = N * NN, PS = N * (N + 1);
    constexpr int NH = DEV ? 3 : 2;
    static_assert(PL >= NH, "plane kernel needs a slot per output plane");
    __shared__ cplx<T> L[PL * PS];
    constexpr int NK = N / 2 + 1;                   // planes computed per task
    const int k1 = blockIdx.x % NK, t = blockIdx.x / NK;
    const int slot = threadIdx.x / N, c = threadIdx.x - slot * N;
    cplx<T>* P = L + slot * PS;
    const int NSB = N / SB;
    cplx<T> G[N];
    sfor<0, N>([&](auto I) { G[I] = mk(T(0), T(0)); });
    const int nround = (Nrho + PL - 1) / PL;
    for (int rd = 0; rd < nround; ++rd) {
        const int p = rd * PL + slot;
        const bool live = p < Nrho;
        __syncthreads();
        if (live) {
            const cplx<T>* src = acc_in + ((size_t(t) * Nrho + p) * QG) * n + size_t(k1) * NN;
            // the thread's N positions pos = c + N i, all loads of a q-group in
            // flight together
            cplx<T> v[N];
            sfor<0, N>([&](auto I) { v[I] = src[c + int(I) * N]; });
            for (int g = 1; g < QG; ++g)
                sfor<0, N>([&](auto I) { v[I] = v[I] + src[size_t(g) * n + c + int(I) * N]; });
            // stored position pos = (r N + y) NSB + j with z = SB j + r: j is the
            // thread's c mod NSB and r N + y = c / NSB + SB i
            const int j = c % NSB, rb = c / NSB;
            sfor<0, N>([&](auto I) {
                const int ry = rb + SB * int(I);
                const int r = ry / N, y = ry - r * N;
                P[y * (N + 1) + SB * j + r] = v[I];
            });
        }
        __syncthreads();
        if (live) lds_fft_row<T, N, -1>(P, c);     // z -> k3
        __syncthreads();
        if (live) {
            lds_fft_col<T, N, -1>(P, c);            // y -> k2 (column c = k3)
            const T* w = wtab + size_t(p) * n + size_t(k1) * NN + c;   // [k1][k2][k3 = c]
            sfor<0, N>([&](auto K2) { G[K2] = G[K2] + w[size_t(K2) * N] * P[K2 * (N + 1) + c]; });
        }
    }
    // reduce the slots' partial columns: slot 0 ends with G^(k1, :, c)
    __syncthreads();
    sfor<0, N>([&](auto K2) { P[K2 * (N + 1) + c] = G[K2]; });
    __syncthreads();
    if (slot == 0)
        for (int o = 1; o < PL; ++o)
            sfor<0, N>([&](auto K2) { G[K2] = G[K2] + L[o * PS + K2 * (N + 1) + c]; });
    __syncthreads();   // every slot's partial has been read before slot 1 reuses its plane
    if (slot == 0) {
        sfor<0, N>([&](auto K2) { P[K2 * (N + 1) + c] = G[K2]; });
    } else if (slot == 1) {
        // loss plane: bb2 * G^_g at l1 = k1, column c = l3
        const cplx<T>* Fg = Fhat + (size_t(t) * 2 + (fg_same ? 0 : 1)) * n;
        if constexpr (DEV) {
            const cplx<T>* Fh = Fhat + size_t(t) * 2 * n;
            const cplx<T>* F0 = Fhat0 + size_t(t) * 2 * n;
            sfor<0, N>([&](auto L2) {
                const size_t o = (size_t(c) * N + k1) * N + L2;
                P[L2 * (N + 1) + c] = bb2[size_t(k1) * NN + L2 * N + c] * (Fh[o] + F0[o]);
            });
        } else {
            sfor<0, N>([&](auto L2) {
                P[L2 * (N + 1) + c] = bb2[size_t(k1) * NN + L2 * N + c] *
                                      Fg[(size_t(c) * N + k1) * N + L2];
            });
        }
    } else if (DEV && slot == 2) {
        const cplx<T>* Fh = Fhat + size_t(t) * 2 * n;
        sfor<0, N>([&](auto L2) {
            P[L2 * (N + 1) + c] = bb2[size_t(k1) * NN + L2 * N + c] *
                                  Fh[(size_t(c) * N + k1) * N + L2];
        });
    }
    __syncthreads();
    if (slot < NH) lds_fft_col<T, N, +1>(P, c);      // k2 -> y
    __syncthreads();
    if (slot < NH) lds_fft_row<T, N, +1>(P, c);      // k3 -> z
    __syncthreads();
    // H(N - k1, y, z) = conj H(k1, y, z) for 0 < k1 < N/2
    const bool mirror = k1 > 0 && 2 * k1 < N;
    for (int i = threadIdx.x; i < NH * NN; i += NT) {
        const int which = i / NN, r = i - which * NN;     // slot 0 = gain, 1 = loss, 2 = L[h]
        const int dst = which == 0 ? 1 : (which == 1 ? 0 : 2);
        const cplx<T> v = L[which * PS + (r / N) * (N + 1) + r % N];
        H[((size_t(t) * NH + dst) * N + k1) * NN + r] = v;
        if (mirror) H[((size_t(t) * NH + dst) * N + (N - k1)) * NN + r] = conj(v);
    }
}

// ================================================================ kernel D
// Line (g, y, z): gain(x) = Re IFFT_k1(sum_{t in g} Hg[t](:,y,z)),
// loss(x) = sum_{t in g} f_t(x,y,z) Re IFFT_k1(Hl[t](:,y,z));
// Q[gslice[g]](x,y,z) += prefac (gain - loss).
// DEV (reference-state form, u0 the reference state, Q0 = Q(f0, f0) per
// output slice in FP64): loss = sum_t (h_t L[f_t] + f0_t L[h_t]) and
// Q[gslice[g]] += Q0[gslice[g]] + prefac (gain - loss), gain = G(h, s).
// LANES = 2 (N divisible by 4): a lane pair per line; lane h loads the k1 =
// 2 k + h inputs of each inverse (pair_dft_dit) and owns the points
// x = (N/2) h + k, so a task set's few lines keep twice the waves busy.
constexpr int line_lanes(int N) { return N % 4 == 0 ? 2 : 1; }
constexpr int kLineWG = 64;   // threads per work-group of k_line_output

template <typename T, int N, bool DEV>
__global__ void __launch_bounds__(kLineWG)
k_line_output(const T* __restrict__ u, const T* __restrict__ u0, const double* __restrict__ Q0,
              TaskView tv, const cplx<T>* __restrict__ H, long nlines,
              const int* __restrict__ gstart, const int* __restrict__ gslice, T prefac,
              T* __restrict__ Q) {
    constexpr int NN = N * N, n = N * NN, NH = DEV ? 3 : 2;
    constexpr int LN = line_lanes(N), PN = N / LN;
    const long tid = long(blockIdx.x) * kLineWG + threadIdx.x;
    const long gid = tid / LN;                    // both lanes of a pair share the line
    const int h = LN == 2 ? int(tid & 1) : 0;
    if (gid >= nlines) return;
    const int g = int(gid / NN), yz = int(gid - long(g) * NN);
    const int ta = gstart[g], tb = gstart[g + 1];
    const int x0 = PN * h;                        // this lane's points x = x0 + k
    auto kin = [&](int k) { return LN == 2 ? 2 * k + h : k; };   // input k1 of slot k
    // o = Re-usable inverse transform over k1 of a[] (slots as kin)
    auto inverse = [&](const cplx<T>* a, cplx<T>* o) {
        if constexpr (LN == 2)
            pair_dft_dit<T, N, +1>(a, o, h);
        else
            RegDFT<T, N, +1>::template run<1, 0>(a, o);
    };
    cplx<T> a[PN], o[PN];
    T loss[PN];
    sfor<0, PN>([&](auto I) { loss[I] = T(0); });
    // loss terms first, task by task (one inverse pencil live at a time),
    // then the gain planes summed over the group and inverted once
    for (int t = ta; t < tb; ++t) {
        const cplx<T>* hl = H + (size_t(t) * NH) * n + yz;
        const int nt = tv.nterm[2 * t];
        const int* sl = tv.slice + size_t(t) * 2 * tv.K;
        const T* cf = static_cast<const T*>(tv.coef) + size_t(t) * 2 * tv.K;
        if constexpr (DEV) {
            // the task's f terms in the outer loop and the line's points in
            // the inner one, so that no per-point value array stays live
            sfor<0, PN>([&](auto K) { a[K] = hl[2 * size_t(n) + size_t(kin(K)) * NN]; });
            inverse(a, o);                                          // L[h]: loss += f0 L[h]
            for (int k = 0; k < nt; ++k) {
                const T c = cf[k];
                const T* p0 = u0 + size_t(sl[k]) * n + size_t(x0) * NN + yz;
                sfor<0, PN>([&](auto X) { loss[X] += c * p0[size_t(X) * NN] * o[X].x; });
            }
            sfor<0, PN>([&](auto K) { a[K] = hl[size_t(kin(K)) * NN]; });
            inverse(a, o);                                          // L[f]: loss += h L[f]
            for (int k = 0; k < nt; ++k) {
                const T c = cf[k];
                const T* p = u + size_t(sl[k]) * n + size_t(x0) * NN + yz;
                const T* p0 = u0 + size_t(sl[k]) * n + size_t(x0) * NN + yz;
                sfor<0, PN>([&](auto X) {
                    loss[X] += c * (p[size_t(X) * NN] - p0[size_t(X) * NN]) * o[X].x;
                });
            }
        } else {
            sfor<0, PN>([&](auto K) { a[K] = hl[size_t(kin(K)) * NN]; });
            inverse(a, o);
            for (int k = 0; k < nt; ++k) {
                const T c = cf[k];
                const T* p = u + size_t(sl[k]) * n + size_t(x0) * NN + yz;
                sfor<0, PN>([&](auto X) { loss[X] += c * p[size_t(X) * NN] * o[X].x; });
            }
        }
    }
    sfor<0, PN>([&](auto I) { a[I] = mk(T(0), T(0)); });
    for (int t = ta; t < tb; ++t) {
        const cplx<T>* hg = H + (size_t(t) * NH + 1) * n + yz;
        sfor<0, PN>([&](auto K) { a[K] = a[K] + hg[size_t(kin(K)) * NN]; });
    }
    inverse(a, o);
    T* dst = Q + size_t(gslice[g]) * n + size_t(x0) * NN + yz;
    if constexpr (DEV) {
        const double* q0 = Q0 + size_t(gslice[g]) * n + size_t(x0) * NN + yz;
        // in chunks of CH points, so the FP64 loads of Q0 are not all live at once
        constexpr int CH = (PN % 8 == 0) ? 8 : ((PN % 4 == 0) ? 4 : 2);
        sfor<0, PN / CH>([&](auto C) {
            sfor<0, CH>([&](auto Xc) {
                constexpr int X = int(C) * CH + int(Xc);
                dst[size_t(X) * NN] += T(q0[size_t(X) * NN] +
                                         double(prefac * (o[X].x - loss[X])));
            });
            __builtin_amdgcn_sched_barrier(0);
        });
    } else {
        sfor<0, PN>([&](auto X) { dst[size_t(X) * NN] += prefac * (o[X].x - loss[X]); });
    }
}

// ================================================================ the plan
struct CollisionPlan {
    int N, M, Nrho, dtype, variant, ntask_max, ngroup_max;
    int SA, SB, SC, SD, QG;
    // gain schedule (autotuner knobs): residue split of the f = g gain kernel
    // (1: the smallest split whose LDS image fits, 2: twice that), and the
    // XCD-aware work-group order (1) or the dispatch order (0)
    int gain_smul = 1, gain_remap = 1;
    int QG_cap = 0;         // q groups the acc scratch holds
    double prefac;
    size_t n;
    void* phase = nullptr;  // [Nrho][M][3][N] cplx<T>
    void* phz = nullptr;    // [Nrho][M][N] cplx<T>: (cos + sin, cos - sin) of the z phase
    void* wtab = nullptr;   // [Nrho][k3][k1][k2] T
    void* bb2 = nullptr;    // [l3][l1][l2] T
    void* Fhat = nullptr;   // [ntask][2][n] cplx<T>
    void* acc = nullptr;    // [ntask][Nrho][QG][n] cplx<T>
    void* Iw = nullptr;     // [ntask][2][n] cplx<T>  forward-transform intermediate
    void* H = nullptr;      // [ntask][2][n] cplx<T>  plane-inverted loss / gain
    void* nyq = nullptr;    // [ntask][2][Nrho][M][6][N*N] T  Nyquist-plane imaginary parts
    // reference-state form: F0^ ([ntask][2][n] cplx<T>, slot 0 used) and
    // u+-_f0 ([ntask][Nrho][M][n] UPair<T>) of the reference state, valid for
    // ref_ntask tasks of the reference state at ref_u0 once set
    void* Fhat0 = nullptr;
    void* ucache = nullptr;
    int ref_cap = 0;        // tasks the reference buffers hold
    int ref_ntask = 0;
    const void* ref_u0 = nullptr;
    const int* ref_slice = nullptr;
    int num_cu = 256;
    // optional stage timing: 6 events per call (before forward, nyquist,
    // gain, plane, line, and after line)
    hipEvent_t* ev = nullptr;
    int ev_max = 0, ev_used = 0;
};

constexpr int kTimingEvents = 6;

inline void timing_mark(CollisionPlan* pl, int stage, hipStream_t st) {
    if (pl->ev && pl->ev_used < pl->ev_max)
        (void)hipEventRecord(pl->ev[pl->ev_used * kTimingEvents + stage], st);
}

// -------------------------------------------------------------- launch

template <typename T, int N>
struct Impl {
    static constexpr int S1 = pick_split_gain<T, N, 1>();   // gain residue split, f = g
    static constexpr int S2 = pick_split_gain<T, N, 2>();   // gain residue split, bilinear
    // the f = g gain's alternative schedule: twice the split (when it divides
    // N) with an unpadded LDS cube, so that two work-groups may share a CU
    static constexpr int S1x = (N % (2 * S1) == 0) ? 2 * S1 : S1;
    static constexpr int PL = planes_per_wg(N);
    // the f = g gain on the lane-pair engine: FP32 at N = 32 on the padded
    // cube of split S1 (1024 threads, four waves per SIMD); FP64 at N = 32 on
    // an unpadded cube of split 4 (512 threads, two waves per SIMD, where
    // k_gain at S1 = 8 runs half a wave per SIMD).  The reference-state modes
    // (FP32 only) keep S1.
    static constexpr bool kPair = std::is_same<T, float>::value && N == 32 &&
                                  2 * N * (N / S1) <= 1024;
    static constexpr bool kPair64 = std::is_same<T, double>::value && N == 32;
    static constexpr int SP = kPair64 ? 4 : S1;          // the f = g pair kernel's split
    static constexpr int PADP = kPair64 ? 0 : 1;
    static_assert(!kPair64 || N * N * (N / SP) * 16 + NyLayout<N, SP>::SIZE * 8 <= 160 * 1024,
                  "the FP64 pair kernel's LDS image exceeds a CU");

    // Q0 == nullptr: Q += Q(f, g).  Q0 != nullptr (reference-state form, ug the
    // reference state u0 whose F0^ and u+-_f0 the plan holds): Q += Q0 +
    // G(h) - h L[f] - f0 L[h], h = f - f0, which equals Q(f, f) when
    // Q0 = Q(f0, f0).
    static int launch(CollisionPlan* pl, const void* u, const void* ug, void* Q, TaskView tv,
                      int ntask, int ngroup, const int* gstart, const int* gslice, int fg_same,
                      hipStream_t st, const double* Q0) {
        const bool dev = Q0 != nullptr;
        if (dev) fg_same = 1;                  // one transform per (p, q): h
        const int nfg = fg_same ? 1 : 2;
        constexpr long NN = long(N) * N;
        timing_mark(pl, 0, st);
        if (int rc = forward(pl, u, ug, tv, ntask, nfg, dev, (cplx<T>*)pl->Fhat, st)) return rc;
        timing_mark(pl, 1, st);
        hipLaunchKernelGGL((k_nyq<T, N>), dim3(ntask * nfg * pl->Nrho * pl->M), dim3(6 * N), 0,
                           st, (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase, pl->M,
                           pl->Nrho, nfg, (T*)pl->nyq);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 2, st);
        const bool alt = fg_same && !dev && pl->gain_smul == 2 && S1x != S1;
        const bool pair = fg_same && !alt && (kPair || kPair64);
        const int S = alt ? S1x : (pair && !dev ? SP : (fg_same ? S1 : S2));
        const int nb = ntask * pl->Nrho * S * pl->QG;
        const int rm = pl->gain_remap;
        if (dev) {
            if constexpr (kPair)
                hipLaunchKernelGGL((k_gain_pair<T, N, S1, kGainDelta>), dim3(nb),
                                   dim3(2 * N * (N / S1)), 0, st, (const cplx<T>*)pl->Fhat,
                                   (const cplx<T>*)pl->phase, (const cplx<T>*)pl->phz,
                                   (const T*)pl->nyq, pl->M, pl->Nrho,
                                   pl->QG, (cplx<T>*)pl->acc, (UPair<T>*)pl->ucache, rm);
            else
                hipLaunchKernelGGL((k_gain<T, N, S1, 1, kGainDelta>), dim3(nb),
                                   dim3(round64(N * (N / S1))), 0, st, (const cplx<T>*)pl->Fhat,
                                   (const cplx<T>*)pl->phase, (const T*)pl->nyq, pl->M, pl->Nrho,
                                   pl->QG, (cplx<T>*)pl->acc, (UPair<T>*)pl->ucache, rm);
        } else if (pair) {
            if constexpr (kPair || kPair64)
                hipLaunchKernelGGL((k_gain_pair<T, N, SP, kGainAccumulate, PADP>), dim3(nb),
                                   dim3(2 * N * (N / SP)), 0, st, (const cplx<T>*)pl->Fhat,
                                   (const cplx<T>*)pl->phase, (const cplx<T>*)pl->phz,
                                   (const T*)pl->nyq, pl->M, pl->Nrho,
                                   pl->QG, (cplx<T>*)pl->acc, (UPair<T>*)nullptr, rm);
        }
        else if (alt)
            hipLaunchKernelGGL((k_gain<T, N, S1x, 1, kGainAccumulate, 0>), dim3(nb),
                               dim3(round64(N * (N / S1x))), 0,
                               st, (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase,
                               (const T*)pl->nyq, pl->M, pl->Nrho, pl->QG, (cplx<T>*)pl->acc,
                               (UPair<T>*)nullptr, rm);
        else if (fg_same)
            hipLaunchKernelGGL((k_gain<T, N, S1, 1>), dim3(nb), dim3(round64(N * (N / S1))), 0,
                               st, (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase,
                               (const T*)pl->nyq, pl->M, pl->Nrho, pl->QG, (cplx<T>*)pl->acc,
                               (UPair<T>*)nullptr, rm);
        else
            hipLaunchKernelGGL((k_gain<T, N, S2, 2>), dim3(nb), dim3(round64(N * (N / S2))), 0,
                               st, (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase,
                               (const T*)pl->nyq, pl->M, pl->Nrho, pl->QG, (cplx<T>*)pl->acc,
                               (UPair<T>*)nullptr, rm);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 3, st);
        if (dev)
            hipLaunchKernelGGL((k_plane_weight<T, N, true>), dim3(ntask * (N / 2 + 1)),
                               dim3(PL * N), 0, st,
                               (const cplx<T>*)pl->acc, (const T*)pl->wtab,
                               (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->Fhat0,
                               (const T*)pl->bb2, fg_same, pl->Nrho, pl->QG, S, (cplx<T>*)pl->H);
        else
            hipLaunchKernelGGL((k_plane_weight<T, N, false>), dim3(ntask * (N / 2 + 1)),
                               dim3(PL * N), 0,
                               st, (const cplx<T>*)pl->acc, (const T*)pl->wtab,
                               (const cplx<T>*)pl->Fhat, (const cplx<T>*)nullptr,
                               (const T*)pl->bb2, fg_same, pl->Nrho, pl->QG, S, (cplx<T>*)pl->H);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 4, st);
        const long nlo = long(ngroup) * NN;
        const long nlb = (nlo * line_lanes(N) + kLineWG - 1) / kLineWG;
        if (dev)
            hipLaunchKernelGGL((k_line_output<T, N, true>), dim3(nlb), dim3(kLineWG), 0, st,
                               (const T*)u, (const T*)ug, Q0, tv, (const cplx<T>*)pl->H, nlo,
                               gstart, gslice, T(pl->prefac), (T*)Q);
        else
            hipLaunchKernelGGL((k_line_output<T, N, false>), dim3(nlb), dim3(kLineWG), 0, st,
                               (const T*)u, (const T*)ug, Q0, tv, (const cplx<T>*)pl->H, nlo,
                               gstart, gslice, T(pl->prefac), (T*)Q);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 5, st);
        if (pl->ev && pl->ev_used < pl->ev_max) ++pl->ev_used;
        return 0;
    }

    // F^ of every task into dst ([t][2][n], nfg functions per task); dev: the
    // single function h = f - f0 (u = f, ug = f0)
    static int forward(CollisionPlan* pl, const void* u, const void* ug, TaskView tv, int ntask,
                       int nfg, bool dev, cplx<T>* dst, hipStream_t st) {
        constexpr long NN = long(N) * N;
        const long npl = long(ntask) * nfg * N;
        constexpr int FP = fwd_planes_per_wg(N);
        if (dev)
            hipLaunchKernelGGL((k_fwd_plane<T, N, true>), dim3((npl + FP - 1) / FP), dim3(FP * N),
                               0, st, (const T*)u, (const T*)ug, tv, nfg, npl, (cplx<T>*)pl->Iw);
        else
            hipLaunchKernelGGL((k_fwd_plane<T, N, false>), dim3((npl + FP - 1) / FP),
                               dim3(FP * N), 0, st, (const T*)u, (const T*)ug, tv, nfg, npl,
                               (cplx<T>*)pl->Iw);
        DGFS_HIP_CHECK(hipGetLastError());
        const long nli = long(ntask) * nfg * NN;
        hipLaunchKernelGGL((k_fwd_line<T, N>), dim3((nli + 255) / 256), dim3(256), 0, st, nli,
                           nfg, (const cplx<T>*)pl->Iw, dst);
        DGFS_HIP_CHECK(hipGetLastError());
        return 0;
    }

    // the reference state f0 (= task f-terms of u0): F0^ and u+-_f0 of every
    // (task, p, q) into the plan's reference buffers
    static int reference_set(CollisionPlan* pl, const void* u0, TaskView tv, int ntask,
                             hipStream_t st) {
        if (int rc = forward(pl, u0, u0, tv, ntask, 1, false, (cplx<T>*)pl->Fhat0, st)) return rc;
        hipLaunchKernelGGL((k_nyq<T, N>), dim3(ntask * pl->Nrho * pl->M), dim3(6 * N), 0, st,
                           (const cplx<T>*)pl->Fhat0, (const cplx<T>*)pl->phase, pl->M, pl->Nrho,
                           1, (T*)pl->nyq);
        DGFS_HIP_CHECK(hipGetLastError());
        const int nb = ntask * pl->Nrho * S1 * pl->QG;
        // the store pass lays the pairs out for the delta kernel that reads them
        if constexpr (kPair)
            hipLaunchKernelGGL((k_gain_pair<T, N, S1, kGainStore>), dim3(nb),
                               dim3(2 * N * (N / S1)), 0, st, (const cplx<T>*)pl->Fhat0,
                               (const cplx<T>*)pl->phase, (const cplx<T>*)pl->phz,
                               (const T*)pl->nyq, pl->M, pl->Nrho,
                               pl->QG, (cplx<T>*)nullptr, (UPair<T>*)pl->ucache, pl->gain_remap);
        else
            hipLaunchKernelGGL((k_gain<T, N, S1, 1, kGainStore>), dim3(nb),
                               dim3(round64(N * (N / S1))), 0, st, (const cplx<T>*)pl->Fhat0,
                               (const cplx<T>*)pl->phase, (const T*)pl->nyq, pl->M, pl->Nrho,
                               pl->QG, (cplx<T>*)nullptr, (UPair<T>*)pl->ucache, pl->gain_remap);
        DGFS_HIP_CHECK(hipGetLastError());
        return 0;
    }
};

template <typename T, int N>
int reference_set_launch(CollisionPlan* pl, const void* u0, TaskView tv, int ntask,
                         hipStream_t st) {
    return Impl<T, N>::reference_set(pl, u0, tv, ntask, st);
}

template <typename T, int N>
int collide_launch(CollisionPlan* pl, const void* u, const void* ug, void* Q, TaskView tv,
                   int ntask, int ngroup, const int* gs, const int* gsl, int fgs, hipStream_t st,
                   const double* Q0) {
    return Impl<T, N>::launch(pl, u, ug, Q, tv, ntask, ngroup, gs, gsl, fgs, st, Q0);
}

#define DGFS_COLLIDE_EXTERN(T, N)                                                        \
    extern template int collide_launch<T, N>(CollisionPlan*, const void*, const void*, void*, \
                                             TaskView, int, int, const int*, const int*, int, \
                                             hipStream_t, const double*);                     \
    extern template int reference_set_launch<T, N>(CollisionPlan*, const void*, TaskView, int, \
                                                   hipStream_t);
#define DGFS_COLLIDE_INST(T, N)                                                          \
    template int collide_launch<T, N>(CollisionPlan*, const void*, const void*, void*,     \
                                      TaskView, int, int, const int*, const int*, int,     \
                                      hipStream_t, const double*);                         \
    template int reference_set_launch<T, N>(CollisionPlan*, const void*, TaskView, int,    \
                                            hipStream_t);
#define DGFS_COLLIDE_SIZES(X, T) X(T, 4) X(T, 6) X(T, 8) X(T, 12) X(T, 16) X(T, 20) X(T, 24) X(T, 32)

DGFS_COLLIDE_SIZES(DGFS_COLLIDE_EXTERN, float)
DGFS_COLLIDE_EXTERN(float, 48)        // N = 48: FP32 only (FP64 plane buffers exceed LDS)
DGFS_COLLIDE_SIZES(DGFS_COLLIDE_EXTERN, double)

}  // namespace dgfs
