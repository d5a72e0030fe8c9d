// Fused fast-spectral VHS collision operator, batched over collision tasks.
//
// Reference behaviour (jaisw7/dgfs1D_gpu): DGFSVHSScatteringModelStd.fs
// (std/scattering.py:346-429) and DGFSVHSGLLScatteringModelStd.fs
// (:251-313) with kernels vhs-gll.mako:43-433.  The reference runs 17 launches
// and 5 batched B*N^3 cuFFT transforms per node, nodes one after another.
// Here one launch of each of four kernels covers every node of an RHS:
//
//   A  fwd   : F^ = FFT3(f)/n, G^ = FFT3(g)/n for every task (f, g are the
//              U/V contractions of the coefficient array, fused into the load)
//   B  gain  : for every (task, radius p, residue r, q-group):
//                acc_p(x) = sum_q 1/2 (u+f u-g + u-f u+g)(x),
//                u+-h = IFFT3(exp(+-i a_pq) H^)   (3 LDS passes, registers)
//              cos/sin products of the reference are the same quantity:
//              t1 t2 + t3 t4 = 1/2 (u+f u-g + u-f u+g).  The q-sum is taken
//              in physical space before the forward transform (linearity).
//   C  weight: Gp(k) = sum_{p in group} (2 w_p sw/n) bb1_p(k) FFT3(acc_p)(k)
//   D  out   : Q[out] += prefac (Re IFFT3(sum G) - f Re IFFT3(bb2 G^))
//
// Every kernel is an instance of Fft3Engine (fft3.h).  Tables (phases, weights,
// bb2) are built once in FP64 on the host at plan creation.
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

// XCD-aware block remap: consecutive logical ids land on one XCD (8 XCDs,
// round-robin dispatch), so the work-groups of one task share an L2.
DGFS_DEV int xcd_remap(int bid, int nblk) {
    const int per = nblk / 8;
    if (bid < per * 8) return (bid % 8) * per + bid / 8;
    return bid;
}

// ================================================================ kernel A
template <typename T, int N, int S>
__global__ void __launch_bounds__(round64(N * (N / S)))
k_forward(const T* __restrict__ u, const T* __restrict__ ug, TaskView tv, int nfg,
          cplx<T>* __restrict__ Fhat) {
    using E = Fft3Engine<T, N, S, round64(N * (N / S)), -1>;
    constexpr int n = N * N * N;
    __shared__ cplx<T> V[E::LDS_ELEMS];
    const int bid = blockIdx.x;
    const int r = bid % S;
    const int which = (bid / S) % nfg;
    const int t = bid / (S * nfg);
    typename E::Split sp;
    sp.init(r);
    const T inv_n = T(1) / T(n);
    cplx<T>* out = Fhat + (size_t(t) * 2 + which) * n;
    const T* src = which ? ug : u;
    // abstract axes: a3 = x (pass 1, residue on l1), a2 = z, a1 = y
    E::run(V, sp,
        [](int p, int& i1, int& i2) { i1 = p / N; i2 = p - (p / N) * N; },  // (y, z)
        [&](int i1, int i2, int i3) {
            const int v = (i3 * N + i1) * N + i2;   // f[x][y][z]
            return mk(task_value<T>(tv, src, t, which, v) * inv_n, T(0));
        },
        IdentityPre{},
        [](int P, int& o2, int& j) { o2 = P / (N / S); j = P - (P / (N / S)) * (N / S); },
        IdentityPre{},
        [&](int o2, int j, const cplx<T>* o) {
            // F^ layout [l3][l1][l2]: o2 = l3, l1 = S*j + r, o1 = l2
            cplx<T>* dst = out + (size_t(o2) * N + (S * j + r)) * N;
            sfor<0, N>([&](auto I) { dst[I] = o[I]; });
        });
}

// ================================================================ kernel B
template <typename T, int N, int S, bool FGSAME>
__global__ void __launch_bounds__(round64(N * (N / S)))
k_gain(const cplx<T>* __restrict__ Fhat, const cplx<T>* __restrict__ phase,
       int M, int Nrho, int QG, cplx<T>* __restrict__ acc_out) {
    using E = Fft3Engine<T, N, S, round64(N * (N / S)), +1>;
    constexpr int n = N * N * N;
    constexpr int NS = N / S;
    __shared__ cplx<T> V[E::LDS_ELEMS];
    const int nblk = gridDim.x;
    const int lb = xcd_remap(blockIdx.x, nblk);
    const int qg = lb % QG;
    const int r = (lb / QG) % S;
    const int p = (lb / (QG * S)) % Nrho;
    const int t = lb / (QG * S * Nrho);
    typename E::Split sp;
    sp.init(r);

    const cplx<T>* Ff = Fhat + size_t(t) * 2 * n;
    const cplx<T>* Fg = FGSAME ? Ff : Ff + n;
    const int qchunk = (M + QG - 1) / QG;
    const int q0 = qg * qchunk;
    const int q1 = (q0 + qchunk < M) ? q0 + qchunk : M;

    cplx<T> acc[N], R1[N];
    sfor<0, N>([&](auto I) { acc[I] = mk(T(0), T(0)); R1[I] = acc[I]; });

    auto map1 = [](int pp, int& i1, int& i2) { i1 = pp / N; i2 = pp - (pp / N) * N; };  // (l1, l2)
    auto map3 = [](int P, int& o2, int& j) { o2 = P / NS; j = P - (P / NS) * NS; };      // (y, j)

    for (int q = q0; q < q1; ++q) {
        const cplx<T>* ph = phase + (size_t(p) * M + q) * 3 * N;   // [axis][l]
        constexpr int NVAR = FGSAME ? 2 : 4;
        sfor<0, NVAR>([&](auto Vi) {
            constexpr int var = Vi;
            // variants: 0 (f,+) 1 (g,-) 2 (f,-) 3 (g,+); FGSAME: 0 (f,+) 1 (f,-)
            constexpr bool minus = (var == 1 || var == 2);
            const cplx<T>* F = (var == 1 || var == 3) ? Fg : Ff;
            E::run(V, sp, map1,
                [&](int i1, int i2, int i3) {
                    return cmul_conj_if(F[(size_t(i3) * N + i1) * N + i2], ph[2 * N + i3], minus);
                },
                [&](auto I, cplx<T> v) { return cmul_conj_if(v, ph[N + int(I)], minus); },
                map3,
                [&](auto I, cplx<T> v) { return cmul_conj_if(v, ph[int(I)], minus); },
                [&](int, int, const cplx<T>* o) {
                    if constexpr (var == 0 || var == 2) {
                        sfor<0, N>([&](auto I) { R1[I] = o[I]; });
                    } else {
                        sfor<0, N>([&](auto I) {
                            acc[I].x += R1[I].x * o[I].x - R1[I].y * o[I].y;
                            acc[I].y += R1[I].x * o[I].y + R1[I].y * o[I].x;
                        });
                    }
                });
        });
    }
    if (threadIdx.x < E::NPEN) {
        int o2, j;
        map3(threadIdx.x, o2, j);
        const T scale = FGSAME ? T(1) : T(0.5);
        // acc layout [t][p][qg][x][y][pos], pos = r*NS + j (z = S*j + r)
        cplx<T>* dst = acc_out + ((size_t(t) * Nrho + p) * QG + qg) * n;
        sfor<0, N>([&](auto I) {
            dst[(size_t(I) * N + o2) * N + r * NS + j] = scale * acc[I];
        });
    }
}

// ================================================================ kernel C
template <typename T, int N, int S, int SB>
__global__ void __launch_bounds__(round64(N * (N / S)))
k_weight(const cplx<T>* __restrict__ acc_in, const T* __restrict__ wtab,
         int Nrho, int QG, int PG, cplx<T>* __restrict__ Gp) {
    using E = Fft3Engine<T, N, S, round64(N * (N / S)), -1>;
    constexpr int n = N * N * N;
    constexpr int NS = N / S;
    constexpr int NSB = N / SB;     // gain-kernel residue layout of z
    __shared__ cplx<T> V[E::LDS_ELEMS];
    const int bid = blockIdx.x;
    const int r = bid % S;
    const int pg = (bid / S) % PG;
    const int t = bid / (S * PG);
    typename E::Split sp;
    sp.init(r);
    const int pchunk = (Nrho + PG - 1) / PG;
    const int p0 = pg * pchunk;
    const int p1 = (p0 + pchunk < Nrho) ? p0 + pchunk : Nrho;

    cplx<T> accG[N];
    sfor<0, N>([&](auto I) { accG[I] = mk(T(0), T(0)); });
    // pass-1 pencil (i1 = y, i2 = z), thread order follows the stored position
    auto map1 = [](int pp, int& i1, int& i2) {
        i1 = pp / N;
        const int pos = pp - (pp / N) * N;
        i2 = (pos % NSB) * SB + pos / NSB;
    };
    auto map3 = [](int P, int& o2, int& j) { o2 = P / NS; j = P - (P / NS) * NS; };  // (k3, j)
    for (int p = p0; p < p1; ++p) {
        const cplx<T>* src = acc_in + (size_t(t) * Nrho + p) * QG * n;
        const T* w = wtab + size_t(p) * n;
        E::run(V, sp, map1,
            [&](int i1, int i2, int i3) {
                // acc[x = i3][y = i1][pos(z = i2)]
                const size_t off = (size_t(i3) * N + i1) * N + (i2 % SB) * NSB + i2 / SB;
                cplx<T> s = src[off];
                for (int g = 1; g < QG; ++g) s = s + src[size_t(g) * n + off];
                return s;
            },
            IdentityPre{}, map3, IdentityPre{},
            [&](int o2, int j, const cplx<T>* o) {
                // k3 = o2, k1 = S*j + r, k2 = o1; weight layout [p][k3][k1][k2]
                const T* wr = w + (size_t(o2) * N + (S * j + r)) * N;
                sfor<0, N>([&](auto I) { accG[I] = accG[I] + wr[I] * o[I]; });
            });
    }
    if (threadIdx.x < E::NPEN) {
        int o2, j;
        map3(threadIdx.x, o2, j);
        cplx<T>* dst = Gp + ((size_t(t) * PG + pg) * N + o2) * N * N + size_t(S * j + r) * N;
        sfor<0, N>([&](auto I) { dst[I] = accG[I]; });
    }
}

// ================================================================ kernel D
template <typename T, int N, int S>
__global__ void __launch_bounds__(round64(N * (N / S)))
k_output(const T* __restrict__ u, TaskView tv, const cplx<T>* __restrict__ Fhat,
         const cplx<T>* __restrict__ Gp, const T* __restrict__ bb2, int PG, int fg_same,
         const int* __restrict__ gstart, const int* __restrict__ gslice, T prefac,
         T* __restrict__ Q) {
    using E = Fft3Engine<T, N, S, round64(N * (N / S)), +1>;
    constexpr int n = N * N * N;
    constexpr int NS = N / S;
    __shared__ cplx<T> V[E::LDS_ELEMS];
    const int bid = blockIdx.x;
    const int r = bid % S;
    const int g = bid / S;
    typename E::Split sp;
    sp.init(r);
    const int ta = gstart[g], tb = gstart[g + 1];

    auto map1 = [](int pp, int& i1, int& i2) { i1 = pp / N; i2 = pp - (pp / N) * N; };  // (l1, l2)
    auto map3 = [](int P, int& o2, int& j) { o2 = P / NS; j = P - (P / NS) * NS; };      // (y, j)

    T loss[N];
    sfor<0, N>([&](auto I) { loss[I] = T(0); });
    for (int t = ta; t < tb; ++t) {
        const cplx<T>* Fg = Fhat + (size_t(t) * 2 + (fg_same ? 0 : 1)) * n;
        E::run(V, sp, map1,
            [&](int i1, int i2, int i3) {
                const size_t off = (size_t(i3) * N + i1) * N + i2;   // [l3][l1][l2]
                return bb2[off] * Fg[off];
            },
            IdentityPre{}, map3, IdentityPre{},
            [&](int o2, int j, const cplx<T>* o) {
                const int z = S * j + r;
                sfor<0, N>([&](auto I) {
                    const int v = (int(I) * N + o2) * N + z;
                    loss[I] += task_value<T>(tv, u, t, 0, v) * o[I].x;
                });
            });
    }
    E::run(V, sp, map1,
        [&](int i1, int i2, int i3) {
            const size_t off = (size_t(i3) * N + i1) * N + i2;
            cplx<T> s = mk(T(0), T(0));
            for (int t = ta; t < tb; ++t)
                for (int pg = 0; pg < PG; ++pg) s = s + Gp[(size_t(t) * PG + pg) * n + off];
            return s;
        },
        IdentityPre{}, map3, IdentityPre{},
        [&](int o2, int j, const cplx<T>* o) {
            const int z = S * j + r;
            T* dst = Q + size_t(gslice[g]) * n;
            sfor<0, N>([&](auto I) {
                const int v = (int(I) * N + o2) * N + z;
                dst[v] += prefac * (o[I].x - loss[I]);
            });
        });
}

// ================================================================ the plan
struct CollisionPlan {
    int N, M, Nrho, dtype, variant, ntask_max, ngroup_max;
    int SA, SB, SC, SD, QG, PG;
    double prefac;
    size_t n;
    void* phase = nullptr;  // [Nrho][M][3][N] cplx<T>
    void* wtab = nullptr;   // [Nrho][k3][k1][k2] T
    void* bb2 = nullptr;    // [l3][l1][l2] T
    void* Fhat = nullptr;   // [ntask][2][n] cplx<T>
    void* acc = nullptr;    // [ntask][Nrho][QG][n] cplx<T>
    void* Gp = nullptr;     // [ntask][PG][n] cplx<T>
    int num_cu = 256;
    // optional stage timing: 5 events per call (before A, B, C, D, after D)
    hipEvent_t* ev = nullptr;
    int ev_max = 0, ev_used = 0;
};

inline void timing_mark(CollisionPlan* pl, int stage, hipStream_t st) {
    if (pl->ev && pl->ev_used < pl->ev_max)
        (void)hipEventRecord(pl->ev[pl->ev_used * 5 + stage], st);
}

// -------------------------------------------------------------- launch

template <typename T, int N>
struct Impl {
    static constexpr int S = pick_split<T>(N);
    static constexpr int NT = round64(N * (N / S));

    static int launch(CollisionPlan* pl, const void* u, const void* ug, void* Q, TaskView tv,
                      int ntask, int ngroup, const int* gstart, const int* gslice, int fg_same,
                      hipStream_t st) {
        const int nfg = fg_same ? 1 : 2;
        timing_mark(pl, 0, st);
        hipLaunchKernelGGL((k_forward<T, N, S>), dim3(ntask * nfg * S), dim3(NT), 0, st,
                           (const T*)u, (const T*)ug, tv, nfg, (cplx<T>*)pl->Fhat);
        DGFS_HIP_CHECK(hipGetLastError());
        const int nb = ntask * pl->Nrho * S * pl->QG;
        timing_mark(pl, 1, st);
        if (fg_same)
            hipLaunchKernelGGL((k_gain<T, N, S, true>), dim3(nb), dim3(NT), 0, st,
                               (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase, pl->M,
                               pl->Nrho, pl->QG, (cplx<T>*)pl->acc);
        else
            hipLaunchKernelGGL((k_gain<T, N, S, false>), dim3(nb), dim3(NT), 0, st,
                               (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->phase, pl->M,
                               pl->Nrho, pl->QG, (cplx<T>*)pl->acc);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 2, st);
        hipLaunchKernelGGL((k_weight<T, N, S, S>), dim3(ntask * pl->PG * S), dim3(NT), 0, st,
                           (const cplx<T>*)pl->acc, (const T*)pl->wtab, pl->Nrho, pl->QG,
                           pl->PG, (cplx<T>*)pl->Gp);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 3, st);
        hipLaunchKernelGGL((k_output<T, N, S>), dim3(ngroup * S), dim3(NT), 0, st,
                           (const T*)u, tv, (const cplx<T>*)pl->Fhat, (const cplx<T>*)pl->Gp,
                           (const T*)pl->bb2, pl->PG, fg_same, gstart, gslice,
                           T(pl->prefac), (T*)Q);
        DGFS_HIP_CHECK(hipGetLastError());
        timing_mark(pl, 4, st);
        if (pl->ev && pl->ev_used < pl->ev_max) ++pl->ev_used;
        return 0;
    }
};


template <typename T, int N>
int collide_launch(CollisionPlan* pl, const void* u, const void* ug, void* Q, TaskView tv,
                   int ntask, int ngroup, const int* gs, const int* gsl, int fgs, hipStream_t st) {
    return Impl<T, N>::launch(pl, u, ug, Q, tv, ntask, ngroup, gs, gsl, fgs, st);
}

#define DGFS_COLLIDE_EXTERN(T, N)                                                        \
    extern template int collide_launch<T, N>(CollisionPlan*, const void*, const void*, void*, \
                                             TaskView, int, int, const int*, const int*, int, \
                                             hipStream_t);
#define DGFS_COLLIDE_INST(T, N)                                                          \
    template int collide_launch<T, N>(CollisionPlan*, const void*, const void*, void*,     \
                                      TaskView, int, int, const int*, const int*, int,     \
                                      hipStream_t);
#define DGFS_COLLIDE_SIZES(X, T) X(T, 4) X(T, 6) X(T, 8) X(T, 12) X(T, 16) X(T, 20) X(T, 24) X(T, 32)

DGFS_COLLIDE_SIZES(DGFS_COLLIDE_EXTERN, float)
DGFS_COLLIDE_SIZES(DGFS_COLLIDE_EXTERN, double)

}  // namespace dgfs
