// Binary-mixture VSS collision plan (kernel tables, scratch) and its C ABI.
// The kernels live in bivss_kernels.h; each (dtype, N) instance is compiled
// in its own translation unit (bi_<dtype>_<N>.hip).
#include "bivss_kernels.h"

#include <cmath>
#include <cstdlib>
#include <vector>

namespace dgfs {

DGFS_DEV double l_of(int i, int N) { return i < N / 2 ? double(i) : double(i - N); }

// b_pq tables times the gain weight (vss.mako:50-93, bi/scattering.py:193-206):
// W[xy][k] = wx[x] sum_pre B[y][pre] exp(-i mpq fac_x szpre . l_k),
// wx = qw_x sw / n * swpre qz_x^(gamma+2).  FP64, one thread per (xy, k).
template <typename T>
__global__ void __launch_bounds__(256)
k_bi_btab(int N, int M, int NXY, int Mpre, const double* __restrict__ fac,
          const double* __restrict__ wx, const double* __restrict__ szpre,
          const double* __restrict__ Bt, double mpq, cplx<T>* __restrict__ wtab) {
    const long n = long(N) * N * N;
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= NXY * n) return;
    const int xy = int(gid / n), k = int(gid - long(xy) * n);
    const int x = xy / M, y = xy - x * M;
    const double l1 = l_of(k / (N * N), N), l2 = l_of((k / N) % N, N), l3 = l_of(k % N, N);
    double sr = 0.0, si = 0.0;
    for (int pp = 0; pp < Mpre; ++pp) {
        const double d = fac[x] * ((szpre[3 * pp] * l1) + (szpre[3 * pp + 1] * l2) +
                                   (szpre[3 * pp + 2] * l3));
        double s, c;
        sincos(mpq * d, &s, &c);
        const double B = Bt[size_t(y) * Mpre + pp];
        sr += B * c;
        si -= B * s;
    }
    wtab[gid] = mk(T(wx[x] * sr), T(wx[x] * si));
}

// c_pq (vss.mako:81-89): c[k] = sum_x wc[x] sum_pre Bsum[pre] exp(-i fac_x szpre . l_k),
// wc = qw_x sw swpre qz_x^(gamma+2), Bsum[pre] = sum_y B[y][pre].
template <typename T>
__global__ void __launch_bounds__(256)
k_bi_ctab(int N, int Nrho, int Mpre, const double* __restrict__ fac,
          const double* __restrict__ wc, const double* __restrict__ szpre,
          const double* __restrict__ Bsum, cplx<T>* __restrict__ ctab) {
    const long n = long(N) * N * N;
    const long k = long(blockIdx.x) * 256 + threadIdx.x;
    if (k >= n) return;
    const double l1 = l_of(int(k / (N * N)), N), l2 = l_of(int((k / N) % N), N),
                 l3 = l_of(int(k % N), N);
    double cr = 0.0, ci = 0.0;
    for (int x = 0; x < Nrho; ++x) {
        double sr = 0.0, si = 0.0;
        for (int pp = 0; pp < Mpre; ++pp) {
            const double d = fac[x] * ((szpre[3 * pp] * l1) + (szpre[3 * pp + 1] * l2) +
                                       (szpre[3 * pp + 2] * l3));
            double s, c;
            sincos(d, &s, &c);
            sr += Bsum[pp] * c;
            si -= Bsum[pp] * s;
        }
        cr += wc[x] * sr;
        ci += wc[x] * si;
    }
    ctab[k] = mk(T(cr), T(ci));
}

template <typename T>
static int bi_dispatch(BiPlan* pl, const void* up, const void* uq, void* Q, TaskView tv,
                       int ntask, int ngroup, const int* gs, const int* gsl, hipStream_t st) {
    switch (pl->N) {
        case 4: return bi_launch<T, 4>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 6: return bi_launch<T, 6>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 8: return bi_launch<T, 8>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 12: return bi_launch<T, 12>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 16: return bi_launch<T, 16>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 20: return bi_launch<T, 20>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 24: return bi_launch<T, 24>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 32: return bi_launch<T, 32>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
        case 48:
            if constexpr (sizeof(T) == 4)
                return bi_launch<T, 48>(pl, up, uq, Q, tv, ntask, ngroup, gs, gsl, st);
            break;
        default: break;
    }
    set_error("bi_collide: unsupported velocity grid N=%d", pl->N);
    return 2;
}

static void bi_free(BiPlan* pl) {
    for (void* p : {pl->phase, pl->wtab, pl->ctab, pl->Fhat, pl->Iw, pl->acc, pl->H, pl->Gp})
        if (p) (void)hipFree(p);
    delete pl;
}

// largest number of xy groups of the plane stage (bi_xgroups_max<N>)
static int bi_gp_groups(int N, int NXY) {
    switch (N) {
        case 4: return bi_xgroups_max<4>(NXY);
        case 6: return bi_xgroups_max<6>(NXY);
        case 8: return bi_xgroups_max<8>(NXY);
        case 12: return bi_xgroups_max<12>(NXY);
        case 16: return bi_xgroups_max<16>(NXY);
        case 20: return bi_xgroups_max<20>(NXY);
        case 24: return bi_xgroups_max<24>(NXY);
        case 32: return bi_xgroups_max<32>(NXY);
        default: return bi_xgroups_max<48>(NXY);
    }
}

// scratch budget of the gain products: tasks are processed in chunks whose
// [chunk][Nrho*M][n] complex products fit in it
constexpr size_t kBiAccBudget = size_t(2) << 30;

}  // namespace dgfs

using dgfs::BiPlan;

extern "C" int dgfs_bi_collision_plan_create(int N, int M, int Nrho, double L, const double* qz,
                                             const double* qw, const double* sz, double sw,
                                             int Mpre, const double* szpre, double mp, double mq,
                                             double alpha, double omega, double prefac,
                                             int dtype, int ntask_max,
                                             dgfs_bi_collision_plan_t* out) {
    DGFS_REQUIRE(out != nullptr, "bi_collision_plan_create: null output handle");
    *out = nullptr;
    DGFS_REQUIRE(N == 4 || N == 6 || N == 8 || N == 12 || N == 16 || N == 20 || N == 24 ||
                     N == 32 || N == 48,
                 "bi_collision_plan_create: unsupported N=%d", N);
    DGFS_REQUIRE(N != 48 || dtype == DGFS_F32,
                 "bi_collision_plan_create: N=48 is built for DGFS_F32 only (its FP64 plane "
                 "buffers exceed the 160 KiB LDS of a work-group)");
    DGFS_REQUIRE(M > 0 && Nrho > 0 && Mpre > 0 && ntask_max > 0,
                 "bi_collision_plan_create: bad sizes");
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "bi_collision_plan_create: bad dtype");
    DGFS_REQUIRE(qz && qw && sz && szpre, "bi_collision_plan_create: null table");
    DGFS_REQUIRE(mp > 0 && mq > 0, "bi_collision_plan_create: masses must be positive");
    const double eta = alpha - 1.0, gam = 2.0 * (1.0 - omega);
    const size_t n = size_t(N) * N * N;
    const int NXY = Nrho * M;
    const size_t es = dtype == DGFS_F64 ? 8 : 4, cs = 2 * es;

    // ---- host parts of the tables (own code; formulas of vss.mako:29-93)
    std::vector<double> Bt(size_t(M) * Mpre), Bsum(Mpre, 0.0);
    for (int y = 0; y < M; ++y)
        for (int pp = 0; pp < Mpre; ++pp) {
            const double base = 1.0 + sz[3 * y] * szpre[3 * pp] + sz[3 * y + 1] * szpre[3 * pp + 1] +
                                sz[3 * y + 2] * szpre[3 * pp + 2];
            // an exactly antipodal pair (1 + s.s' = 0) can round below zero,
            // where pow with a fractional exponent is NaN: clamp at its exact value
            Bt[size_t(y) * Mpre + pp] = std::pow(base > 0.0 ? base : 0.0, eta);
        }
    for (int pp = 0; pp < Mpre; ++pp)
        for (int y = 0; y < M; ++y) Bsum[pp] += Bt[size_t(y) * Mpre + pp];
    const double swpre = 4.0 * M_PI / Mpre;
    std::vector<double> fac(Nrho), wx(Nrho), wc(Nrho);
    for (int x = 0; x < Nrho; ++x) {
        fac[x] = M_PI / L * qz[x];
        const double fb = swpre * std::pow(qz[x], gam + 2.0);
        wx[x] = qw[x] * sw / double(n) * fb;
        wc[x] = qw[x] * sw * fb;
    }
    const double mqf = mq / (mp + mq), mpf = mp / (mp + mq);
    std::vector<double> ph(size_t(NXY) * 2 * 3 * N * 2);
    for (int x = 0; x < Nrho; ++x)
        for (int y = 0; y < M; ++y)
            for (int side = 0; side < 2; ++side)
                for (int ax = 0; ax < 3; ++ax)
                    for (int i = 0; i < N; ++i) {
                        const double l = i < N / 2 ? double(i) : double(i - N);
                        const double th = (side == 0 ? mqf : -mpf) * fac[x] * sz[3 * y + ax] * l;
                        const size_t o = ((((size_t(x) * M + y) * 2 + side) * 3 + ax) * N + i) * 2;
                        ph[o] = std::cos(th);
                        ph[o + 1] = std::sin(th);
                    }

    auto* pl = new BiPlan();
    pl->N = N; pl->M = M; pl->Nrho = Nrho; pl->Mpre = Mpre; pl->dtype = dtype;
    pl->ntask_max = ntask_max; pl->prefac = prefac; pl->mp = mp; pl->mq = mq; pl->n = n;
    const bool slabs = (dtype == DGFS_F64 ? dgfs::pick_split<double>(N)
                                          : dgfs::pick_split<float>(N)) > 1;
    const size_t per_task = size_t(NXY) * n * cs * (slabs ? 2 : 1);
    size_t budget = dgfs::kBiAccBudget;
    if (const char* mb = std::getenv("DGFS_BI_SCRATCH_MB")) {
        const long v = std::atol(mb);
        if (v > 0) budget = size_t(v) << 20;
    }
    size_t chunk = budget / per_task;
    if (chunk < 1) chunk = 1;
    if (chunk > size_t(ntask_max)) chunk = size_t(ntask_max);
    pl->chunk = int(chunk);

    void *d_fac = nullptr, *d_wx = nullptr, *d_wc = nullptr, *d_sp = nullptr, *d_B = nullptr,
         *d_Bs = nullptr;
    auto fail = [&](int rc) {
        for (void* p : {d_fac, d_wx, d_wc, d_sp, d_B, d_Bs})
            if (p) (void)hipFree(p);
        dgfs::bi_free(pl);
        return rc;
    };
    auto up64 = [&](void** d, const double* h, size_t cnt) -> bool {
        return hipMalloc(d, cnt * 8) == hipSuccess &&
               hipMemcpy(*d, h, cnt * 8, hipMemcpyHostToDevice) == hipSuccess;
    };
    if (!up64(&d_fac, fac.data(), Nrho) || !up64(&d_wx, wx.data(), Nrho) ||
        !up64(&d_wc, wc.data(), Nrho) || !up64(&d_sp, szpre, size_t(Mpre) * 3) ||
        !up64(&d_B, Bt.data(), Bt.size()) || !up64(&d_Bs, Bsum.data(), Mpre)) {
        dgfs::set_error("bi_collision_plan_create: table upload failed");
        return fail(1);
    }
    if (hipMalloc(&pl->phase, ph.size() / 2 * cs) != hipSuccess ||
        hipMalloc(&pl->wtab, size_t(NXY) * n * cs) != hipSuccess ||
        hipMalloc(&pl->ctab, n * cs) != hipSuccess ||
        hipMalloc(&pl->Fhat, chunk * 2 * n * cs) != hipSuccess ||
        hipMalloc(&pl->Iw, chunk * 2 * n * cs) != hipSuccess ||
        hipMalloc(&pl->acc, chunk * per_task) != hipSuccess ||
        hipMalloc(&pl->H, size_t(ntask_max) * 2 * n * cs) != hipSuccess ||
        hipMalloc(&pl->Gp, chunk * size_t(dgfs::bi_gp_groups(N, NXY)) * n * cs) != hipSuccess) {
        dgfs::set_error("bi_collision_plan_create: allocation failed");
        return fail(1);
    }
    if (es == 8) {
        if (hipMemcpy(pl->phase, ph.data(), ph.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
            dgfs::set_error("bi_collision_plan_create: phase upload failed");
            return fail(1);
        }
    } else {
        std::vector<float> f(ph.begin(), ph.end());
        if (hipMemcpy(pl->phase, f.data(), f.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
            dgfs::set_error("bi_collision_plan_create: phase upload failed");
            return fail(1);
        }
    }
    const long nb = (long(NXY) * long(n) + 255) / 256, nc = (long(n) + 255) / 256;
    if (es == 8) {
        hipLaunchKernelGGL(dgfs::k_bi_btab<double>, dim3(nb), dim3(256), 0, 0, N, M, NXY, Mpre,
                           (const double*)d_fac, (const double*)d_wx, (const double*)d_sp,
                           (const double*)d_B, mqf, (dgfs::cplx<double>*)pl->wtab);
        hipLaunchKernelGGL(dgfs::k_bi_ctab<double>, dim3(nc), dim3(256), 0, 0, N, Nrho, Mpre,
                           (const double*)d_fac, (const double*)d_wc, (const double*)d_sp,
                           (const double*)d_Bs, (dgfs::cplx<double>*)pl->ctab);
    } else {
        hipLaunchKernelGGL(dgfs::k_bi_btab<float>, dim3(nb), dim3(256), 0, 0, N, M, NXY, Mpre,
                           (const double*)d_fac, (const double*)d_wx, (const double*)d_sp,
                           (const double*)d_B, mqf, (dgfs::cplx<float>*)pl->wtab);
        hipLaunchKernelGGL(dgfs::k_bi_ctab<float>, dim3(nc), dim3(256), 0, 0, N, Nrho, Mpre,
                           (const double*)d_fac, (const double*)d_wc, (const double*)d_sp,
                           (const double*)d_Bs, (dgfs::cplx<float>*)pl->ctab);
    }
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) {
        dgfs::set_error("bi_collision_plan_create: table kernels failed");
        return fail(1);
    }
    for (void* p : {d_fac, d_wx, d_wc, d_sp, d_B, d_Bs}) (void)hipFree(p);
    *out = pl;
    return 0;
}

extern "C" int dgfs_bi_collision_plan_query(dgfs_bi_collision_plan_t plan, int* split,
                                            int* chunk, size_t* scratch_bytes) {
    auto* pl = static_cast<BiPlan*>(plan);
    DGFS_REQUIRE(pl, "bi_collision_plan_query: null plan");
    const size_t cs = pl->dtype == DGFS_F64 ? 16 : 8;
    if (split) *split = pl->dtype == DGFS_F64 ? dgfs::pick_split<double>(pl->N)
                                              : dgfs::pick_split<float>(pl->N);
    if (chunk) *chunk = pl->chunk;
    if (scratch_bytes) {
        const int S = pl->dtype == DGFS_F64 ? dgfs::pick_split<double>(pl->N)
                                            : dgfs::pick_split<float>(pl->N);
        const size_t acc = size_t(pl->Nrho) * pl->M * (S > 1 ? 2 : 1);
        *scratch_bytes = (size_t(pl->chunk) * (4 + acc) + size_t(pl->ntask_max) * 2) * pl->n * cs;
    }
    return 0;
}

extern "C" int dgfs_bi_collision_plan_destroy(dgfs_bi_collision_plan_t plan) {
    if (plan) dgfs::bi_free(static_cast<BiPlan*>(plan));
    return 0;
}

extern "C" int dgfs_bi_collide(dgfs_bi_collision_plan_t plan, const void* up, const void* uq,
                               void* Q, int K, int Ne, int ntask, const int* task_nterm,
                               const int* task_slice, const void* task_coef, int ngroup,
                               const int* group_start, const int* group_slice,
                               hipStream_t stream) {
    auto* pl = static_cast<BiPlan*>(plan);
    DGFS_REQUIRE(pl, "bi_collide: null plan");
    DGFS_REQUIRE(ntask >= 0 && ntask <= pl->ntask_max, "bi_collide: ntask=%d exceeds plan max %d",
                 ntask, pl->ntask_max);
    DGFS_REQUIRE(ngroup >= 0 && ngroup <= ntask, "bi_collide: bad ngroup=%d", ngroup);
    DGFS_REQUIRE(K > 0 && Ne > 0, "bi_collide: bad K/Ne");
    if (ntask == 0) return 0;
    DGFS_REQUIRE(up && uq && Q && task_nterm && task_slice && task_coef && group_start &&
                     group_slice,
                 "bi_collide: null pointer argument");
    dgfs::TaskView tv{K, Ne, int(pl->n), task_nterm, task_slice, task_coef};
    if (pl->dtype == DGFS_F64)
        return dgfs::bi_dispatch<double>(pl, up, uq, Q, tv, ntask, ngroup, group_start,
                                         group_slice, stream);
    return dgfs::bi_dispatch<float>(pl, up, uq, Q, tv, ntask, ngroup, group_start, group_slice,
                                    stream);
}
