// Collision plan (tables, scratch) and the C ABI entry points.  The kernels
// live in collision_kernels.h; each (dtype, N) instance is compiled in its own
// translation unit (coll_<dtype>_<N>.hip) so the library builds in parallel.
#include "collision_kernels.h"

#include <cmath>
#include <type_traits>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

namespace dgfs {

static thread_local std::string g_err;
void set_error(const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}
const char* last_error() { return g_err.c_str(); }

template <typename T>
static int dispatch(CollisionPlan* pl, const void* u, const void* ug, void* Q, TaskView tv,
                    int ntask, int ngroup, const int* gs, const int* gsl, int fgs,
                    hipStream_t st, const double* Q0) {
#define DGFS_CASE(NN)                                                                     \
    case NN:                                                                              \
        return collide_launch<T, NN>(pl, u, ug, Q, tv, ntask, ngroup, gs, gsl, fgs, st, Q0);
    switch (pl->N) {
        DGFS_CASE(4) DGFS_CASE(6) DGFS_CASE(8) DGFS_CASE(12) DGFS_CASE(16) DGFS_CASE(20)
        DGFS_CASE(24) DGFS_CASE(32)
        case 48:
            if constexpr (std::is_same<T, float>::value)
                return collide_launch<T, 48>(pl, u, ug, Q, tv, ntask, ngroup, gs, gsl, fgs, st,
                                             Q0);
            break;
        default: break;
    }
#undef DGFS_CASE
    set_error("collision: unsupported velocity grid N=%d", pl->N);
    return 2;
}

template <typename T>
static int dispatch_reference_set(CollisionPlan* pl, const void* u0, TaskView tv, int ntask,
                                  hipStream_t st) {
#define DGFS_CASE(NN)                                                                     \
    case NN:                                                                              \
        return reference_set_launch<T, NN>(pl, u0, tv, ntask, st);
    switch (pl->N) {
        DGFS_CASE(4) DGFS_CASE(6) DGFS_CASE(8) DGFS_CASE(12) DGFS_CASE(16) DGFS_CASE(20)
        DGFS_CASE(24) DGFS_CASE(32)
        case 48:
            if constexpr (std::is_same<T, float>::value)
                return reference_set_launch<T, 48>(pl, u0, tv, ntask, st);
            break;
        default: break;
    }
#undef DGFS_CASE
    set_error("collision: unsupported velocity grid N=%d", pl->N);
    return 2;
}

static bool supported_N(int N) {
    return N == 4 || N == 6 || N == 8 || N == 12 || N == 16 || N == 20 || N == 24 || N == 32 ||
           N == 48;
}

template <typename T>
static int split_for(int N) { return pick_split<T>(N); }

}  // namespace dgfs

// ================================================================ C ABI
using dgfs::CollisionPlan;

extern "C" const char* dgfs_last_error(void) { return dgfs::last_error(); }

extern "C" int dgfs_collision_plan_create(int N, int M, int Nrho, double L, const double* qz,
                                          const double* qw, const double* sz, double sw,
                                          double gamma, double prefac, int dtype, int variant,
                                          int ntask_max, dgfs_collision_plan_t* out) {
    DGFS_REQUIRE(out != nullptr, "collision_plan_create: null output handle");
    *out = nullptr;
    DGFS_REQUIRE(dgfs::supported_N(N), "collision_plan_create: unsupported N=%d", N);
    DGFS_REQUIRE(N != 48 || dtype == DGFS_F32,
                 "collision_plan_create: N=48 is built for DGFS_F32 only (its FP64 plane "
                 "buffers exceed a work-group's 160 KiB of LDS)");
    DGFS_REQUIRE(M > 0 && Nrho > 0 && ntask_max > 0, "collision_plan_create: bad sizes");
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "collision_plan_create: bad dtype");
    DGFS_REQUIRE(variant == DGFS_VHS || variant == DGFS_VHS_GLL,
                 "collision_plan_create: bad variant");
    DGFS_REQUIRE(qz && qw && sz, "collision_plan_create: null table");

    auto* pl = new CollisionPlan();
    pl->N = N; pl->M = M; pl->Nrho = Nrho; pl->dtype = dtype; pl->variant = variant;
    pl->ntask_max = ntask_max; pl->prefac = prefac;
    const size_t n = size_t(N) * N * N;
    pl->n = n;
    const int S = dtype == DGFS_F64 ? dgfs::split_for<double>(N) : dgfs::split_for<float>(N);
    // (the gain kernel's own split is a compile-time choice per (dtype, N, f==g);
    //  S here sizes the q-groups)
    pl->SB = S;
    pl->SA = pl->SC = pl->SD = dgfs::planes_per_wg(N);
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) == hipSuccess) pl->num_cu = prop.multiProcessorCount;
    // q-groups: enough gain work-groups to give every CU ~3 of them
    {
        const long base = long(ntask_max) * Nrho * S;
        int QG = 1;
        while (QG < M && base * QG < 3L * pl->num_cu) ++QG;
        pl->QG = QG;
    }

    // ---- host tables in FP64 (vhs-gll.mako:43-99 restated with own code)
    std::vector<double> l0(N);
    for (int i = 0; i < N; ++i) l0[i] = (i < N / 2) ? double(i) : double(i - N);
    std::vector<double> ph(size_t(Nrho) * M * 3 * N * 2);
    for (int p = 0; p < Nrho; ++p) {
        const double fac = M_PI / L * qz[p] / 2.0;
        for (int q = 0; q < M; ++q)
            for (int ax = 0; ax < 3; ++ax)
                for (int i = 0; i < N; ++i) {
                    const double a = fac * sz[q * 3 + ax] * l0[i];
                    const size_t o = (((size_t(p) * M + q) * 3 + ax) * N + i) * 2;
                    ph[o] = std::cos(a);
                    ph[o + 1] = std::sin(a);
                }
    }
    auto sinc = [](double s) { return std::sin(s + 1e-15) / (s + 1e-15); };
    std::vector<double> wt(size_t(Nrho) * n), b2(n, 0.0);
    for (int i1 = 0; i1 < N; ++i1)
        for (int i2 = 0; i2 < N; ++i2)
            for (int i3 = 0; i3 < N; ++i3) {
                const double ln = std::sqrt(l0[i1] * l0[i1] + l0[i2] * l0[i2] + l0[i3] * l0[i3]);
                const size_t o = (size_t(i1) * N + i2) * N + i3;   // natural [l1][l2][l3]
                double bsum = 0.0;
                for (int p = 0; p < Nrho; ++p) {
                    const double zg = std::pow(qz[p], gamma + 2.0);
                    const double bb1 = zg * 4.0 * M_PI * sinc(M_PI / L * qz[p] / 2.0 * ln);
                    wt[size_t(p) * n + o] = 2.0 * qw[p] * sw / double(n) * bb1;
                    bsum += qw[p] * zg * 16.0 * M_PI * M_PI * sinc(M_PI / L * qz[p] * ln);
                }
                b2[o] = bsum;
            }
    const size_t es = dtype == DGFS_F64 ? 8 : 4;
    auto upload = [&](void** dptr, const std::vector<double>& h) -> int {
        DGFS_HIP_CHECK(hipMalloc(dptr, h.size() * es));
        if (es == 8) {
            DGFS_HIP_CHECK(hipMemcpy(*dptr, h.data(), h.size() * 8, hipMemcpyHostToDevice));
        } else {
            std::vector<float> f(h.begin(), h.end());
            DGFS_HIP_CHECK(hipMemcpy(*dptr, f.data(), f.size() * 4, hipMemcpyHostToDevice));
        }
        return 0;
    };
    // the gain's z-axis factors: with e = A B (B the z phase), Re e + Im e =
    // A.x (B.x + B.y) + A.y (B.x - B.y); [Nrho][M][N] pairs (B.x + B.y, B.x - B.y)
    std::vector<double> pz(size_t(Nrho) * M * N * 2);
    for (int p = 0; p < Nrho; ++p)
        for (int q = 0; q < M; ++q)
            for (int i = 0; i < N; ++i) {
                const size_t o = (((size_t(p) * M + q) * 3 + 2) * N + i) * 2;
                const size_t d = ((size_t(p) * M + q) * N + i) * 2;
                pz[d] = ph[o] + ph[o + 1];
                pz[d + 1] = ph[o] - ph[o + 1];
            }
    int rc = upload(&pl->phase, ph);
    if (!rc) rc = upload(&pl->phz, pz);
    if (!rc) rc = upload(&pl->wtab, wt);
    if (!rc) rc = upload(&pl->bb2, b2);
    if (rc) { dgfs_collision_plan_destroy(pl); return rc; }
    const size_t cs = 2 * es;
    hipError_t e1 = hipMalloc(&pl->Fhat, size_t(ntask_max) * 2 * n * cs);
    hipError_t e2 = hipMalloc(&pl->acc, size_t(ntask_max) * Nrho * pl->QG * n * cs);
    pl->QG_cap = pl->QG;
    hipError_t e3 = hipMalloc(&pl->Iw, size_t(ntask_max) * 2 * n * cs);
    hipError_t e4 = hipMalloc(&pl->H, size_t(ntask_max) * 3 * n * cs);   // 3 planes: reference-state form
    hipError_t e5 = hipMalloc(&pl->nyq, size_t(ntask_max) * 2 * Nrho * M * 6 * N * N * es);
    if (e1 != hipSuccess || e2 != hipSuccess || e3 != hipSuccess || e4 != hipSuccess ||
        e5 != hipSuccess) {
        dgfs::set_error("collision_plan_create: scratch allocation failed");
        dgfs_collision_plan_destroy(pl);
        return 1;
    }
    *out = pl;
    return 0;
}

static void free_events(CollisionPlan* pl) {
    if (!pl->ev) return;
    for (int i = 0; i < pl->ev_max * dgfs::kTimingEvents; ++i) (void)hipEventDestroy(pl->ev[i]);
    delete[] pl->ev;
    pl->ev = nullptr;
    pl->ev_max = pl->ev_used = 0;
}

extern "C" int dgfs_collision_timing(dgfs_collision_plan_t plan, int max_calls) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collision_timing: null plan");
    DGFS_REQUIRE(max_calls >= 0, "collision_timing: bad max_calls");
    free_events(pl);
    if (max_calls == 0) return 0;
    constexpr int E = dgfs::kTimingEvents;
    pl->ev = new hipEvent_t[size_t(max_calls) * E];
    for (int i = 0; i < max_calls * E; ++i) DGFS_HIP_CHECK(hipEventCreate(&pl->ev[i]));
    pl->ev_max = max_calls;
    pl->ev_used = 0;
    return 0;
}

extern "C" int dgfs_collision_timing_read(dgfs_collision_plan_t plan, double* ms, int* ncalls) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl && ms && ncalls, "collision_timing_read: null argument");
    constexpr int E = dgfs::kTimingEvents;
    for (int s = 0; s < E - 1; ++s) ms[s] = 0.0;
    *ncalls = pl->ev_used;
    for (int c = 0; c < pl->ev_used; ++c) {
        DGFS_HIP_CHECK(hipEventSynchronize(pl->ev[c * E + E - 1]));
        for (int s = 0; s < E - 1; ++s) {
            float t = 0.f;
            DGFS_HIP_CHECK(hipEventElapsedTime(&t, pl->ev[c * E + s], pl->ev[c * E + s + 1]));
            ms[s] += t;
        }
    }
    pl->ev_used = 0;
    return 0;
}

extern "C" int dgfs_collision_plan_destroy(dgfs_collision_plan_t plan) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    if (!pl) return 0;
    free_events(pl);
    for (void* p : {pl->phase, pl->phz, pl->wtab, pl->bb2, pl->Fhat, pl->acc, pl->Iw, pl->H, pl->nyq,
                    pl->Fhat0, pl->ucache})
        if (p) (void)hipFree(p);
    delete pl;
    return 0;
}

extern "C" int dgfs_collision_plan_query(dgfs_collision_plan_t plan, int* split, int* qgroups,
                                         int* pgroups, size_t* scratch_bytes) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collision_plan_query: null plan");
    const size_t cs = pl->dtype == DGFS_F64 ? 16 : 8;
    if (split) *split = pl->SB;
    if (qgroups) *qgroups = pl->QG;
    if (pgroups) *pgroups = pl->SD;
    if (scratch_bytes)
        *scratch_bytes = size_t(pl->ntask_max) * pl->n * cs * (7 + pl->Nrho * pl->QG);
    return 0;
}

extern "C" int dgfs_collision_plan_set_qgroups(dgfs_collision_plan_t plan, int qgroups) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collision_plan_set_qgroups: null plan");
    DGFS_REQUIRE(qgroups >= 1 && qgroups <= pl->M, "collision_plan_set_qgroups: need 1 <= qg <= M");
    if (qgroups > pl->QG_cap) {
        const size_t cs = pl->dtype == DGFS_F64 ? 16 : 8;
        DGFS_HIP_CHECK(hipDeviceSynchronize());
        DGFS_HIP_CHECK(hipFree(pl->acc));
        pl->acc = nullptr;
        DGFS_HIP_CHECK(hipMalloc(&pl->acc, size_t(pl->ntask_max) * pl->Nrho * qgroups * pl->n * cs));
        pl->QG_cap = qgroups;
    }
    pl->QG = qgroups;
    return 0;
}

extern "C" int dgfs_collision_plan_set_schedule(dgfs_collision_plan_t plan, int split_mult,
                                                int xcd_remap) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collision_plan_set_schedule: null plan");
    DGFS_REQUIRE(split_mult == 1 || split_mult == 2,
                 "collision_plan_set_schedule: split_mult must be 1 or 2");
    DGFS_REQUIRE(xcd_remap == 0 || xcd_remap == 1,
                 "collision_plan_set_schedule: xcd_remap must be 0 or 1");
    pl->gain_smul = split_mult;
    pl->gain_remap = xcd_remap;
    return 0;
}

extern "C" int dgfs_collide(dgfs_collision_plan_t plan, const void* u, const void* ug, void* Q,
                            int K, int Ne,
                            int ntask, const int* task_nterm, const int* task_slice,
                            const void* task_coef, int ngroup, const int* group_start,
                            const int* group_slice, int fg_same, hipStream_t stream) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collide: null plan");
    DGFS_REQUIRE(ntask >= 0 && ntask <= pl->ntask_max, "collide: ntask=%d exceeds plan max %d",
                 ntask, pl->ntask_max);
    DGFS_REQUIRE(ngroup >= 0 && ngroup <= ntask, "collide: bad ngroup=%d", ngroup);
    DGFS_REQUIRE(K > 0 && Ne > 0, "collide: bad K/Ne");
    if (ntask == 0) return 0;
    DGFS_REQUIRE(u && Q && task_nterm && task_slice && task_coef && group_start && group_slice,
                 "collide: null pointer argument");
    if (pl->variant == DGFS_VHS_GLL) fg_same = 1;   // vhs-gll evaluates Q(f, f)
    if (ug == nullptr) ug = u;
    if (ug != u) fg_same = 0;
    dgfs::TaskView tv{K, Ne, int(pl->n), task_nterm, task_slice, task_coef};
    if (pl->dtype == DGFS_F64)
        return dgfs::dispatch<double>(pl, u, ug, Q, tv, ntask, ngroup, group_start,
                                      group_slice, fg_same, stream, nullptr);
    return dgfs::dispatch<float>(pl, u, ug, Q, tv, ntask, ngroup, group_start, group_slice,
                                 fg_same, stream, nullptr);
}

extern "C" int dgfs_collision_reference_set(dgfs_collision_plan_t plan, const void* u0, int K,
                                            int Ne, int ntask, const int* task_nterm,
                                            const int* task_slice, const void* task_coef,
                                            hipStream_t stream) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collision_reference_set: null plan");
    DGFS_REQUIRE(ntask > 0 && ntask <= pl->ntask_max,
                 "collision_reference_set: ntask=%d outside 1..%d", ntask, pl->ntask_max);
    DGFS_REQUIRE(K > 0 && Ne > 0, "collision_reference_set: bad K/Ne");
    DGFS_REQUIRE(u0 && task_nterm && task_slice && task_coef,
                 "collision_reference_set: null pointer argument");
    pl->ref_u0 = nullptr;
    pl->ref_ntask = 0;
    if (pl->ref_cap < pl->ntask_max) {
        // first use: the reference state's buffers for the plan's task capacity
        const size_t es = pl->dtype == DGFS_F64 ? 8 : 4;
        if (pl->Fhat0) (void)hipFree(pl->Fhat0);
        if (pl->ucache) (void)hipFree(pl->ucache);
        pl->Fhat0 = pl->ucache = nullptr;
        const size_t nf = size_t(pl->ntask_max) * 2 * pl->n * 2 * es;
        const size_t nc = size_t(pl->ntask_max) * pl->Nrho * pl->M * pl->n * 4 * es;
        if (hipMalloc(&pl->Fhat0, nf) != hipSuccess || hipMalloc(&pl->ucache, nc) != hipSuccess) {
            if (pl->Fhat0) (void)hipFree(pl->Fhat0);
            pl->Fhat0 = nullptr;
            pl->ucache = nullptr;
            pl->ref_cap = 0;
            dgfs::set_error("collision_reference_set: allocation of %zu bytes failed", nf + nc);
            return 1;
        }
        pl->ref_cap = pl->ntask_max;
    }
    dgfs::TaskView tv{K, Ne, int(pl->n), task_nterm, task_slice, task_coef};
    const int rc = pl->dtype == DGFS_F64
                       ? dgfs::dispatch_reference_set<double>(pl, u0, tv, ntask, stream)
                       : dgfs::dispatch_reference_set<float>(pl, u0, tv, ntask, stream);
    if (rc) return rc;
    pl->ref_u0 = u0;
    pl->ref_ntask = ntask;
    pl->ref_slice = task_slice;
    return 0;
}

extern "C" int dgfs_collide_reference(dgfs_collision_plan_t plan, const void* u, const void* u0,
                                      const double* Q0, void* Q, int K, int Ne, int ntask,
                                      const int* task_nterm, const int* task_slice,
                                      const void* task_coef, int ngroup,
                                      const int* group_start, const int* group_slice,
                                      hipStream_t stream) {
    auto* pl = static_cast<CollisionPlan*>(plan);
    DGFS_REQUIRE(pl, "collide_reference: null plan");
    DGFS_REQUIRE(ntask >= 0 && ntask <= pl->ntask_max,
                 "collide_reference: ntask=%d exceeds plan max %d", ntask, pl->ntask_max);
    DGFS_REQUIRE(ngroup >= 0 && ngroup <= ntask, "collide_reference: bad ngroup=%d", ngroup);
    DGFS_REQUIRE(K > 0 && Ne > 0, "collide_reference: bad K/Ne");
    if (ntask == 0) return 0;
    DGFS_REQUIRE(u && u0 && Q0 && Q && task_nterm && task_slice && task_coef && group_start &&
                     group_slice,
                 "collide_reference: null pointer argument");
    DGFS_REQUIRE(u0 != u, "collide_reference: the reference state must be its own array");
    DGFS_REQUIRE(pl->ref_u0 == u0 && pl->ref_ntask == ntask && pl->ref_slice == task_slice,
                 "collide_reference: the plan holds no reference state for this u0 and task "
                 "table (call dgfs_collision_reference_set after every change of u0)");
    dgfs::TaskView tv{K, Ne, int(pl->n), task_nterm, task_slice, task_coef};
    if (pl->dtype == DGFS_F64)
        return dgfs::dispatch<double>(pl, u, u0, Q, tv, ntask, ngroup, group_start,
                                      group_slice, 0, stream, Q0);
    return dgfs::dispatch<float>(pl, u, u0, Q, tv, ntask, ngroup, group_start, group_slice, 0,
                                 stream, Q0);
}
