/* dgfs_hip.h -- C ABI of libdgfs_hip.so, the MI355X (gfx950) kernels of the
 * DGFS 1D3V single-species operator path.
 *
 * Conventions (all entry points):
 *   - return int status: 0 = OK, 1 = HIP runtime error, 2 = invalid argument;
 *     the message of the last failure on the calling thread is returned by
 *     dgfs_last_error().
 *   - every array argument is a DEVICE pointer owned by the caller (PyTorch
 *     tensors in the Python host layer); plans own their tables and scratch,
 *     allocated once at plan creation, so launch calls never allocate and
 *     never synchronise the host (they are stream-ordered on `stream`).
 *   - dtype: DGFS_F64 (double) or DGFS_F32 (float); all floating arrays of a
 *     call share the plan's dtype unless the parameter says otherwise.
 *   - coefficient arrays use the reference layout (K, Ne, Nv) row-major with
 *     the velocity index fastest, slice s = mode*Ne + elem at offset s*Nv
 *     (dgfs1D/std/kernels/std.mako:106, vhs-gll.mako:408); the velocity index
 *     is v = I*N*N + J*N + K with I <-> c_x slowest
 *     (dgfs1D/std/velocitymesh.py:66-72).
 */
#ifndef DGFS_HIP_H
#define DGFS_HIP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DGFS_F64 0
#define DGFS_F32 1

#define DGFS_VHS 0      /* scattering-model type 'vhs'     (bilinear Q(f,g)) */
#define DGFS_VHS_GLL 1  /* scattering-model type 'vhs-gll' (Q(f,f))          */

typedef void* dgfs_collision_plan_t;

const char* dgfs_last_error(void);

/* ---------------------------------------------------------------- collision
 * Replaces DGFSVHSGLLScatteringModelStd.perform_precomputation
 * (dgfs1D/std/scattering.py:106-246) and the precompute_aa / precompute_bb
 * kernels (dgfs1D/std/kernels/scattering/vhs-gll.mako:43-99): builds the
 * phase, radial-weight (bb1) and loss (bb2) tables in FP64 on the host and
 * allocates scratch for up to ntask_max collision tasks.
 *   N      velocity points per axis (4, 6, 8, 12, 16, 20, 24, 32)
 *   qz,qw  N_rho radial nodes/weights on [0, R]   (velocitymesh.py:93-96)
 *   sz     M x 3 half-sphere directions, sw = 2 pi / M (velocitymesh.py:98-105)
 *   gamma  2 (1 - omega);  prefac as scattering.py:81-82
 */
int dgfs_collision_plan_create(int N, int M, int Nrho, double L, const double* qz,
                               const double* qw, const double* sz, double sw,
                               double gamma, double prefac, int dtype, int variant,
                               int ntask_max, dgfs_collision_plan_t* plan);
int dgfs_collision_plan_destroy(dgfs_collision_plan_t plan);
/* split factor S of the residue decomposition, q/p group counts, scratch */
int dgfs_collision_plan_query(dgfs_collision_plan_t plan, int* split, int* qgroups,
                              int* pgroups, size_t* scratch_bytes);

/* Batched collision: replaces the per-node loop
 *     for m: U/V transform; for r in sigModes[m], e: sm.fs(d_f, d_g, d_ux, e, r, m)
 * of rhs() (dgfs1D/std/std.py:352-356) and fs() itself
 * (dgfs1D/std/scattering.py:251-313 'vhs-gll', :346-429 'vhs').
 * Task t evaluates Q(f_t, g_t) with
 *     f_t = sum_a coef[t][0][a] * u[slice[t][0][a]],  a < nterm[t][0]
 *     g_t = sum_b coef[t][1][b] * u[slice[t][1][b]],  b < nterm[t][1]
 * (the U[m]^T / V[m] contraction of the SVD-decomposed H tensor fused into
 * the load, basis.py:163-177) and tasks are grouped by output slice:
 *     Q[group_slice[g]] += sum_{t in [group_start[g], group_start[g+1])}
 *                          prefac (Re Gain(f_t, g_t) - f_t Re Loss(g_t))
 * Device arrays: task_nterm [ntask][2] int, task_slice [ntask][2][K] int,
 * task_coef [ntask][2][K] dtype, group_start [ngroup+1] int,
 * group_slice [ngroup] int.  fg_same != 0 declares f_t == g_t for every task
 * (halves the work; forced for DGFS_VHS_GLL).  Distinct groups must have
 * distinct output slices. */
int dgfs_collide(dgfs_collision_plan_t plan, const void* u, void* Q, int K, int Ne, int ntask,
                 const int* task_nterm, const int* task_slice, const void* task_coef,
                 int ngroup, const int* group_start, const int* group_slice, int fg_same,
                 hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DGFS_HIP_H */
