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
 *   N      velocity points per axis (4, 6, 8, 12, 16, 20, 24, 32; 48 with DGFS_F32)
 *   qz,qw  N_rho radial nodes/weights on [0, R]   (velocitymesh.py:93-96)
 *   sz     M x 3 half-sphere directions, sw = 2 pi / M (velocitymesh.py:98-105)
 *   gamma  2 (1 - omega);  prefac as scattering.py:81-82
 */
int dgfs_collision_plan_create(int N, int M, int Nrho, double L, const double* qz,
                               const double* qw, const double* sz, double sw,
                               double gamma, double prefac, int dtype, int variant,
                               int ntask_max, dgfs_collision_plan_t* plan);
int dgfs_collision_plan_destroy(dgfs_collision_plan_t plan);
/* split S of the gain kernel's residue decomposition, its q-group count, the
 * split of the forward/weight/output kernels, and the scratch bytes */
int dgfs_collision_plan_query(dgfs_collision_plan_t plan, int* split, int* qgroups,
                              int* pgroups, size_t* scratch_bytes);
/* Tuning knob (no reference counterpart): the number of q groups the gain
 * kernel splits the M sphere directions into (more groups = more
 * work-groups, more partial sums).  1 <= qgroups <= M.  Grows the scratch
 * when needed (synchronises the device): call between launches, not while
 * capturing a graph. */
int dgfs_collision_plan_set_qgroups(dgfs_collision_plan_t plan, int qgroups);
/* Tuning knobs of the f = g gain kernel (no reference counterpart):
 * split_mult = 1 uses the smallest residue split S whose LDS image fits a
 * work-group, 2 uses 2S where 2S divides N with an unpadded LDS cube (more,
 * smaller work-groups, up to two per CU, each reloading the spectrum);
 * xcd_remap = 1 orders work-groups so that those of
 * one task share an XCD's L2, 0 keeps the dispatch order.  Results are the
 * same up to summation order.  Defaults (1, 1). */
int dgfs_collision_plan_set_schedule(dgfs_collision_plan_t plan, int split_mult, int xcd_remap);

/* Batched collision: replaces the per-node loop
 *     for m: U/V transform; for r in sigModes[m], e: sm.fs(d_f, d_g, d_ux, e, r, m)
 * of rhs() (dgfs1D/std/std.py:352-356) and fs() itself
 * (dgfs1D/std/scattering.py:251-313 'vhs-gll', :346-429 'vhs').
 * Task t evaluates Q(f_t, g_t) with
 *     f_t = sum_a coef[t][0][a] * u [slice[t][0][a]],  a < nterm[t][0]
 *     g_t = sum_b coef[t][1][b] * ug[slice[t][1][b]],  b < nterm[t][1]
 * (ug = NULL means ug = u; a distinct ug is the d_f / d_g pair of fs()).
 * (the U[m]^T / V[m] contraction of the SVD-decomposed H tensor fused into
 * the load, basis.py:163-177) and tasks are grouped by output slice:
 *     Q[group_slice[g]] += sum_{t in [group_start[g], group_start[g+1])}
 *                          prefac (Re Gain(f_t, g_t) - f_t Re Loss(g_t))
 * Device arrays: task_nterm [ntask][2] int, task_slice [ntask][2][K] int,
 * task_coef [ntask][2][K] dtype, group_start [ngroup+1] int,
 * group_slice [ngroup] int.  fg_same != 0 declares f_t == g_t for every task
 * (halves the work; forced for DGFS_VHS_GLL).  Distinct groups must have
 * distinct output slices. */
int dgfs_collide(dgfs_collision_plan_t plan, const void* u, const void* ug, void* Q, int K,
                 int Ne, int ntask,
                 const int* task_nterm, const int* task_slice, const void* task_coef,
                 int ngroup, const int* group_start, const int* group_slice, int fg_same,
                 hipStream_t stream);

/* Reference-state form of the same collision (no counterpart in the
 * reference, whose fs() is FP64 end to end; it replaces dgfs_collide for
 * f = g tasks when the plan computes in FP32).  With f_t the task's f
 * contraction of u and f0_t the same contraction of the reference state u0:
 *     Q[group_slice[g]] += Q0[group_slice[g]]
 *         + sum_{t in g} prefac (Re [Gain(f_t) - Gain(f0_t)] - h_t Re Loss(f_t)
 *                                - f0_t Re Loss(h_t)),      h = f - f0,
 * which equals dgfs_collide's Q(f, f) exactly when Q0 holds Q(f0, f0) (Gain
 * is a quadratic form and Loss is linear).  Q0: FP64, the (K, Ne, Nv)
 * layout, typically one FP64 dgfs_collide call on u0.  Every FP32 rounding
 * error of the spectral pipeline then scales with |h| instead of |f|: near
 * equilibrium the gain and the loss each exceed Q by orders of magnitude,
 * and FP32 rounding of those two large terms is what limits the direct form.
 *
 * dgfs_collision_reference_set(plan, u0, ...) must be called, on the same
 * stream, after every change of u0 and before the dgfs_collide_reference
 * calls that use it (same u0 pointer, ntask and task_slice pointer; other
 * calls are refused): it keeps F0^ and the inverse transforms
 * u+-_f0 = IFFT3(exp(+-i a_pq) F0^) of every (task, p, q) in the plan
 * (allocated on first use: ntask_max*N_rho*M*N^3*4 values of the plan's
 * dtype), so that each dgfs_collide_reference call transforms h alone:
 * Gain(f) - Gain(f0) = sum_pq w (u+h (u-h + u-f0) + u-h u+f0).  The g terms of
 * the task table are ignored.  u0 must be a different array from u. */
int dgfs_collision_reference_set(dgfs_collision_plan_t plan, const void* u0, int K, int Ne,
                                 int ntask, const int* task_nterm, const int* task_slice,
                                 const void* task_coef, hipStream_t stream);
int dgfs_collide_reference(dgfs_collision_plan_t plan, const void* u, const void* u0,
                           const double* Q0, void* Q, int K, int Ne, int ntask,
                           const int* task_nterm, const int* task_slice, const void* task_coef,
                           int ngroup, const int* group_start, const int* group_slice,
                           hipStream_t stream);

/* Stage timing (no counterpart in the reference, which only times the whole
 * loop, std/std.py:394-423): dgfs_collision_timing(plan, max_calls) creates
 * 6*max_calls HIP events (0 frees them); the next max_calls dgfs_collide calls
 * record an event before each of the five stages (forward transform, Nyquist
 * planes, gain, plane weight, line output) and after the last, on the launch
 * stream.  timing_read waits for the recorded events and returns the summed
 * elapsed ms per stage. */
int dgfs_collision_timing(dgfs_collision_plan_t plan, int max_calls);
int dgfs_collision_timing_read(dgfs_collision_plan_t plan, double* ms /*[5]*/, int* ncalls);

/* ------------------------------------------------- binary-mixture collision
 * Replaces DGFSBiVSSScatteringModel.perform_precomputation for one species
 * pair (p, q) (dgfs1D/bi/scattering.py:95-206, kernels precompute_a /
 * precompute_bc_pq of bi/kernels/scattering/vss.mako:29-93): the phase
 * tables and the complex VSS kernels b_pq[x][y][k], c_pq[k] are built in
 * FP64 on the device (pre-rule sums), then held in the dtype.
 *   N        velocity points per axis (4, 6, 8, 12, 16, 20, 24, 32; 48 with DGFS_F32)
 *   qz,qw    N_rho radial nodes/weights on [0, R]  (bi/velocitymesh.py:101-103)
 *   sz       M x 3 full-sphere directions, sw = 4 pi / M (bi/velocitymesh.py:111-118)
 *   szpre    Mpre x 3 full-sphere pre-rule (scattering.py:55-61), swpre = 4 pi / Mpre
 *   mp, mq   masses of the species pair (molar mass / molarMass0)
 *   alpha, omega  VSS parameters of the pair; prefac as scattering.py:86-90
 * Tasks are processed in chunks whose gain products fit a 2 GiB scratch
 * (environment variable DGFS_BI_SCRATCH_MB, read at plan creation, sets
 * another budget; at least one task per chunk). */
typedef void* dgfs_bi_collision_plan_t;
int dgfs_bi_collision_plan_create(int N, int M, int Nrho, double L, const double* qz,
                                  const double* qw, const double* sz, double sw, int Mpre,
                                  const double* szpre, double mp, double mq, double alpha,
                                  double omega, double prefac, int dtype, int ntask_max,
                                  dgfs_bi_collision_plan_t* plan);
int dgfs_bi_collision_plan_destroy(dgfs_bi_collision_plan_t plan);
/* split S of the gain kernel, tasks per chunk, scratch bytes */
int dgfs_bi_collision_plan_query(dgfs_bi_collision_plan_t plan, int* split, int* chunk,
                                 size_t* scratch_bytes);
/* Q_p[out] += prefac (gain - f_p loss) of the pair for every task: replaces
 * the per-node fs(cases[q], d_f, d_g, d_ux, e, r, m) loop of rhs()
 * (dgfs1D/bi/bi.py:329-337, scattering.py:261-321).  The task table is that
 * of dgfs_collide: f terms contract up (species p), g terms contract uq
 * (species q), the loss uses the f value.  Call once per partner q into the
 * same Q to form Q_p = sum_q Q_pq. */
int dgfs_bi_collide(dgfs_bi_collision_plan_t plan, const void* up, const void* uq, void* Q,
                    int K, int Ne, int ntask, const int* task_nterm, const int* task_slice,
                    const void* task_coef, int ngroup, const int* group_start,
                    const int* group_slice, hipStream_t stream);

/* Linear minmod slope limiter (reference limitLin,
 * dgfs1D/astd/kernels/limiters.mako:17-76, applied by astd/abstd.py:402-416
 * after each step when [time-integrator] limiter is set; nodal bases).
 * u, uavg (cell means at every node: the basis' cell-average operator),
 * ulx (slope of the linear part: derivative operator times invjac) and ulim
 * are (K, Ne, Nv) arrays of the dtype; xrel[k*Ne + e] = x_k,e - (cell
 * centre) and h[e] the element widths are FP64 DEVICE arrays.  Elements 0 and Ne-1
 * are copied; an interior element whose end values leave the minmod band of
 * its neighbours' means (threshold 1e-6) is replaced by its mean plus the
 * minmod slope.  ulim may alias u. */
int dgfs_limit_linear(int dtype, int K, int Ne, int Nv, const double* xrel, const double* h,
                      const void* u, const void* uavg, const void* ulx, void* ulim,
                      hipStream_t stream);

/* Positivity limiter of the upwind IMEX driver (reference limitPos,
 * dgfs1D/astd/kernels/limiters.mako:178-203, applied by astd/abstdu.py:432-442
 * after the linear limiter): per element and velocity, with a the cell mean
 * (uavg, as for dgfs_limit_linear) and m the smallest nodal value,
 * theta = min(1, |(a - 1e-13)/(a - m)|) (absolute value in float, as the
 * reference's fabsf) and ulim = a + theta (u - a).  ulim may alias u. */
int dgfs_limit_positivity(int dtype, int K, int Ne, int Nv, const void* u, const void* uavg,
                          void* ulim, hipStream_t stream);

/* ---------------------------------------------------------------- transport
 * Boundary kinds (dgfs1D/std/bc.py):
 *   COPY          'dgfs-periodic' / 'dgfs-cyclic' on one rank: ghost = interior
 *                 (bcs/dgfs-periodic.mako:12, dgfs-cyclic.mako:12)
 *   DIFFUSE       'dgfs-wall-diffuse': ghost = fac<0 ? n_w f0 : interior,
 *                 n_w = -sum_{fac>=0} fac f cw / sum_{fac<0} fac f0 cw   (bc.py:25-78)
 *   DIFFUSE_EXPR  'dgfs-wall-expr-diffuse': as DIFFUSE with the unnormalised
 *                 exp(-|c-u|^2/T) of (ux,uy,uz,T) evaluated in the kernel (bc.py:136-198)
 *   INLET         'dgfs-inlet' / 'dgfs-inlet-nondim': ghost = fac<0 ? f0 : interior
 *   VANISH        'dgfs-vanish': ghost = fac>=0 ? 0 : interior
 * with fac = (c_x - ushift) * nl, nl = -1 (left) / +1 (right).  The n_w sums
 * accumulate in FP64. */
#define DGFS_BC_COPY 0
#define DGFS_BC_DIFFUSE 1
#define DGFS_BC_DIFFUSE_EXPR 2
#define DGFS_BC_INLET 3
#define DGFS_BC_VANISH 4

typedef struct {
    int kind;
    double nl, ushift;
    const void* f0;            /* device, Nv values of dtype (DIFFUSE, INLET) */
    double ux, uy, uz, T;      /* DIFFUSE_EXPR, non-dimensional */
} dgfs_wall_t;

typedef struct {
    int dtype, K, Ne, Nv;
    const double* D;           /* device K*K: derivative operator, ux = D u (std.py:109-114) */
    const double* BL;          /* device K: left-face reconstruction  Bqf[:,0]      */
    const double* BR;          /* device K: right-face reconstruction Bqf[:,Nqf-1]  */
    const double* gLD;         /* device K: lift vectors (basis.py gLD, gRD) */
    const double* gRD;
    const double* invjac;      /* device Ne: 2/h_e (mesh.py:121) */
    const void* cx;            /* device Nv velocity components (dtype) */
    const void* cy;
    const void* cz;
    double cw;                 /* velocity cell volume (velocitymesh.py:60) */
    dgfs_wall_t left, right;
    double* work;              /* device, dgfs_transport_work_bytes(Nv) */
    double* nw;                /* device, 2 doubles: wall number densities */
    const double* Dm;          /* device K*K or NULL.  Non-null selects the upwind
                                * summation-by-parts form of the upwind driver
                                * (dgfs1D/astd/abstdu.py:360-378, upwind.mako):
                                * ux = -(Dm c+ u + D c- u) with D = Dp^T, Dm = Dm^T of
                                * basis.py:303-312, face jumps jL = c- (uR - uL),
                                * jR = -c+ (uR - uL), then the same lift and 1/J */
    const double* De;          /* device K*K or NULL.  Non-null (and Dm NULL) selects the
                                * entropy-stable volume term of the entropy driver
                                * (dgfs1D/astd/abstde2.py:214-217, eDeriv of
                                * limiters.mako): -2 sum_m De[l][m] c L(u_l, u_m) with
                                * L the logarithmic mean and De the basis derivMat,
                                * with the standard flux, lift and 1/J */
} dgfs_transport_t;

size_t dgfs_transport_work_bytes(int Nv);

/* Fused transport of rhs() (dgfs1D/std/std.py:283-348): face reconstruction
 * (bwdTransFace + extract_left/right), boundary update/apply/insert,
 * upwind flux, -c_x D u, lift with gLD/gRD and the 1/J scaling
 * (dgfs1D/std/kernels/std.mako:4-190).  Writes ux (K, Ne, Nv); t is the
 * stage time (kept for time-dependent boundaries evaluated on the host). */
int dgfs_transport_rhs(const dgfs_transport_t* tp, const void* u, void* ux, double t,
                       hipStream_t stream);

/* Batched transport (no reference counterpart: the reference runs one deck
 * per process, dgfs1D/std/std.py:37-423).  The transport of dgfs_transport_rhs
 * for nprob independent problems of one dtype, each described by its own
 * dgfs_transport_t (K, Ne, Nv, operators, velocity grid, walls; its work and
 * nw fields are not used), with state u + offset[p] and output ux + offset[p]
 * (offsets in elements of the dtype).  The walls' parameters are read at
 * creation: expression walls (DGFS_BC_DIFFUSE_EXPR, time-dependent) are
 * refused.  dgfs_transport_batch_rhs runs three launches for the whole batch
 * (diffuse-wall partial sums of every problem, their wall densities, the
 * transport of every (problem, element, velocity)); it allocates nothing and
 * may be captured in a graph.  dgfs_transport_batch_nw copies the 2*nprob
 * wall densities of the last call to the host (synchronising). */
typedef void* dgfs_transport_batch_t;
int dgfs_transport_batch_create(const dgfs_transport_t* tps, int nprob, const long* offset,
                                dgfs_transport_batch_t* batch);
int dgfs_transport_batch_destroy(dgfs_transport_batch_t batch);
int dgfs_transport_batch_rhs(dgfs_transport_batch_t batch, const void* u, void* ux,
                             hipStream_t stream);
int dgfs_transport_batch_nw(dgfs_transport_batch_t batch, double* nw_host);

/* out[k][j] (+)= sum_m A[k][m] in[m][j], j < ncol: the GiMMiK operators of
 * dgfs1D/util.py:20-25 (fwdTrans, bwdTrans, invMass, U/V, interpolation).
 * A is a device array of Kout*Kin doubles (row-major). */
int dgfs_matop(int dtype, int Kout, int Kin, long ncol, const double* A, const void* in,
               void* out, int accumulate, hipStream_t stream);

/* x0 = a0 x0 + a1 x1 + a2 x2 + a3 x3 (first nv terms; a0 == 0 never reads x0)
 * -- dgfs1D/axnpby.py:13-43 with the K sub-arrays flattened (n = K Ne Nv). */
int dgfs_axnpby(int dtype, long n, int nv, void* x0, const void* x1, const void* x2,
                const void* x3, double a0, double a1, double a2, double a3, hipStream_t stream);

/* Compensated FP32 stage update (no reference counterpart; extends
 * axnpby.py:13-43 for FP32 states carried as an unevaluated sum hi + lo):
 *     (x0, x0_lo) = a0 (x0 + x0_lo) + sum_{i<nv} a_i (x_i + x_i_lo)
 * in double-float arithmetic built from FP32 error-free transformations
 * (two_sum, fma product residual); relative error ~2^-44.  Any *_lo may be
 * NULL (zero low word; x0_lo NULL stores only the rounded sum).  a0 == 0 does
 * not read x0.  Integrators use it so that dt*L increments are not lost to
 * the FP32 rounding of the state. */
int dgfs_axnpby_compensated(long n, int nv, float* x0, float* x0_lo, const float* x1,
                            const float* x1_lo, const float* x2, const float* x2_lo,
                            const float* x3, const float* x3_lo, double a0, double a1, double a2,
                            double a3, hipStream_t stream);

/* FP64 moments at Nqr reconstruction points per element
 * (dgfs1D/std/moments.py:11-62, non-dimensional, before the dimensional
 * scaling of :63-70).  im: device K*Nqr (interpMat); out: device
 * [Nqr*Ne][9] = {rho, ux, uy, T, qx, qy, Pxx, Pyy, Pxy}, node = r*Ne + e. */
size_t dgfs_moments_work_bytes(int Nv, int nnode);
int dgfs_moments(int dtype, int K, int Ne, int Nqr, int Nv, const double* im, const void* u,
                 const void* cx, const void* cy, const void* cz, double cw, double* work,
                 double* out, hipStream_t stream);

/* out2 = {sum (cur-prev)^2, sum prev^2} in FP64 (dgfs1D/std/residual.py:21-22). */
size_t dgfs_residual_work_bytes(void);
int dgfs_residual(int dtype, long n, const void* cur, const void* prev, double* work,
                  double* out2, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Conservation projection (FP64).  No reference counterpart: the reference's
 * spectral Q conserves mass, momentum and energy only to spectral accuracy
 * (std/scattering.py:251-429); its one discrete-conservation mechanism is the
 * BGK Newton solve (std/scattering.py:647-695).
 *
 * For each of nslice slices Q_s (Nv points), with Phi = {1, cx, cy, cz, |c|^2}
 * (5 x Nv) and gram_inv = (Phi Phi^T)^{-1} (5x5 row-major, host memory):
 *     out_s (+)= Q_s - Phi^T gram_inv Phi Q_s
 * the Euclidean-nearest slice whose five moments vanish.  Moments, the 5x5
 * solve and the correction are evaluated in FP64.  out may alias Q when
 * accumulate == 0.  work: dgfs_projection_work_bytes(nslice, Nv) bytes. */
size_t dgfs_projection_work_bytes(int nslice, int Nv);
int dgfs_project_conservative(int dtype, int nslice, int Nv, const void* cx, const void* cy,
                              const void* cz, const double* gram_inv, const void* Q, void* out,
                              int accumulate, double* work, hipStream_t stream);

/* dgfs_project_conservative without the work argument: it allocates nothing.
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The workspace is a fixed array of
 * DGFS_PROJECTION_SCRATCH_BYTES in the library's device image, one per GPU,
 * held until the process ends; calls needing more
 * (dgfs_projection_work_bytes(nslice, Nv) > DGFS_PROJECTION_SCRATCH_BYTES,
 * e.g. nslice > 241 at Nv = 48^3) return status 2.  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (concurrent streams
 * or host threads) race.  Results are identical to dgfs_project_conservative. */
#define DGFS_PROJECTION_SCRATCH_BYTES (4u << 20)
int dgfs_project_conservative_noalloc(int dtype, int nslice, int Nv, const void* cx,
                                      const void* cy, const void* cz, const double* gram_inv,
                                      const void* Q, void* out, int accumulate,
                                      hipStream_t stream);

/* Mixture conservation projection (FP64; no reference counterpart).  Each
 * Q_pq of the binary VSS operator keeps species p's mass, but momentum and
 * energy are kept only by the pair sum, so for each slice s of the two
 * species' collision terms Q0_s, Q1_s (same node) the Euclidean-nearest pair
 * with
 *     sum Q0 = 0,  sum Q1 = 0,
 *     m0 sum c Q0 + m1 sum c Q1 = 0,  m0 sum |c|^2 Q0 + m1 sum |c|^2 Q1 = 0
 * is out0_s (+)= Q0_s - Phi0^T l, out1_s (+)= Q1_s - Phi1^T l with the six
 * constraint rows Phi_p = (delta_p0, delta_p1, m_p c, m_p |c|^2) and
 * l = gram_inv (Phi0 Q0 + Phi1 Q1), gram_inv = (Phi0 Phi0^T + Phi1 Phi1^T)^{-1}
 * (6x6 row-major, host memory, on the grid as the device holds it).
 * work: dgfs_projection_mixture_work_bytes(nslice, Nv) bytes. */
size_t dgfs_projection_mixture_work_bytes(int nslice, int Nv);
int dgfs_project_conservative_mixture(int dtype, int nslice, int Nv, const void* cx,
                                      const void* cy, const void* cz, double m0, double m1,
                                      const double* gram_inv, const void* Q0, const void* Q1,
                                      void* out0, void* out1, int accumulate, double* work,
                                      hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Shakov relaxation model 'shakov-direct-gll' (std/scattering.py:1636-1833,
 * std/kernels/scattering/shakov-direct-gll.mako:15-167): its fs(d_f, d_g, d_Q,
 * elem, modein, modeout) for every task of a batch in two launches.
 *
 * Task table: the layout of dgfs_collide, f-terms only (task_nterm[t*2],
 * task_slice[(t*2)*K + a], task_coef[(t*2)*K + a]); tasks sorted by output
 * slice, group g = tasks [group_start[g], group_start[g+1]) writing slice
 * group_slice[g].  For each task, with f_t = sum_a coef_a u[slice_a]:
 *   rho = cw sum f, u_i = sum f c_i / sum f, T = sum f |c-u|^2 cw / (1.5 rho),
 *   q_i = cw sum f |c-u|^2 (c-u)_i,
 *   f_S = rho/(pi T)^{3/2} exp(-|c-u|^2/T) (1 + 4(1-Pr)/5 (q.(c-u))/(rho T^2) (|c-u|^2/T - 5/2))
 *   Q[group slice] += prefac rho T^{1-omega} (f_S - f_t)
 * Moments accumulate in FP64.  work: dgfs_shakov_work_bytes(ntask, Nv) bytes. */
size_t dgfs_shakov_work_bytes(int ntask, int Nv);
int dgfs_shakov(int dtype, int K, int Ne, int Nv, int ntask, const int* task_nterm,
                const int* task_slice, const void* task_coef, int ngroup,
                const int* group_start, const int* group_slice, const void* u, const void* cx,
                const void* cy, const void* cz, double cw, double prefac, double omega,
                double Pr, void* Q, double* work, hipStream_t stream);

/* ES-BGK relaxation 'esbgk-direct-gll' (replaces fs() of
 * DGFSESBGKDirectGLLScatteringModelStd, dgfs1D/std/scattering.py:1540-1584,
 * kernels std/kernels/scattering/esbgk-direct-gll.mako:15-229), with the task
 * and group semantics of dgfs_shakov:
 *   Q[group slice] += prefac rho T^{1-omega} (G_t - f_t),
 *   G = rho / ((2 pi)^{3/2} sqrt(det L)) exp(-(c-u)^T L^{-1} (c-u) / 2),
 *   L = cw/rho ((1 - 1/Pr) sum f C C^T + (1/Pr) sum f_M C C^T), C = c - u,
 * f_M the node's discrete Maxwellian of (rho, u, T); prefac is the caller's
 * (t0 Pr p0 / mu(T0) in the reference).  Moments and L in FP64; Pr > 0.
 * work: dgfs_esbgk_work_bytes(ntask, Nv) bytes. */
size_t dgfs_esbgk_work_bytes(int ntask, int Nv);

/* Conservative BGK 'bgk-gll' (replaces fs() of DGFSBGKGLLScatteringModelStd,
 * dgfs1D/std/scattering.py:600-706, kernels bgk-gll.mako:83-229), task and
 * group semantics of dgfs_shakov: per task the discrete Maxwellian
 * fe = a0 exp(-a1 |c-u|^2 + a2 (c-u)_x + a3 (c-u)_y + a4 (c-u)_z) whose
 * discrete mass, momentum and energy equal f's (Newton's method from
 * a = (rho/(pi T)^1.5, 1/T, 0, 0, 0), at most 20 iterations, until every
 * moment residual is below 1e-12 of its scale), then
 *   Q[group slice] += prefac rho T^{1-omega} (fe - f_t).
 * FP64 moments and solve.  grid_n > 0 declares cx, cy, cz the tensor grid
 * of grid_n^3 = Nv points, cx[p] = c0[i], cy[p] = c0[j], cz[p] = c0[k] at
 * p = (i grid_n + j) grid_n + k (the reference's velocity mesh, whose points
 * are always laid out so; grid_n <= 64 for the tensor kernels, larger grids
 * take the point sums): the Newton sums then factor into one-dimensional
 * sums over c0 (= cz[0 .. grid_n)), one wave per task.  grid_n = 0: any
 * points; each task's velocity sums are split over up to 64 work-groups of
 * one launch sized to be resident at once (the work-groups of a task meet
 * once per iteration).  status: device array of ntask ints,
 * overwritten with per-task bits: 1 = a residual or the parameters became NaN
 * or infinite (the reference raises 'NaN encountered', scattering.py:689),
 * 2 = not converged within the iteration limit (the reference stops there
 * silently), 4 = a work-group wait gave up after one second (other kernels
 * held the device's compute units that long).  The call itself does not
 * synchronise; the caller reads status when it needs the verdict.
 * work: dgfs_bgk_work_bytes(ntask, Nv) bytes. */
size_t dgfs_bgk_work_bytes(int ntask, int Nv);

/* Conservative ES-BGK 'esbgk-gll' (replaces fs() of
 * DGFSESBGKGLLScatteringModelStd, dgfs1D/std/scattering.py:752-1185, kernels
 * esbgk-gll.mako:218-416), semantics of dgfs_shakov: the conservative BGK
 * target fe of dgfs_bgk_conservative, then T_ij = cw/rho sum (c-u)_i (c-u)_j
 * ((1 - 1/Pr) f + (1/Pr) fe) and the discrete Gaussian G whose mass,
 * momentum and second moments equal rho, rho u, rho (u_i u_j + T_ij)
 * (Newton's method, at most 40 iterations); Q += prefac rho T^{1-omega} (G - f),
 * prefac the caller's (t0 Pr p0 / mu(T0) in the reference).  Pr > 0.
 * grid_n: as dgfs_bgk_conservative's; on the tensor grid the ES-BGK sums are
 * taken column by column, the z-sums as Horner sums of a geometric factor
 * (one work-group per task).  The ES-BGK Newton starts from the
 * continuous Gaussian with the target covariance (the reference starts from
 * the Maxwellian of rho, u, T, esbgk-gll.mako:261-270; the root is the same).
 * status: as dgfs_bgk_conservative's, the bits of both solves combined.
 * work: dgfs_esbgk_conservative_work_bytes(ntask, Nv) bytes. */
size_t dgfs_esbgk_conservative_work_bytes(int ntask, int Nv);
int dgfs_esbgk_conservative(int dtype, int K, int Ne, int Nv, int grid_n, int ntask,
                            const int* task_nterm, const int* task_slice,
                            const void* task_coef, int ngroup,
                            const int* group_start, const int* group_slice, const void* u,
                            const void* cx, const void* cy, const void* cz, double cw,
                            double prefac, double omega, double Pr, void* Q, int* status,
                            double* work, hipStream_t stream);
int dgfs_bgk_conservative(int dtype, int K, int Ne, int Nv, int grid_n, int ntask,
                          const int* task_nterm, const int* task_slice,
                          const void* task_coef, int ngroup,
                          const int* group_start, const int* group_slice, const void* u,
                          const void* cx, const void* cy, const void* cz, double cw,
                          double prefac, double omega, void* Q, int* status, double* work,
                          hipStream_t stream);
int dgfs_esbgk(int dtype, int K, int Ne, int Nv, int ntask, const int* task_nterm,
               const int* task_slice, const void* task_coef, int ngroup,
               const int* group_start, const int* group_slice, const void* u, const void* cx,
               const void* cy, const void* cz, double cw, double prefac, double omega, double Pr,
               void* Q, double* work, hipStream_t stream);

/* ---------------------------------------------------------------------------
 * IMEX penalisation with the 14-moment Shakov model
 * (DGFSShakovDirectGLLScatteringModelAstd, dgfs1D/astd/scattering.py:355-397,
 * kernels astd/kernels/scattering/shakov-direct-gll.mako:193-691).
 * Moment vectors are FP64 [nslice][14] = {rho, rho u_x, rho u_y, rho u_z, E,
 * P_xx, P_yy, P_zz, P_xy, P_yz, P_zx, Q_x, Q_y, Q_z}: raw moments
 * sum_v f {1, c, |c|^2, c_i c_j, |c|^2 c} cw of each (mode, element) slice.
 *
 * dgfs_mom14         U = f . mat                     (moment(), scattering.py:155-162)
 * dgfs_shakov_target M = Shakov target of U          (momentNorm + cmaxwellian, mako:193-264)
 * dgfs_relax_collide Q = nu (M - f), nu = prefac rho T^{1-omega}   (collide, mako:266-305)
 * work of dgfs_mom14: dgfs_mom14_work_bytes(nslice, Nv) bytes. */
size_t dgfs_mom14_work_bytes(int nslice, int Nv);
int dgfs_mom14(int dtype, int nslice, int Nv, const void* f, const void* cx, const void* cy,
               const void* cz, double cw, double* work, double* U, hipStream_t stream);
int dgfs_shakov_target(int dtype, int nslice, int Nv, const double* U, const void* cx,
                       const void* cy, const void* cz, double Pr, void* M, hipStream_t stream);
/* ES-BGK target of the moments (the first 11 of the 14; esbgk penalisation,
 * dgfs1D/astd/scattering.py:287-354, momentNorm + cmaxwellian of
 * astd/kernels/scattering/esbgk-direct-gll.mako:157-255):
 *   Theta = T/Pr I + 2 (1 - 1/Pr) (P/rho - u u^T),
 *   M = rho / (pi^{3/2} sqrt(det Theta)) exp(-(c-u)^T Theta^{-1} (c-u)).
 * With srate = 1/Pr, dgfs_imex_lm_update / dgfs_imex_ars_update relax the
 * stresses at rate nu/Pr, the reference's ES-BGK moment update. */
int dgfs_esbgk_target(int dtype, int nslice, int Nv, const double* U, const void* cx,
                      const void* cy, const void* cz, double Pr, void* M, hipStream_t stream);
int dgfs_relax_collide(int dtype, int nslice, int Nv, const double* U, const void* M,
                       const void* f, double prefac, double omega, void* Q, hipStream_t stream);

#define DGFS_IMEX_MAXS 8

/* s-step IMEX linear multistep update (updateMomentLM / updateDistLM,
 * astd/scattering.py:262-279; kernels updateMom{s}_LM and
 * updateDistribution{s}_LM, mako:548-691).  History index i = 0..s-1 is
 * oldest first; term i carries a[s-1-i], b[s-1-i], c[s-i]:
 *   what = 0:  Unew[k] = sum_i (-a U_i + dt b LU_i + dt c LM_i)[k] for k < 5;
 *              stress / heat flux relax implicitly with c[0] dt nu(Unew)
 *              (heat flux at rate Pr nu) towards their Shakov values.
 *   what = 1:  fnew = (sum_i (-a F_i + dt b L_i + dt c Pn_i) + c[0] dt nu Mnew)
 *                     / (1 + c[0] dt nu),  nu = prefac rho T^{1-omega} of Unew. */
typedef struct {
    int nstages;
    double dt, prefac, omega, Pr;
    double a[DGFS_IMEX_MAXS], b[DGFS_IMEX_MAXS], c[DGFS_IMEX_MAXS + 1];
    const double* U[DGFS_IMEX_MAXS];
    const double* LU[DGFS_IMEX_MAXS];
    const double* LM[DGFS_IMEX_MAXS];
    double* Unew;
    const void* F[DGFS_IMEX_MAXS];
    const void* L[DGFS_IMEX_MAXS];
    const void* Pn[DGFS_IMEX_MAXS];
    const void* Mnew;
    void* fnew;
    double srate;   /* stress relaxation rate / nu: 1 (Shakov), 1/Pr (ES-BGK); > 0 */
} dgfs_lm_t;
int dgfs_imex_lm_update(int dtype, int nslice, int Nv, const dgfs_lm_t* args, int what,
                        hipStream_t stream);

/* Stage s (>= 1) of an ARS IMEX Runge-Kutta tableau (updateMomentARS /
 * updateDistARS, astd/scattering.py:220-240; kernels updateMom{s}_ARS and
 * updateDistribution{s}_ARS, mako:326-484): explicit weights ah[0..s-1] on
 * the explicit parts L_0..L_{s-1}, implicit weights ai[1..s] on the
 * penalisations of stages 1..s (stage s implicit):
 *   what = 0:  U[s] from U[0], LU[0..s-1] and U[1..s-1]
 *   what = 1:  F[s] = (F[0] + sum dt ah_i L_i + sum_{i<s} dt ai_i nu_i (M_i - F_i)
 *                      + dt ai_s nu_s M_s) / (1 + dt ai_s nu_s)
 * M[i] is the Shakov target of U[i]; F[s] may alias F[0]. */
typedef struct {
    int stage;
    double dt, prefac, omega, Pr;
    double ah[DGFS_IMEX_MAXS];
    double ai[DGFS_IMEX_MAXS + 1];
    double* U[DGFS_IMEX_MAXS + 1];
    const double* LU[DGFS_IMEX_MAXS];
    void* F[DGFS_IMEX_MAXS + 1];
    const void* L[DGFS_IMEX_MAXS];
    const void* M[DGFS_IMEX_MAXS + 1];
    double srate;   /* stress relaxation rate / nu: 1 (Shakov), 1/Pr (ES-BGK); > 0 */
} dgfs_ars_t;
int dgfs_imex_ars_update(int dtype, int nslice, int Nv, const dgfs_ars_t* args, int what,
                         hipStream_t stream);

/* ---------------------------------------------------------------------------
 * Learned collision closure (no reference counterpart; its training targets
 * are dgfs_collide minus dgfs_shakov, std/scattering.py:1636-1833).
 * Network on the G^3 velocity grid (G = 16, 32 or 48): conv3x3x3 1->32, SiLU,
 * 32->32, SiLU, 32->32, SiLU, 32->1; zero padding; tap = (dz+1)*9+(dy+1)*3+(dx+1),
 * out[v] = sum_tap W[tap] . in[v + d(tap)].  Device weights: W1 [27][32] and
 * W4 [27][32] FP32; W2/W3 as bf16 copies W*f [27][32 out][32 in] and
 * tap-flipped transposes W*t [27][32 in][32 out] (dgfs_closure_pack from FP32
 * masters [27][32][32]); biases FP32.  Every forward layer and the 32->32
 * input and weight gradients use bf16 MFMA with FP32 accumulation (the scaled
 * input, W1 and W4 are rounded to bf16 for the forward); activations are bf16
 * channels-last.
 *
 * dgfs_closure_forward: input slice b is scale[b] * u[src[b]] (solver dtype);
 *   act = 6 * nb*G^3*32 bf16 (Z1 Y1 Z2 Y2 Z3 Y3, kept for the backward pass);
 *   mode 0: pred[b][v] = network output (FP32);
 *   mode 1: Q[dst[b]][v] += oscale[b] * output (solver dtype).
 * dgfs_closure_backward: gradients of sum_v dOut[b][v] * pred[b][v] with
 *   respect to every weight and bias (FP32, overwritten); gact = 2 * nb*G^3*32
 *   bf16, fin = nb*G^3 FP32, work = dgfs_closure_work_bytes(G, nb, chunk)
 *   (room for the per-work-group weight-gradient partials, reduced in a fixed
 *   order; chunk >= 1 only sizes the buffer).
 * dgfs_closure_loss: loss[b] = |pred_b - target_b|^2 / max(|ref_b|^2, floor2) (FP64;
 *   ref = NULL means ref = target) and dOut = d(mean_b loss_b)/d(pred);
 *   work = dgfs_closure_loss_work_bytes.
 * dgfs_adam: w -= lr_t m / (sqrt(v) + eps) after the moment updates. */
typedef struct {
    const float* W1; const float* b1;
    const void* W2f; const void* W2t; const float* b2;
    const void* W3f; const void* W3t; const float* b3;
    const float* W4; const float* b4;
} dgfs_closure_t;
typedef struct {
    float* W1; float* b1; float* W2; float* b2; float* W3; float* b3; float* W4; float* b4;
} dgfs_closure_grad_t;
size_t dgfs_closure_work_bytes(int G, int nb, int chunk);
int dgfs_closure_forward(const dgfs_closure_t* net, int dtype, int G, int nb, const void* u,
                         const int* src, const float* scale, void* act, int mode, float* pred,
                         const int* dst, const float* oscale, void* Q, hipStream_t stream);
int dgfs_closure_backward(const dgfs_closure_t* net, const dgfs_closure_grad_t* grad, int dtype,
                          int G, int nb, const void* u, const int* src, const float* scale,
                          const void* act, const float* dOut, void* gact, float* fin, int chunk,
                          float* work, hipStream_t stream);
int dgfs_closure_pack(const float* W2, const float* W3, void* W2f, void* W2t, void* W3f,
                      void* W3t, hipStream_t stream);
size_t dgfs_closure_loss_work_bytes(int G, int nb);
int dgfs_closure_loss(int G, int nb, const float* pred, const float* target, const float* ref,
                      double floor2, float* dOut, double* loss, double* work, hipStream_t stream);
int dgfs_adam(int n, float* w, const float* g, float* m, float* v, double lr_t, double beta1,
              double beta2, double eps, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* DGFS_HIP_H */
