// IMEX penalisation kernels for the DGFS 1D3V single-species path (gfx950):
// the 14-moment Shakov penalisation P(f) = nu (M_S[U(f)] - f) and the
// stage updates of the IMEX linear-multistep and ARS integrators.
//
// Reference behaviour (jaisw7/dgfs1D_gpu):
//   moments U = f . mat (cuBLAS GEMM)   dgfs1D/astd/scattering.py:155-162, :380-397
//   momentNorm + cmaxwellian            astd/kernels/scattering/shakov-direct-gll.mako:193-264
//   collide                             :266-305
//   updateMom{s}_ARS / updateDistribution{s}_ARS   :326-484
//   updateMom{s}_LM / updateDistribution{s}_LM     :548-691
//
// U is kept in FP64 for every dtype: [slice][14] = {rho, rho u (3), E,
// P_xx P_yy P_zz P_xy P_yz P_zx, Q_x Q_y Q_z} (raw moments times cw).
// Distribution arrays are (nslice, Nv) of the solver dtype.
#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {
namespace {

constexpr int kIB = 256;
constexpr int kNA = 14;

template <int NV>
DGFS_DEV void block_sum(double (&v)[NV], double* sh) {
    const int tid = threadIdx.x;
    for (int i = 0; i < NV; ++i) sh[i * kIB + tid] = v[i];
    __syncthreads();
    for (int s = kIB / 2; s > 0; s >>= 1) {
        if (tid < s)
            for (int i = 0; i < NV; ++i) sh[i * kIB + tid] += sh[i * kIB + tid + s];
        __syncthreads();
    }
    for (int i = 0; i < NV; ++i) v[i] = sh[i * kIB];
}

// ------------------------------------------------------------ 14 moments
template <typename T>
__global__ void __launch_bounds__(kIB)
k_mom14_partial(int n, const T* __restrict__ f, const T* __restrict__ cx,
                const T* __restrict__ cy, const T* __restrict__ cz,
                double* __restrict__ part /*[slice][nb][14]*/) {
    __shared__ double sh[kNA * kIB];
    const int nb = gridDim.x, s = blockIdx.y;
    const int v = blockIdx.x * kIB + threadIdx.x;
    double a[kNA];
    for (int i = 0; i < kNA; ++i) a[i] = 0.0;
    if (v < n) {
        const double g = f[size_t(s) * n + v];
        const double x = cx[v], y = cy[v], z = cz[v];
        const double c2 = x * x + y * y + z * z;
        a[0] = g;
        a[1] = g * x; a[2] = g * y; a[3] = g * z;
        a[4] = g * c2;
        a[5] = g * x * x; a[6] = g * y * y; a[7] = g * z * z;
        a[8] = g * x * y; a[9] = g * y * z; a[10] = g * z * x;
        a[11] = g * c2 * x; a[12] = g * c2 * y; a[13] = g * c2 * z;
    }
    block_sum<kNA>(a, sh);
    if (threadIdx.x == 0)
        for (int i = 0; i < kNA; ++i) part[(size_t(s) * nb + blockIdx.x) * kNA + i] = a[i];
}

__global__ void __launch_bounds__(kIB)
k_mom14_final(int nslice, int nb, double cw, const double* __restrict__ part,
              double* __restrict__ U) {
    const int i = blockIdx.x * kIB + threadIdx.x;      // (slice, moment)
    if (i >= nslice * kNA) return;
    const int s = i / kNA, m = i - (i / kNA) * kNA;
    double acc = 0.0;
    for (int b = 0; b < nb; ++b) acc += part[(size_t(s) * nb + b) * kNA + m];
    U[i] = acc * cw;
}

// ------------------------------------------------ macroscopic state of U
struct Macro {
    double rho, ux, uy, uz, T, qx, qy, qz;
};

// temperature of an unnormalised moment vector (mako:278-285)
DGFS_DEV double temp_of(const double* U) {
    return (U[4] - (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / U[0]) / (1.5 * U[0]);
}

// momentNorm (mako:193-230): velocity, temperature, central heat flux
DGFS_DEV Macro macro_of(const double* U) {
    Macro m;
    m.rho = U[0];
    m.ux = U[1] / U[0];
    m.uy = U[2] / U[0];
    m.uz = U[3] / U[0];
    const double u2 = m.ux * m.ux + m.uy * m.uy + m.uz * m.uz;
    m.T = U[4] / (1.5 * U[0]) - u2 / 1.5;
    // rows (xx xy zx), (xy yy yz), (zx yz zz) of the raw second moments
    const double Pu_x = U[5] * m.ux + U[8] * m.uy + U[10] * m.uz;
    const double Pu_y = U[8] * m.ux + U[6] * m.uy + U[9] * m.uz;
    const double Pu_z = U[10] * m.ux + U[9] * m.uy + U[7] * m.uz;
    m.qx = U[11] - 2.0 * Pu_x + m.rho * m.ux * u2 - 1.5 * m.ux * m.rho * m.T;
    m.qy = U[12] - 2.0 * Pu_y + m.rho * m.uy * u2 - 1.5 * m.uy * m.rho * m.T;
    m.qz = U[13] - 2.0 * Pu_z + m.rho * m.uz * u2 - 1.5 * m.uz * m.rho * m.T;
    return m;
}

// Shakov target from U (cmaxwellian, mako:232-264)
DGFS_DEV double shakov_target(const Macro& m, double Pr, double x, double y, double z) {
    const double Cx = x - m.ux, Cy = y - m.uy, Cz = z - m.uz;
    const double C2 = Cx * Cx + Cy * Cy + Cz * Cz;
    const double fe = m.rho / pow(kPi * m.T, 1.5) * exp(-C2 / m.T);
    return fe * (1.0 + 2.0 * (1.0 - Pr) / 5.0 * (m.qx * Cx + m.qy * Cy + m.qz * Cz) /
                           (m.rho * m.T * m.T) * (2.0 * C2 / m.T - 5.0));
}

template <typename T>
__global__ void __launch_bounds__(kIB)
k_target(int n, const double* __restrict__ U, const T* __restrict__ cx,
         const T* __restrict__ cy, const T* __restrict__ cz, double Pr, T* __restrict__ M) {
    const int s = blockIdx.y, v = blockIdx.x * kIB + threadIdx.x;
    if (v >= n) return;
    const Macro m = macro_of(U + size_t(s) * kNA);
    M[size_t(s) * n + v] = T(shakov_target(m, Pr, cx[v], cy[v], cz[v]));
}

// ES-BGK target (esbgk-direct-gll.mako:157-255): Theta = T/Pr I
// + 2 (1 - 1/Pr) (P/rho - u u^T), M = rho/(pi^1.5 sqrt det) exp(-C^T Theta^-1 C)
template <typename T>
__global__ void __launch_bounds__(kIB)
k_es_target(int n, const double* __restrict__ U, const T* __restrict__ cx,
            const T* __restrict__ cy, const T* __restrict__ cz, double Pr, T* __restrict__ M) {
    const int s = blockIdx.y, v = blockIdx.x * kIB + threadIdx.x;
    if (v >= n) return;
    const double* u = U + size_t(s) * kNA;
    const Macro m = macro_of(u);
    const double a = 1.0 / Pr, b = 2.0 * (1.0 - 1.0 / Pr), r = m.rho;
    const double xx = a * m.T + b * (u[5] / r - m.ux * m.ux);
    const double yy = a * m.T + b * (u[6] / r - m.uy * m.uy);
    const double zz = a * m.T + b * (u[7] / r - m.uz * m.uz);
    const double xy = b * (u[8] / r - m.ux * m.uy);
    const double yz = b * (u[9] / r - m.uy * m.uz);
    const double zx = b * (u[10] / r - m.uz * m.ux);
    const double det = xx * (yy * zz - yz * yz) - xy * (xy * zz - yz * zx) +
                       zx * (xy * yz - yy * zx);
    const double Cx = double(cx[v]) - m.ux, Cy = double(cy[v]) - m.uy, Cz = double(cz[v]) - m.uz;
    const double q = (yy * zz - yz * yz) * Cx * Cx + (xx * zz - zx * zx) * Cy * Cy +
                     (xx * yy - xy * xy) * Cz * Cz +
                     2.0 * ((yz * zx - zz * xy) * Cx * Cy + (xy * zx - xx * yz) * Cy * Cz +
                            (xy * yz - yy * zx) * Cz * Cx);
    M[size_t(s) * n + v] = T(r / (pow(kPi, 1.5) * sqrt(det)) * exp(-q / det));
}

// Q = nu (M - f), nu = prefac rho T^{1-omega}  (collide, mako:266-305)
template <typename T>
__global__ void __launch_bounds__(kIB)
k_collide(int n, const double* __restrict__ U, const T* __restrict__ M, const T* __restrict__ f,
          double prefac, double omega, T* __restrict__ Q) {
    const int s = blockIdx.y, v = blockIdx.x * kIB + threadIdx.x;
    if (v >= n) return;
    const double* u = U + size_t(s) * kNA;
    const double nu = prefac * u[0] * pow(temp_of(u), 1.0 - omega);
    const size_t o = size_t(s) * n + v;
    Q[o] = T(nu * (double(M[o]) - double(f[o])));
}

// ------------------------------------------------ relaxation targets of U
// equilibrium value of component j (5..13) given U (mako:583-640)
struct Targets {
    double P[6], Qf[3];   // stress targets, heat-flux targets (without the -Pr U term)
};

DGFS_DEV Targets targets_of(const double* U, double T, double Pr) {
    Targets t;
    const double r = U[0];
    t.P[0] = 0.5 * r * T + U[1] * U[1] / r;
    t.P[1] = 0.5 * r * T + U[2] * U[2] / r;
    t.P[2] = 0.5 * r * T + U[3] * U[3] / r;
    t.P[3] = U[1] * U[2] / r;
    t.P[4] = U[2] * U[3] / r;
    t.P[5] = U[3] * U[1] / r;
    const double m2 = (U[1] * U[1] + U[2] * U[2] + U[3] * U[3]) / (r * r);
    const int q[3][3] = {{0, 3, 5}, {3, 1, 4}, {5, 4, 2}};
    for (int i = 0; i < 3; ++i) {
        const double Pu = U[5 + q[i][0]] * U[1] + U[5 + q[i][1]] * U[2] + U[5 + q[i][2]] * U[3];
        const double ui = U[1 + i];
        t.Qf[i] = (1.0 - Pr) * (-2.0 * Pu / r + ui * m2 - 1.5 * ui * T) + (ui * m2 + 2.5 * ui * T);
    }
    return t;
}

}  // namespace
}  // namespace dgfs

using namespace dgfs;

// --------------------------------------------------------------- LM update
// U_s[k] = sum_{i<s} (-a[s-1-i] U_i[k] + dt b[s-1-i] LU_i[k] + dt c[s-i] LM_i[k]),
// k < 5 explicitly; stress and heat flux relax implicitly with
// nu_s = prefac rho_s T_s^{1-omega} (rate Pr nu_s for the heat flux).
__global__ void __launch_bounds__(kIB)
k_lm_moments(int nslice, dgfs_lm_t P) {
    const int s = blockIdx.x * kIB + threadIdx.x;
    if (s >= nslice) return;
    const int ns = P.nstages;
    double E[kNA];
    for (int k = 0; k < kNA; ++k) {
        double acc = 0.0;
        for (int i = 0; i < ns; ++i) {
            const size_t o = size_t(s) * kNA + k;
            acc += -P.a[ns - 1 - i] * P.U[i][o] + P.dt * P.b[ns - 1 - i] * P.LU[i][o] +
                   P.dt * P.c[ns - i] * P.LM[i][o];
        }
        E[k] = acc;
    }
    double* Un = P.Unew + size_t(s) * kNA;
    for (int k = 0; k < 5; ++k) Un[k] = E[k];
    const double T = temp_of(E);
    const double nu = P.c[0] * P.dt * P.prefac * E[0] * pow(T, 1.0 - P.omega);
    // stresses first; the heat-flux targets use the new stresses (mako:583-640)
    double Us[kNA];
    for (int k = 0; k < kNA; ++k) Us[k] = E[k];
    Targets t = targets_of(Us, T, P.Pr);
    const double nus = nu * P.srate;
    for (int j = 0; j < 6; ++j) {
        Us[5 + j] = (E[5 + j] + nus * t.P[j]) / (1.0 + nus);
        Un[5 + j] = Us[5 + j];
    }
    t = targets_of(Us, T, P.Pr);
    for (int i = 0; i < 3; ++i) Un[11 + i] = (E[11 + i] + nu * t.Qf[i]) / (1.0 + P.Pr * nu);
}

// f_s = (sum_{i<s} (-a[s-1-i] F_i + dt b[s-1-i] L_i + dt c[s-i] Pn_i) + c0 dt nu_s M_s)
//       / (1 + c0 dt nu_s)
template <typename T>
__global__ void __launch_bounds__(kIB)
k_lm_dist(int n, dgfs_lm_t P) {
    const int s = blockIdx.y, v = blockIdx.x * kIB + threadIdx.x;
    if (v >= n) return;
    const int ns = P.nstages;
    const size_t o = size_t(s) * n + v;
    const double* U = P.Unew + size_t(s) * kNA;
    const double nu = P.prefac * U[0] * pow(temp_of(U), 1.0 - P.omega);
    double acc = 0.0;
    for (int i = 0; i < ns; ++i)
        acc += -P.a[ns - 1 - i] * double(static_cast<const T*>(P.F[i])[o]) +
               P.dt * P.b[ns - 1 - i] * double(static_cast<const T*>(P.L[i])[o]) +
               P.dt * P.c[ns - i] * double(static_cast<const T*>(P.Pn[i])[o]);
    const double cdn = P.c[0] * P.dt * nu;
    static_cast<T*>(P.fnew)[o] =
        T((acc + cdn * double(static_cast<const T*>(P.Mnew)[o])) / (1.0 + cdn));
}

// -------------------------------------------------------------- ARS update
// stage s (1-based) of an ARS(s) tableau: explicit weights ah[0..s-1] on
// L_0..L_{s-1}, implicit weights ai[1..s] on the penalisation of U_1..U_s.
__global__ void __launch_bounds__(kIB)
k_ars_moments(int nslice, dgfs_ars_t P) {
    const int s = blockIdx.x * kIB + threadIdx.x;
    if (s >= nslice) return;
    const int ns = P.stage;
    const size_t b = size_t(s) * kNA;
    double* Un = P.U[ns] + b;
    double E[kNA];
    for (int k = 0; k < kNA; ++k) {
        double acc = P.U[0][b + k];
        for (int i = 0; i < ns; ++i) acc += P.dt * P.ah[i] * P.LU[i][b + k];
        E[k] = acc;
    }
    for (int k = 0; k < 5; ++k) Un[k] = E[k];
    // explicit relaxation of the earlier stages (mako:368-452)
    for (int i = 1; i < ns; ++i) {
        const double* Ui = P.U[i] + b;
        const double Ti = temp_of(Ui);
        const double w = P.ai[i] * P.dt * P.prefac * Ui[0] * pow(Ti, 1.0 - P.omega);
        const Targets t = targets_of(Ui, Ti, P.Pr);
        for (int j = 0; j < 6; ++j) E[5 + j] += P.srate * w * (t.P[j] - Ui[5 + j]);
        for (int q = 0; q < 3; ++q) E[11 + q] += w * (t.Qf[q] - P.Pr * Ui[11 + q]);
    }
    // implicit stage (conserved moments of U_s are already final)
    double Us[kNA];
    for (int k = 0; k < 5; ++k) Us[k] = E[k];
    const double Ts = temp_of(Us);
    const double w = P.ai[ns] * P.dt * P.prefac * Us[0] * pow(Ts, 1.0 - P.omega);
    // targets use only the conserved part and, for the heat flux, the new
    // stresses: evaluate the stress first, then the heat flux from them
    for (int j = 0; j < 6; ++j) Us[5 + j] = 0.0;
    for (int q = 0; q < 3; ++q) Us[11 + q] = 0.0;
    Targets t = targets_of(Us, Ts, P.Pr);
    const double ws = w * P.srate;
    for (int j = 0; j < 6; ++j) {
        Un[5 + j] = (E[5 + j] + ws * t.P[j]) / (1.0 + ws);
        Us[5 + j] = Un[5 + j];
    }
    t = targets_of(Us, Ts, P.Pr);
    for (int q = 0; q < 3; ++q) Un[11 + q] = (E[11 + q] + w * t.Qf[q]) / (1.0 + P.Pr * w);
}

template <typename T>
__global__ void __launch_bounds__(kIB)
k_ars_dist(int n, dgfs_ars_t P) {
    const int s = blockIdx.y, v = blockIdx.x * kIB + threadIdx.x;
    if (v >= n) return;
    const int ns = P.stage;
    const size_t o = size_t(s) * n + v;
    const size_t b = size_t(s) * kNA;
    double acc = double(static_cast<const T*>(P.F[0])[o]);
    for (int i = 0; i < ns; ++i) acc += P.ah[i] * P.dt * double(static_cast<const T*>(P.L[i])[o]);
    for (int i = 1; i < ns; ++i) {
        const double* Ui = P.U[i] + b;
        const double nu = P.prefac * Ui[0] * pow(temp_of(Ui), 1.0 - P.omega);
        acc += P.ai[i] * P.dt * nu *
               (double(static_cast<const T*>(P.M[i])[o]) - double(static_cast<const T*>(P.F[i])[o]));
    }
    const double* Us = P.U[ns] + b;
    const double nus = P.ai[ns] * P.dt * P.prefac * Us[0] * pow(temp_of(Us), 1.0 - P.omega);
    static_cast<T*>(P.F[ns])[o] =
        T((acc + nus * double(static_cast<const T*>(P.M[ns])[o])) / (1.0 + nus));
}

// ------------------------------------------------------------------ C ABI
extern "C" size_t dgfs_mom14_work_bytes(int nslice, int Nv) {
    const int nb = (Nv + kIB - 1) / kIB;
    return size_t(nslice) * nb * kNA * sizeof(double);
}

extern "C" int dgfs_mom14(int dtype, int nslice, int Nv, const void* f, const void* cx,
                          const void* cy, const void* cz, double cw, double* work, double* U,
                          hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(nslice >= 0 && nslice <= 65535 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(f && cx && cy && cz && work && U, "null pointer");
    if (nslice == 0) return 0;
    const int nb = (Nv + kIB - 1) / kIB;
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_mom14_partial<double>, dim3(nb, nslice), dim3(kIB), 0, stream, Nv,
                           (const double*)f, (const double*)cx, (const double*)cy,
                           (const double*)cz, work);
    else
        hipLaunchKernelGGL(k_mom14_partial<float>, dim3(nb, nslice), dim3(kIB), 0, stream, Nv,
                           (const float*)f, (const float*)cx, (const float*)cy, (const float*)cz,
                           work);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_mom14_final, dim3((nslice * kNA + kIB - 1) / kIB), dim3(kIB), 0, stream,
                       nslice, nb, cw, (const double*)work, U);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_shakov_target(int dtype, int nslice, int Nv, const double* U, const void* cx,
                                  const void* cy, const void* cz, double Pr, void* M,
                                  hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(nslice >= 0 && nslice <= 65535 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(U && cx && cy && cz && M, "null pointer");
    if (nslice == 0) return 0;
    const dim3 g((Nv + kIB - 1) / kIB, nslice);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_target<double>, g, dim3(kIB), 0, stream, Nv, U, (const double*)cx,
                           (const double*)cy, (const double*)cz, Pr, (double*)M);
    else
        hipLaunchKernelGGL(k_target<float>, g, dim3(kIB), 0, stream, Nv, U, (const float*)cx,
                           (const float*)cy, (const float*)cz, Pr, (float*)M);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_esbgk_target(int dtype, int nslice, int Nv, const double* U, const void* cx,
                                 const void* cy, const void* cz, double Pr, void* M,
                                 hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(nslice >= 0 && nslice <= 65535 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(Pr > 0.0, "esbgk_target: Pr must be positive");
    if (nslice == 0) return 0;
    DGFS_REQUIRE(U && cx && cy && cz && M, "null pointer");
    const dim3 g((Nv + kIB - 1) / kIB, nslice);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_es_target<double>, g, dim3(kIB), 0, stream, Nv, U,
                           (const double*)cx, (const double*)cy, (const double*)cz, Pr,
                           (double*)M);
    else
        hipLaunchKernelGGL(k_es_target<float>, g, dim3(kIB), 0, stream, Nv, U,
                           (const float*)cx, (const float*)cy, (const float*)cz, Pr, (float*)M);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_relax_collide(int dtype, int nslice, int Nv, const double* U, const void* M,
                                  const void* f, double prefac, double omega, void* Q,
                                  hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(nslice >= 0 && nslice <= 65535 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(U && M && f && Q, "null pointer");
    if (nslice == 0) return 0;
    const dim3 g((Nv + kIB - 1) / kIB, nslice);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_collide<double>, g, dim3(kIB), 0, stream, Nv, U, (const double*)M,
                           (const double*)f, prefac, omega, (double*)Q);
    else
        hipLaunchKernelGGL(k_collide<float>, g, dim3(kIB), 0, stream, Nv, U, (const float*)M,
                           (const float*)f, prefac, omega, (float*)Q);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_imex_lm_update(int dtype, int nslice, int Nv, const dgfs_lm_t* args,
                                   int what, hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(args && args->nstages >= 1 && args->nstages <= DGFS_IMEX_MAXS, "bad stages");
    DGFS_REQUIRE(nslice >= 0 && nslice <= 65535 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(what == 0 || what == 1, "what must be 0 (moments) or 1 (distribution)");
    if (nslice == 0) return 0;
    const dgfs_lm_t P = *args;
    DGFS_REQUIRE(P.srate > 0.0, "srate must be positive");
    DGFS_REQUIRE(P.Unew, "null U_s");
    for (int i = 0; i < P.nstages; ++i) {
        if (what == 0) DGFS_REQUIRE(P.U[i] && P.LU[i] && P.LM[i], "null moment history");
        else DGFS_REQUIRE(P.F[i] && P.L[i] && P.Pn[i], "null distribution history");
    }
    if (what == 0) {
        hipLaunchKernelGGL(k_lm_moments, dim3((nslice + kIB - 1) / kIB), dim3(kIB), 0, stream,
                           nslice, P);
    } else {
        DGFS_REQUIRE(P.fnew && P.Mnew, "null f_s / M_s");
        const dim3 g((Nv + kIB - 1) / kIB, nslice);
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_lm_dist<double>, g, dim3(kIB), 0, stream, Nv, P);
        else
            hipLaunchKernelGGL(k_lm_dist<float>, g, dim3(kIB), 0, stream, Nv, P);
    }
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_imex_ars_update(int dtype, int nslice, int Nv, const dgfs_ars_t* args,
                                    int what, hipStream_t stream) {
    DGFS_REQUIRE(dtype == DGFS_F64 || dtype == DGFS_F32, "dtype must be DGFS_F64 or DGFS_F32");
    DGFS_REQUIRE(args && args->stage >= 1 && args->stage < DGFS_IMEX_MAXS, "bad stage");
    DGFS_REQUIRE(nslice >= 0 && nslice <= 65535 && Nv > 0, "bad sizes");
    DGFS_REQUIRE(what == 0 || what == 1, "what must be 0 (moments) or 1 (distribution)");
    if (nslice == 0) return 0;
    const dgfs_ars_t P = *args;
    DGFS_REQUIRE(P.srate > 0.0, "srate must be positive");
    for (int i = 0; i <= P.stage; ++i) DGFS_REQUIRE(P.U[i], "null moment stage");
    for (int i = 0; i < P.stage; ++i) DGFS_REQUIRE(P.LU[i] || what == 1, "null LU");
    if (what == 0) {
        hipLaunchKernelGGL(k_ars_moments, dim3((nslice + kIB - 1) / kIB), dim3(kIB), 0, stream,
                           nslice, P);
    } else {
        for (int i = 0; i <= P.stage; ++i) DGFS_REQUIRE(P.F[i], "null distribution stage");
        for (int i = 0; i < P.stage; ++i) DGFS_REQUIRE(P.L[i], "null explicit stage");
        for (int i = 1; i <= P.stage; ++i) DGFS_REQUIRE(P.M[i], "null target stage");
        const dim3 g((Nv + kIB - 1) / kIB, nslice);
        if (dtype == DGFS_F64)
            hipLaunchKernelGGL(k_ars_dist<double>, g, dim3(kIB), 0, stream, Nv, P);
        else
            hipLaunchKernelGGL(k_ars_dist<float>, g, dim3(kIB), 0, stream, Nv, P);
    }
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}
