// Learned collision closure: a 4-layer 3D convolution network on the
// velocity grid (1 -> 32 -> 32 -> 32 -> 1 channels, 3x3x3 kernels, zero
// padding, SiLU after the first three layers) that predicts the residual
// Q(f, f) - Q_Shakov(f).  No reference counterpart: its targets are this
// library's own spectral operator minus the Shakov model
// (dgfs1D/std/scattering.py:1636-1833).
//
// Layout: activations are channels-last bf16, act[b][z][y][x][c] with c
// fastest (32 channels = 64 B per voxel); b indexes velocity slices (nodes);
// the grid is G^3 with G % 16 == 0.  Weights: master FP32 copies
//   W1 [27][32], W2, W3 [27][32 out][32 in], W4 [27][32], biases [32]/[1]
// with tap = (dz+1)*9 + (dy+1)*3 + (dx+1) and out[v] = sum W[tap] in[v + d(tap)].
//
// The two 32 -> 32 layers (96 % of the arithmetic) run on the bf16 matrix
// cores: v_mfma_f32_16x16x32_bf16 with A = weights (16 out channels x 32 in
// channels, from LDS) and B = activations (32 in channels x 16 voxels along
// x, one 16 B global load per lane), FP32 accumulation.  The same kernel
// computes the input gradient of those layers with tap-flipped, transposed
// weights.  The thin first/last layers and the weight gradients run on FP32
// VALU with FP32 accumulation; partial weight gradients are reduced in a
// fixed order (deterministic).
#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kC = 32;          // hidden channels
constexpr int kTaps = 27;
constexpr int kWP = 40;         // LDS row pitch (bf16) of a weight tile row

DGFS_DEV float silu(float z) { return z / (1.0f + __expf(-z)); }
DGFS_DEV float silu_d(float z) {
    const float s = 1.0f / (1.0f + __expf(-z));
    return s * (1.0f + z * (1.0f - s));
}
DGFS_DEV float bf2f(__bf16 h) { return float(h); }
DGFS_DEV __bf16 f2bf(float f) { return __bf16(f); }

DGFS_DEV void tap_offset(int tap, int& dz, int& dy, int& dx) {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
}

// ------------------------------------------------- 32 -> 32 layer, MFMA
// Work-group = 4 waves over 8 y-rows of one (b, z) plane; wave w owns rows
// y0 + 2w, y0 + 2w + 1 and every 16-voxel x tile of them.
// MODE 0: z = acc + bias -> store Z (bf16) and Y = silu(z) (bf16)
// MODE 1: z = acc + bias -> store Y = z (bf16)
// MODE 2: (input gradient) store Y = acc * silu'(Zp)  (Zp: pre-activation of
//         the layer below, bf16), no bias
template <int G, int MODE>
__global__ void __launch_bounds__(256)
k_conv_mid(int nb, const __bf16* __restrict__ X, const __bf16* __restrict__ W /*[27][32][32]*/,
           const float* __restrict__ bias, const __bf16* __restrict__ Zp,
           __bf16* __restrict__ Zout, __bf16* __restrict__ Y) {
    constexpr int NXT = G / 16;             // x tiles per row
    constexpr int NT = 2 * NXT;             // voxel tiles per wave
    constexpr int YB = G / 8;               // y blocks per plane
    __shared__ __attribute__((aligned(16))) __bf16 Ws[kTaps * kC * kWP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = blockIdx.x;
    const int yb = bid % YB, z = (bid / YB) % G, b = bid / (YB * G);
    // stage the weights (27*32 rows of 32 bf16 = 4 x 16 B chunks each)
    for (int r = tid; r < kTaps * kC * 4; r += 256) {
        const int row = r >> 2, ch = r & 3;
        *reinterpret_cast<uint4*>(&Ws[row * kWP + ch * 8]) =
            *reinterpret_cast<const uint4*>(&W[row * kC + ch * 8]);
    }
    __syncthreads();
    if (b >= nb) return;
    const int kq = lane >> 4, li = lane & 15;
    floatx4 acc[NT][2];
    for (int t = 0; t < NT; ++t)
        for (int j = 0; j < 2; ++j) acc[t][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    const size_t plane = size_t(G) * G;
    const __bf16* Xb = X + size_t(b) * G * plane * kC;
    for (int tap = 0; tap < kTaps; ++tap) {
        int dz, dy, dx;
        tap_offset(tap, dz, dy, dx);
        const int zz = z + dz;
        bf16x8 a[2];
        for (int j = 0; j < 2; ++j)
            a[j] = *reinterpret_cast<const bf16x8*>(
                &Ws[(tap * kC + 16 * j + li) * kWP + 8 * kq]);
        for (int t = 0; t < NT; ++t) {
            const int y = yb * 8 + 2 * wave + t / NXT;
            const int x = (t % NXT) * 16 + li + dx;
            const int yy = y + dy;
            bf16x8 bv;
            if (zz >= 0 && zz < G && yy >= 0 && yy < G && x >= 0 && x < G) {
                bv = *reinterpret_cast<const bf16x8*>(
                    &Xb[((size_t(zz) * G + yy) * G + x) * kC + 8 * kq]);
            } else {
                for (int i = 0; i < 8; ++i) bv[i] = __bf16(0.0f);
            }
            for (int j = 0; j < 2; ++j)
                acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j], bv, acc[t][j], 0, 0, 0);
        }
    }
    // epilogue: lane holds out channels 16j + 4kq + i of voxel (tile base + li)
    for (int t = 0; t < NT; ++t) {
        const int y = yb * 8 + 2 * wave + t / NXT;
        const int x = (t % NXT) * 16 + li;
        const size_t vo = ((size_t(b) * G + z) * plane + size_t(y) * G + x) * kC;
        for (int j = 0; j < 2; ++j) {
            const int c0 = 16 * j + 4 * kq;
            __bf16 yo[4], zo[4];
            for (int i = 0; i < 4; ++i) {
                const float v = acc[t][j][i];
                if (MODE == 0) {
                    const float zv = v + bias[c0 + i];
                    zo[i] = f2bf(zv);
                    yo[i] = f2bf(silu(zv));
                } else if (MODE == 1) {
                    yo[i] = f2bf(v + bias[c0 + i]);
                } else {
                    yo[i] = f2bf(v * silu_d(bf2f(Zp[vo + c0 + i])));
                }
            }
            if (MODE == 0) {
                uint2 pz;
                __builtin_memcpy(&pz, zo, 8);
                *reinterpret_cast<uint2*>(&Zout[vo + c0]) = pz;
            }
            uint2 py;
            __builtin_memcpy(&py, yo, 8);
            *reinterpret_cast<uint2*>(&Y[vo + c0]) = py;
        }
    }
}

// Same layer with the input tile staged in LDS: the work-group's 3 x 10 x
// (G + 2) halo of voxels (zero padded) is loaded once, so each activation is
// read from HBM/L2 once per work-group instead of once per tap.  Voxel pitch
// 40 bf16 (80 B) keeps the 16 lanes of a k-group on disjoint banks.
// LDS: 27*32*40*2 + 3*10*(G+2)*40*2 bytes (150.7 KiB at G = 32).
template <int G, int MODE>
__global__ void __launch_bounds__(256)
k_conv_mid_lds(int nb, const __bf16* __restrict__ X, const __bf16* __restrict__ W,
               const float* __restrict__ bias, const __bf16* __restrict__ Zp,
               __bf16* __restrict__ Zout, __bf16* __restrict__ Y) {
    constexpr int NXT = G / 16, NT = 2 * NXT, YB = G / 8;
    constexpr int XP = G + 2, XR = 10, VP = kWP;          // halo row length, rows, voxel pitch
    __shared__ __attribute__((aligned(16))) __bf16 Ws[kTaps * kC * kWP];
    __shared__ __attribute__((aligned(16))) __bf16 Xs[3 * XR * XP * VP];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int bid = blockIdx.x;
    const int yb = bid % YB, z = (bid / YB) % G, b = bid / (YB * G);
    for (int r = tid; r < kTaps * kC * 4; r += 256) {
        const int row = r >> 2, ch = r & 3;
        *reinterpret_cast<uint4*>(&Ws[row * kWP + ch * 8]) =
            *reinterpret_cast<const uint4*>(&W[row * kC + ch * 8]);
    }
    const size_t plane = size_t(G) * G;
    const __bf16* Xb = X + size_t(b) * G * plane * kC;
    for (int r = tid; r < 3 * XR * XP * 4; r += 256) {
        const int ch = r & 3, vox = r >> 2;
        const int xi = vox % XP, yi = (vox / XP) % XR, zi = vox / (XP * XR);
        const int zz = z + zi - 1, yy = yb * 8 + yi - 1, xx = xi - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (zz >= 0 && zz < G && yy >= 0 && yy < G && xx >= 0 && xx < G)
            v = *reinterpret_cast<const uint4*>(&Xb[((size_t(zz) * G + yy) * G + xx) * kC + ch * 8]);
        *reinterpret_cast<uint4*>(&Xs[vox * VP + ch * 8]) = v;
    }
    __syncthreads();
    if (b >= nb) return;
    const int kq = lane >> 4, li = lane & 15;
    floatx4 acc[NT][2];
    for (int t = 0; t < NT; ++t)
        for (int j = 0; j < 2; ++j) acc[t][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int tap = 0; tap < kTaps; ++tap) {
        int dz, dy, dx;
        tap_offset(tap, dz, dy, dx);
        bf16x8 a[2];
        for (int j = 0; j < 2; ++j)
            a[j] = *reinterpret_cast<const bf16x8*>(
                &Ws[(tap * kC + 16 * j + li) * kWP + 8 * kq]);
        for (int t = 0; t < NT; ++t) {
            const int yl = 2 * wave + t / NXT;                 // row within the block
            const int xi = (t % NXT) * 16 + li + dx + 1;       // halo column
            const int vox = ((dz + 1) * XR + (yl + dy + 1)) * XP + xi;
            const bf16x8 bv = *reinterpret_cast<const bf16x8*>(&Xs[vox * VP + 8 * kq]);
            for (int j = 0; j < 2; ++j)
                acc[t][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[j], bv, acc[t][j], 0, 0, 0);
        }
    }
    for (int t = 0; t < NT; ++t) {
        const int y = yb * 8 + 2 * wave + t / NXT;
        const int x = (t % NXT) * 16 + li;
        const size_t vo = ((size_t(b) * G + z) * plane + size_t(y) * G + x) * kC;
        for (int j = 0; j < 2; ++j) {
            const int c0 = 16 * j + 4 * kq;
            __bf16 yo[4], zo[4];
            for (int i = 0; i < 4; ++i) {
                const float v = acc[t][j][i];
                if (MODE == 0) {
                    const float zv = v + bias[c0 + i];
                    zo[i] = f2bf(zv);
                    yo[i] = f2bf(silu(zv));
                } else if (MODE == 1) {
                    yo[i] = f2bf(v + bias[c0 + i]);
                } else {
                    yo[i] = f2bf(v * silu_d(bf2f(Zp[vo + c0 + i])));
                }
            }
            if (MODE == 0) {
                uint2 pz;
                __builtin_memcpy(&pz, zo, 8);
                *reinterpret_cast<uint2*>(&Zout[vo + c0]) = pz;
            }
            uint2 py;
            __builtin_memcpy(&py, yo, 8);
            *reinterpret_cast<uint2*>(&Y[vo + c0]) = py;
        }
    }
}

// ------------------------------------------------------- 1 -> 32, VALU
// f (solver dtype) of slice src[b] scaled by scale[b]; stores Z1 and Y1.
template <typename T>
__global__ void __launch_bounds__(256)
k_conv_in(int G, long nvox, const T* __restrict__ u, const int* __restrict__ src,
          const float* __restrict__ scale, const float* __restrict__ W1 /*[27][32]*/,
          const float* __restrict__ b1, __bf16* __restrict__ Z, __bf16* __restrict__ Y) {
    __shared__ float Ws[kTaps * kC + kC];
    for (int i = threadIdx.x; i < kTaps * kC; i += 256) Ws[i] = W1[i];
    for (int i = threadIdx.x; i < kC; i += 256) Ws[kTaps * kC + i] = b1[i];
    __syncthreads();
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= nvox) return;
    const int n = G * G * G;
    const int b = int(gid / n), v = int(gid - long(b) * n);
    const int x = v % G, y = (v / G) % G, z = v / (G * G);
    const T* f = u + size_t(src[b]) * n;
    const float s = scale[b];
    float acc[kC];
    for (int c = 0; c < kC; ++c) acc[c] = Ws[kTaps * kC + c];
    for (int tap = 0; tap < kTaps; ++tap) {
        int dz, dy, dx;
        tap_offset(tap, dz, dy, dx);
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        if (zz < 0 || zz >= G || yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
        const float fv = s * float(f[(zz * G + yy) * G + xx]);
        for (int c = 0; c < kC; ++c) acc[c] += fv * Ws[tap * kC + c];
    }
    __bf16 zo[kC], yo[kC];
    for (int c = 0; c < kC; ++c) {
        zo[c] = f2bf(acc[c]);
        yo[c] = f2bf(silu(acc[c]));
    }
    uint4* zp = reinterpret_cast<uint4*>(&Z[size_t(gid) * kC]);
    uint4* yp = reinterpret_cast<uint4*>(&Y[size_t(gid) * kC]);
    for (int q = 0; q < 4; ++q) {
        uint4 t1, t2;
        __builtin_memcpy(&t1, &zo[8 * q], 16);
        __builtin_memcpy(&t2, &yo[8 * q], 16);
        zp[q] = t1;
        yp[q] = t2;
    }
}

// ------------------------------------------------------- 32 -> 1, VALU
// out[b][v] = b4 + sum W4 . Y3; mode 0: store FP32 prediction pred[b][v];
// mode 1: Q[dst[b]][v] += oscale[b] * out (solver dtype)
template <typename T>
__global__ void __launch_bounds__(256)
k_conv_out(int G, long nvox, const __bf16* __restrict__ Y3, const float* __restrict__ W4,
           const float* __restrict__ b4, int mode, float* __restrict__ pred,
           const int* __restrict__ dst, const float* __restrict__ oscale, T* __restrict__ Q) {
    __shared__ float Ws[kTaps * kC];
    for (int i = threadIdx.x; i < kTaps * kC; i += 256) Ws[i] = W4[i];
    __syncthreads();
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= nvox) return;
    const int n = G * G * G;
    const int b = int(gid / n), v = int(gid - long(b) * n);
    const int x = v % G, y = (v / G) % G, z = v / (G * G);
    const __bf16* Yb = Y3 + size_t(b) * n * kC;
    float acc = b4[0];
    for (int tap = 0; tap < kTaps; ++tap) {
        int dz, dy, dx;
        tap_offset(tap, dz, dy, dx);
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        if (zz < 0 || zz >= G || yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
        const uint4* p = reinterpret_cast<const uint4*>(&Yb[size_t((zz * G + yy) * G + xx) * kC]);
        for (int q = 0; q < 4; ++q) {
            const uint4 w = p[q];
            __bf16 h[8];
            __builtin_memcpy(h, &w, 16);
            for (int i = 0; i < 8; ++i) acc += bf2f(h[i]) * Ws[tap * kC + 8 * q + i];
        }
    }
    if (mode == 0) pred[gid] = acc;
    else {
        T* q = Q + size_t(dst[b]) * n + v;
        *q = T(float(*q) + oscale[b] * acc);
    }
}

// d(out)/d(Y3) then times silu'(Z3): dZ3[b][v][c] = silu'(Z3) sum_tap dOut[v - d] W4[tap][c]
__global__ void __launch_bounds__(256)
k_out_bwd(int G, long nvox, const float* __restrict__ dOut, const float* __restrict__ W4,
          const __bf16* __restrict__ Z3, __bf16* __restrict__ dZ3) {
    __shared__ float Ws[kTaps * kC];
    for (int i = threadIdx.x; i < kTaps * kC; i += 256) Ws[i] = W4[i];
    __syncthreads();
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= nvox) return;
    const int n = G * G * G;
    const int b = int(gid / n), v = int(gid - long(b) * n);
    const int x = v % G, y = (v / G) % G, z = v / (G * G);
    float acc[kC];
    for (int c = 0; c < kC; ++c) acc[c] = 0.f;
    for (int tap = 0; tap < kTaps; ++tap) {
        int dz, dy, dx;
        tap_offset(tap, dz, dy, dx);
        const int zz = z - dz, yy = y - dy, xx = x - dx;
        if (zz < 0 || zz >= G || yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
        const float g = dOut[size_t(b) * n + (zz * G + yy) * G + xx];
        for (int c = 0; c < kC; ++c) acc[c] += g * Ws[tap * kC + c];
    }
    const uint4* zp = reinterpret_cast<const uint4*>(&Z3[size_t(gid) * kC]);
    uint4* op = reinterpret_cast<uint4*>(&dZ3[size_t(gid) * kC]);
    for (int q = 0; q < 4; ++q) {
        const uint4 w = zp[q];
        __bf16 h[8], o[8];
        __builtin_memcpy(h, &w, 16);
        for (int i = 0; i < 8; ++i) o[i] = f2bf(acc[8 * q + i] * silu_d(bf2f(h[i])));
        uint4 t;
        __builtin_memcpy(&t, o, 16);
        op[q] = t;
    }
}

// ------------------------------------------------ weight gradients (VALU)
// part[chunk][tap][co][ci] = sum_{v in chunk} G[v][co] X[v + d(tap)][ci]
// (G, X channels-last bf16 with 32 channels); block (chunk, tap), 64 threads,
// each a 4x4 block of (co, ci).  Tap 13 (centre) also sums G for the bias.
__global__ void __launch_bounds__(64)
k_wgrad_mid(int G, int nb, int chunk, const __bf16* __restrict__ Gr, const __bf16* __restrict__ X,
            float* __restrict__ part, float* __restrict__ bpart) {
    const int c = blockIdx.x, tap = blockIdx.y, t = threadIdx.x;
    const int co0 = (t >> 3) * 4, ci0 = (t & 7) * 4;
    int dz, dy, dx;
    tap_offset(tap, dz, dy, dx);
    const int n = G * G * G;
    const long total = long(nb) * n;
    const long v0 = long(c) * chunk;
    const long v1 = v0 + chunk < total ? v0 + chunk : total;
    float acc[4][4] = {};
    float bacc[4] = {0.f, 0.f, 0.f, 0.f};
    for (long gv = v0; gv < v1; ++gv) {
        const int b = int(gv / n), v = int(gv - long(b) * n);
        const int x = v % G, y = (v / G) % G, z = v / (G * G);
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        __bf16 gh[4];
        const uint2 gw = *reinterpret_cast<const uint2*>(&Gr[size_t(gv) * kC + co0]);
        __builtin_memcpy(gh, &gw, 8);
        if (tap == 13)
            for (int i = 0; i < 4; ++i) bacc[i] += bf2f(gh[i]);
        if (zz < 0 || zz >= G || yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
        __bf16 xh[4];
        const uint2 xw = *reinterpret_cast<const uint2*>(
            &X[(size_t(b) * n + (zz * G + yy) * G + xx) * kC + ci0]);
        __builtin_memcpy(xh, &xw, 8);
        for (int i = 0; i < 4; ++i)
            for (int k = 0; k < 4; ++k) acc[i][k] += bf2f(gh[i]) * bf2f(xh[k]);
    }
    float* p = part + (size_t(c) * kTaps + tap) * kC * kC;
    for (int i = 0; i < 4; ++i)
        for (int k = 0; k < 4; ++k) p[(co0 + i) * kC + ci0 + k] = acc[i][k];
    if (tap == 13 && ci0 == 0)
        for (int i = 0; i < 4; ++i) bpart[size_t(c) * kC + co0 + i] = bacc[i];
}

// thin layers: part[chunk][tap][c] = sum_v g[v] x[v + d][c] where exactly one
// operand has one channel.  GC1 = true: g is FP32 with one channel (layer 4:
// g = dOut, x = Y3 bf16); false: g = dZ1 bf16 32 channels, x = the scaled
// input f (FP32 one channel, layer 1).  Threads own one channel each.
template <bool GC1>
__global__ void __launch_bounds__(64)
k_wgrad_thin(int G, int nb, int chunk, const float* __restrict__ g1, const __bf16* __restrict__ g32,
             const float* __restrict__ x1, const __bf16* __restrict__ x32,
             float* __restrict__ part, float* __restrict__ bpart) {
    const int c = blockIdx.x, tap = blockIdx.y, t = threadIdx.x;
    if (t >= kC) return;
    int dz, dy, dx;
    tap_offset(tap, dz, dy, dx);
    const int n = G * G * G;
    const long total = long(nb) * n;
    const long v0 = long(c) * chunk;
    const long v1 = v0 + chunk < total ? v0 + chunk : total;
    float acc = 0.f, bacc = 0.f;
    for (long gv = v0; gv < v1; ++gv) {
        const int b = int(gv / n), v = int(gv - long(b) * n);
        const int x = v % G, y = (v / G) % G, z = v / (G * G);
        const float gval = GC1 ? g1[gv] : bf2f(g32[size_t(gv) * kC + t]);
        if (tap == 13) bacc += gval;
        const int zz = z + dz, yy = y + dy, xx = x + dx;
        if (zz < 0 || zz >= G || yy < 0 || yy >= G || xx < 0 || xx >= G) continue;
        const size_t o = size_t(b) * n + (zz * G + yy) * G + xx;
        const float xv = GC1 ? bf2f(x32[o * kC + t]) : x1[o];
        acc += gval * xv;
    }
    part[(size_t(c) * kTaps + tap) * kC + t] = acc;
    if (tap == 13) {
        if (GC1) {
            if (t == 0) bpart[c] = bacc;               // one bias: [chunk]
        } else {
            bpart[size_t(c) * kC + t] = bacc;          // [chunk][32]
        }
    }
}

// out[i] = sum_c part[c][i] (fixed order), i < len
__global__ void __launch_bounds__(256)
k_reduce_chunks(int nchunk, int len, const float* __restrict__ part, float* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= len) return;
    float acc = 0.f;
    for (int c = 0; c < nchunk; ++c) acc += part[size_t(c) * len + i];
    out[i] = acc;
}

// scaled input of layer 1 as FP32 (kept for its weight gradient)
template <typename T>
__global__ void __launch_bounds__(256)
k_gather_scaled(int n, long nvox, const T* __restrict__ u, const int* __restrict__ src,
                const float* __restrict__ scale, float* __restrict__ out) {
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= nvox) return;
    const int b = int(gid / n), v = int(gid - long(b) * n);
    out[gid] = scale[b] * float(u[size_t(src[b]) * n + v]);
}

// master FP32 [27][32 out][32 in] -> bf16 forward copy and the tap-flipped,
// transposed copy used for the input gradient: Wt[tap][in][out] = W[26-tap][out][in]
__global__ void __launch_bounds__(256)
k_pack_mid(const float* __restrict__ W, __bf16* __restrict__ Wf, __bf16* __restrict__ Wt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= kTaps * kC * kC) return;
    const int tap = i / (kC * kC), r = (i / kC) % kC, c = i % kC;
    Wf[i] = f2bf(W[i]);
    Wt[((kTaps - 1 - tap) * kC + c) * kC + r] = f2bf(W[i]);
}

// relative squared error per slice and its gradient, relative to a
// reference field r (the target itself, or the full collision integral):
// loss_b = |p - t|^2 / max(|r|^2, floor2), dOut = 2 (p - t) / (max(..) nb);
// the floor keeps near-equilibrium slices (|r| -> 0) from dominating.
__global__ void __launch_bounds__(256)
k_loss_partial(int n, const float* __restrict__ p, const float* __restrict__ t,
               const float* __restrict__ ref, double* __restrict__ part /*[b][nblk][2]*/) {
    __shared__ double sh[2 * 256];
    const int b = blockIdx.y, nblk = gridDim.x;
    const int v = blockIdx.x * 256 + threadIdx.x;
    double e = 0.0, s = 0.0;
    if (v < n) {
        const double d = double(p[size_t(b) * n + v]) - double(t[size_t(b) * n + v]);
        const double tt = ref[size_t(b) * n + v];
        e = d * d;
        s = tt * tt;
    }
    sh[threadIdx.x] = e;
    sh[256 + threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if (threadIdx.x < k) {
            sh[threadIdx.x] += sh[threadIdx.x + k];
            sh[256 + threadIdx.x] += sh[256 + threadIdx.x + k];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[(size_t(b) * nblk + blockIdx.x) * 2] = sh[0];
        part[(size_t(b) * nblk + blockIdx.x) * 2 + 1] = sh[256];
    }
}

__global__ void __launch_bounds__(256)
k_loss_grad(int n, int nb, int nblk, double floor2, const double* __restrict__ part,
            const float* __restrict__ p, const float* __restrict__ t, float* __restrict__ dOut,
            double* __restrict__ loss) {
    const int b = blockIdx.y;
    double e = 0.0, s = 0.0;
    for (int k = 0; k < nblk; ++k) {
        e += part[(size_t(b) * nblk + k) * 2];
        s += part[(size_t(b) * nblk + k) * 2 + 1];
    }
    s = s > floor2 ? s : floor2;
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) loss[b] = e / s;
    if (v >= n) return;
    const size_t o = size_t(b) * n + v;
    dOut[o] = float(2.0 * (double(p[o]) - double(t[o])) / (s * double(nb)));
}

// Adam on FP32 master weights (m, v moments), bias-corrected step size lr_t
__global__ void __launch_bounds__(256)
k_adam(int n, float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
       float* __restrict__ v, float lr_t, float b1, float b2, float eps) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr_t * mi / (sqrtf(vi) + eps);
}

}  // namespace
}  // namespace dgfs

using namespace dgfs;

namespace {

int dispatch_mid(int G, int mode, int nb, const void* X, const void* W, const float* bias,
                 const void* Zp, void* Zout, void* Y, hipStream_t st) {
    const dim3 grid(nb * G * (G / 8));
    const __bf16* x = (const __bf16*)X;
    const __bf16* w = (const __bf16*)W;
    const __bf16* zp = (const __bf16*)Zp;
    __bf16* zo = (__bf16*)Zout;
    __bf16* y = (__bf16*)Y;
#define DGFS_MID(GG, KER)                                                                     \
    if (G == GG) {                                                                            \
        if (mode == 0)                                                                        \
            hipLaunchKernelGGL((KER<GG, 0>), grid, dim3(256), 0, st, nb, x, w, bias, zp, zo,  \
                               y);                                                            \
        else if (mode == 1)                                                                   \
            hipLaunchKernelGGL((KER<GG, 1>), grid, dim3(256), 0, st, nb, x, w, bias, zp, zo,  \
                               y);                                                            \
        else                                                                                  \
            hipLaunchKernelGGL((KER<GG, 2>), grid, dim3(256), 0, st, nb, x, w, bias, zp, zo,  \
                               y);                                                            \
        DGFS_HIP_CHECK(hipGetLastError());                                                    \
        return 0;                                                                             \
    }
    // LDS-staged input where the tile fits (G <= 32); direct loads at G = 48
    DGFS_MID(16, k_conv_mid_lds)
    DGFS_MID(32, k_conv_mid_lds)
    DGFS_MID(48, k_conv_mid)
#undef DGFS_MID
    DGFS_REQUIRE(false, "closure: grid size must be 16, 32 or 48");
}

}  // namespace

extern "C" size_t dgfs_closure_work_bytes(int G, int nb, int chunk) {
    const long total = long(nb) * G * G * G;
    const long nchunk = (total + chunk - 1) / chunk;
    return size_t(nchunk) * (kTaps * kC * kC + kC) * sizeof(float);
}

extern "C" int dgfs_closure_forward(const dgfs_closure_t* net, int dtype, int G, int nb,
                                    const void* u, const int* src, const float* scale,
                                    void* act, int mode, float* pred, const int* dst,
                                    const float* oscale, void* Q, hipStream_t st) {
    DGFS_REQUIRE(net && act && u && src && scale, "closure_forward: null pointer");
    DGFS_REQUIRE(G == 16 || G == 32 || G == 48, "closure: grid size must be 16, 32 or 48");
    DGFS_REQUIRE(nb >= 0 && (mode == 0 || mode == 1), "closure_forward: bad arguments");
    DGFS_REQUIRE(mode == 1 || pred, "closure_forward: mode 0 needs pred");
    DGFS_REQUIRE(mode == 0 || (dst && oscale && Q), "closure_forward: mode 1 needs dst, oscale, Q");
    if (nb == 0) return 0;
    const long nvox = long(nb) * G * G * G;
    DGFS_REQUIRE(nvox / 256 < 2147483647L && long(nb) * G * (G / 8) < 2147483647L, "too large");
    const size_t a = size_t(nvox) * kC;         // one activation tensor (bf16 elements)
    __bf16* A = (__bf16*)act;                   // [Z1 Y1 Z2 Y2 Z3 Y3], 6 tensors
    const dim3 gv((nvox + 255) / 256);
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_conv_in<double>, gv, dim3(256), 0, st, G, nvox, (const double*)u, src,
                           scale, net->W1, net->b1, A, A + a);
    else
        hipLaunchKernelGGL(k_conv_in<float>, gv, dim3(256), 0, st, G, nvox, (const float*)u, src,
                           scale, net->W1, net->b1, A, A + a);
    DGFS_HIP_CHECK(hipGetLastError());
    int rc = dispatch_mid(G, 0, nb, A + a, net->W2f, net->b2, nullptr, A + 2 * a, A + 3 * a, st);
    if (rc) return rc;
    rc = dispatch_mid(G, 0, nb, A + 3 * a, net->W3f, net->b3, nullptr, A + 4 * a, A + 5 * a, st);
    if (rc) return rc;
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_conv_out<double>, gv, dim3(256), 0, st, G, nvox, A + 5 * a, net->W4,
                           net->b4, mode, pred, dst, oscale, (double*)Q);
    else
        hipLaunchKernelGGL(k_conv_out<float>, gv, dim3(256), 0, st, G, nvox, A + 5 * a, net->W4,
                           net->b4, mode, pred, dst, oscale, (float*)Q);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_closure_backward(const dgfs_closure_t* net, const dgfs_closure_grad_t* gr,
                                     int dtype, int G, int nb, const void* u, const int* src,
                                     const float* scale, const void* act, const float* dOut,
                                     void* gact, float* fin, int chunk, float* work,
                                     hipStream_t st) {
    DGFS_REQUIRE(net && gr && act && dOut && gact && fin && work && u && src && scale,
                 "closure_backward: null pointer");
    DGFS_REQUIRE(G == 16 || G == 32 || G == 48, "closure: grid size must be 16, 32 or 48");
    DGFS_REQUIRE(chunk > 0 && nb >= 0, "closure_backward: bad arguments");
    if (nb == 0) return 0;
    const int n = G * G * G;
    const long nvox = long(nb) * n;
    const long nch = (nvox + chunk - 1) / chunk;
    DGFS_REQUIRE(nch <= 65535, "closure_backward: chunk too small for this batch");
    const size_t a = size_t(nvox) * kC;
    const __bf16* A = (const __bf16*)act;       // Z1 Y1 Z2 Y2 Z3 Y3
    __bf16* GA = (__bf16*)gact;                 // two gradient tensors (ping-pong)
    const dim3 gv((nvox + 255) / 256);
    const int nc = int(nch);
    float* part = work;
    float* bpart = work + size_t(nc) * kTaps * kC * kC;
    // layer 4: dW4, db4 and dZ3
    hipLaunchKernelGGL(k_wgrad_thin<true>, dim3(nc, kTaps), dim3(64), 0, st, G, nb, chunk, dOut,
                       (const __bf16*)nullptr, (const float*)nullptr, A + 5 * a, part, bpart);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC + 255) / 256), dim3(256), 0, st, nc,
                       kTaps * kC, part, gr->W4);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, nc, 1, bpart, gr->b4);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_out_bwd, gv, dim3(256), 0, st, G, nvox, dOut, net->W4, A + 4 * a, GA);
    DGFS_HIP_CHECK(hipGetLastError());
    // layer 3: dW3 = dZ3 x Y2, dZ2 = silu'(Z2) convT(dZ3, W3)
    hipLaunchKernelGGL(k_wgrad_mid, dim3(nc, kTaps), dim3(64), 0, st, G, nb, chunk, GA, A + 3 * a,
                       part, bpart);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC * kC + 255) / 256), dim3(256), 0, st, nc,
                       kTaps * kC * kC, part, gr->W3);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, nc, kC, bpart, gr->b3);
    DGFS_HIP_CHECK(hipGetLastError());
    int rc = dispatch_mid(G, 2, nb, GA, net->W3t, nullptr, A + 2 * a, nullptr, GA + a, st);
    if (rc) return rc;
    // layer 2: dW2 = dZ2 x Y1, dZ1 = silu'(Z1) convT(dZ2, W2)
    hipLaunchKernelGGL(k_wgrad_mid, dim3(nc, kTaps), dim3(64), 0, st, G, nb, chunk, GA + a, A + a,
                       part, bpart);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC * kC + 255) / 256), dim3(256), 0, st, nc,
                       kTaps * kC * kC, part, gr->W2);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, nc, kC, bpart, gr->b2);
    DGFS_HIP_CHECK(hipGetLastError());
    rc = dispatch_mid(G, 2, nb, GA + a, net->W2t, nullptr, A, nullptr, GA, st);
    if (rc) return rc;
    // layer 1: dW1 = dZ1 x f, db1
    if (dtype == DGFS_F64)
        hipLaunchKernelGGL(k_gather_scaled<double>, gv, dim3(256), 0, st, n, nvox,
                           (const double*)u, src, scale, fin);
    else
        hipLaunchKernelGGL(k_gather_scaled<float>, gv, dim3(256), 0, st, n, nvox,
                           (const float*)u, src, scale, fin);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_wgrad_thin<false>, dim3(nc, kTaps), dim3(64), 0, st, G, nb, chunk,
                       (const float*)nullptr, GA, fin, (const __bf16*)nullptr, part, bpart);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC + 255) / 256), dim3(256), 0, st, nc,
                       kTaps * kC, part, gr->W1);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, nc, kC, bpart, gr->b1);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_closure_pack(const float* W2, const float* W3, void* W2f, void* W2t,
                                 void* W3f, void* W3t, hipStream_t st) {
    DGFS_REQUIRE(W2 && W3 && W2f && W2t && W3f && W3t, "closure_pack: null pointer");
    const dim3 g((kTaps * kC * kC + 255) / 256);
    hipLaunchKernelGGL(k_pack_mid, g, dim3(256), 0, st, W2, (__bf16*)W2f, (__bf16*)W2t);
    hipLaunchKernelGGL(k_pack_mid, g, dim3(256), 0, st, W3, (__bf16*)W3f, (__bf16*)W3t);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" size_t dgfs_closure_loss_work_bytes(int G, int nb) {
    const int n = G * G * G;
    return size_t(nb) * ((n + 255) / 256) * 2 * sizeof(double);
}

extern "C" int dgfs_closure_loss(int G, int nb, const float* pred, const float* target,
                                 const float* ref, double floor2, float* dOut, double* loss,
                                 double* work, hipStream_t st) {
    DGFS_REQUIRE(pred && target && dOut && loss && work, "closure_loss: null pointer");
    DGFS_REQUIRE(nb >= 0 && nb <= 65535 && G > 0, "closure_loss: bad sizes");
    if (nb == 0) return 0;
    const int n = G * G * G, nblk = (n + 255) / 256;
    hipLaunchKernelGGL(k_loss_partial, dim3(nblk, nb), dim3(256), 0, st, n, pred, target,
                       ref ? ref : target, work);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_loss_grad, dim3(nblk, nb), dim3(256), 0, st, n, nb, nblk, floor2, work,
                       pred, target, dOut, loss);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

extern "C" int dgfs_adam(int n, float* w, const float* g, float* m, float* v, double lr_t,
                         double beta1, double beta2, double eps, hipStream_t st) {
    DGFS_REQUIRE(n >= 0 && w && g && m && v, "adam: bad arguments");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, st, n, w, g, m, v, float(lr_t),
                       float(beta1), float(beta2), float(eps));
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}
