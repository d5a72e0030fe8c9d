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
// Forward: all four layers run on the bf16 matrix cores
// (v_mfma_f32_16x16x32_bf16, FP32 accumulation) in three z-streaming kernels:
// layers 1 + 2 fused (layer 1's K dimension is the 27 taps of the bf16-rounded
// scaled input; its output never leaves LDS), layer 3, and layer 4 (the taps
// as MFMA rows, then a shifted 27-term sum).  Weights are bf16 on the device
// (W2/W3 packed copies; W1/W4 rounded from the FP32 masters in-kernel).
// Backward: the 32 -> 32 input gradients reuse the layer-3 kernel with
// tap-flipped, transposed weights; the thin layers' gradients and all weight
// gradients run on FP32 VALU with FP32 accumulation, partial weight gradients
// reduced in a fixed order (deterministic).
#include "dgfs_common.h"
#include "dgfs_hip.h"

namespace dgfs {
namespace {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float floatx4;

constexpr int kC = 32;          // hidden channels
constexpr int kTaps = 27;
// LDS voxel pitch (bf16) of the activation planes: 96 B.  ds_read_b128 of 16
// consecutive voxels (one k-group of a B fragment) is free of bank conflicts
// at this pitch on gfx950, from any starting voxel; 64 B and 80 B pitches
// conflict (measured with the SQ_LDS_BANK_CONFLICT counter).
constexpr int kVP = 48;

// sigmoid from the hardware 2^x and reciprocal (v_exp_f32, v_rcp_f32, about
// 1 ulp each; the results are rounded to bf16 right after)
DGFS_DEV float sigm(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-1.4426950408889634f * z));
}
DGFS_DEV float silu(float z) { return z * sigm(z); }
DGFS_DEV float silu_d(float z) {
    const float s = sigm(z);
    return s * (1.0f + z * (1.0f - s));
}
DGFS_DEV float bf2f(__bf16 h) { return float(h); }
DGFS_DEV __bf16 f2bf(float f) { return __bf16(f); }

DGFS_DEV void tap_offset(int tap, int& dz, int& dy, int& dx) {
    dz = tap / 9 - 1;
    dy = (tap / 3) % 3 - 1;
    dx = tap % 3 - 1;
}

// ================================================= z-streaming tile kernels
// A task is RB rows (y) x all G columns (x) of ZC consecutive planes (z) of
// one slice.  A work-group marches through the planes of its task keeping a
// ring of three zero-padded input planes ((RB + 2) x (G + 2) voxels) in LDS:
// each input plane is fetched from HBM once per task (plus one halo plane at
// each end of the chunk and one halo row on each side), and the fetch of the
// next plane is issued before the matrix-core work of the current one so that
// its latency is hidden.  Its 4 waves own RB/4 rows each; a wave's MFMA tiles
// are 16 consecutive x of one row, and the B fragments of each input row it
// reads serve all three dy taps of the rows it owns (voxel pitch kVP).  The
// 32 -> 32
// weights stay in registers (27 taps x 2 A fragments = 216 VGPRs per lane,
// one wave per SIMD), so LDS serves only the activations; work-groups are
// persistent (one per CU) and load the weights once.
template <int G_, int RB_, int ZC_ = 8>
struct Zs {
    static constexpr int G = G_;
    static constexpr int RB = RB_;                        // rows per task
    static constexpr int RW = RB / 4;                     // rows per wave
    static constexpr int NXT = G / 16;                    // x tiles per row
    static constexpr int NT = RW * NXT;                   // output tiles per wave
    static constexpr int ZC = ZC_;                        // planes per task
    static constexpr int HR = RB + 2, HX = G + 2;         // halo plane rows, columns
    static constexpr int PS = HR * HX * kVP;              // halo plane, bf16 elements
    static constexpr int NCH = HR * HX * 4;               // its 16 B chunks
    static constexpr int NPF = (NCH + 255) / 256;         // chunks per thread
    static constexpr int YB = G / RB, ZB = G / ZC;
    static constexpr int NBLK = YB * ZB;                  // tasks per slice
    static_assert(G % 16 == 0 && G % ZC == 0 && G % RB == 0 && RB % 4 == 0, "tile shape");
};
// 32 -> 32 layers (and layer 1 fused in front of layer 2)
template <int G>
using ZsMid = Zs<G, (G == 48 ? 4 : 8)>;
// layer 4: one output voxel per thread, 16 planes per task
template <int G>
using ZsOut = Zs<G, (G == 48 ? 4 : 256 / G), 16>;

struct ZsTile {
    int b, y0, z0;
};

template <typename S>
DGFS_DEV ZsTile zs_decode(int task) {
    ZsTile t;
    t.b = task / S::NBLK;
    const int r = task - t.b * S::NBLK;
    t.y0 = (r % S::YB) * S::RB;
    t.z0 = (r / S::YB) * S::ZC;
    return t;
}

// one task per work-group
template <typename S>
DGFS_DEV ZsTile zs_tile(int nb) {
    return zs_decode<S>(xcd_remap(blockIdx.x, nb * S::NBLK));
}

// persistent work-groups: round r runs task r * gridDim.x + (XCD-remapped
// block id), so the tasks in flight on one XCD are neighbours in the slice
DGFS_DEV int zs_task(int round) { return round * gridDim.x + xcd_remap(blockIdx.x, gridDim.x); }

// per-lane A fragments of [27][32 out][32 in] bf16 weights:
// wf[tap][j] = W[tap][16j + li][8kq .. 8kq + 7]
DGFS_DEV void load_wfrag(bf16x8 (&wf)[kTaps][2], const __bf16* __restrict__ W, int li, int kq) {
#pragma unroll
    for (int tap = 0; tap < kTaps; ++tap)
#pragma unroll
        for (int j = 0; j < 2; ++j)
            wf[tap][j] =
                *reinterpret_cast<const bf16x8*>(&W[(tap * kC + 16 * j + li) * kC + 8 * kq]);
}

// halo plane p (rows y0-1 .. y0+RB, columns -1 .. G) of a 32-channel slice
template <typename S>
DGFS_DEV void plane_fetch(uint4 (&pf)[S::NPF], const __bf16* __restrict__ Xb, int p, int y0) {
    constexpr int G = S::G;
#pragma unroll
    for (int i = 0; i < S::NPF; ++i) {
        const int k = threadIdx.x + 256 * i;
        const int ch = k & 3, vox = k >> 2;
        const int xi = vox % S::HX, yi = vox / S::HX;
        const int y = y0 + yi - 1, x = xi - 1;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (k < S::NCH && p >= 0 && p < G && y >= 0 && y < G && x >= 0 && x < G)
            v = *reinterpret_cast<const uint4*>(&Xb[((size_t(p) * G + y) * G + x) * kC + ch * 8]);
        pf[i] = v;
    }
}

template <typename S>
DGFS_DEV void plane_store(__bf16* slot, const uint4 (&pf)[S::NPF]) {
#pragma unroll
    for (int i = 0; i < S::NPF; ++i) {
        const int k = threadIdx.x + 256 * i;
        if (k < S::NCH) *reinterpret_cast<uint4*>(&slot[(k >> 2) * kVP + (k & 3) * 8]) = pf[i];
    }
}

// 32 -> 32 convolution of output plane z from the ring (slots p mod 3), A
// fragments from registers.  Nine steps (dz, dx); a step reads the B
// fragments of the RW + 2 input rows under the wave's RW output rows once and
// applies each to the (output row, dy) pairs it feeds.  Software-pipelined:
// the fragments of step k + 1 are read from LDS while the MFMAs of step k run.
// epi(step) is called in the region of step's MFMAs: the previous plane's
// epilogue (VALU and stores) issues between them instead of after them.
// (Deriving the dx = +-1 fragments from the dx = 0 ones with DPP row shifts
// reads LDS three times less but measured slower: the shifts cost more than
// the reads they save.)
template <typename S, typename Epi>
DGFS_DEV void mid_plane(floatx4 (&acc)[S::NT][2], const bf16x8 (&wf)[kTaps][2],
                        const __bf16* ring, int z, int wave, int li, int kq, Epi&& epi) {
    constexpr int NR = S::RW + 2;                          // input rows of a wave
    constexpr int NB = NR * S::NXT;                        // B fragments per step
    const __bf16* slot[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) slot[d] = ring + ((z + d + 2) % 3) * S::PS;   // planes z-1, z, z+1
    const int base = (wave * S::RW * S::HX + li) * kVP + 8 * kq;
    auto load = [&](int step, bf16x8(&b)[NB]) {
        const int dz = step / 3, dx = step % 3;
#pragma unroll
        for (int ir = 0; ir < NR; ++ir)
#pragma unroll
            for (int xt = 0; xt < S::NXT; ++xt)
                b[ir * S::NXT + xt] = *reinterpret_cast<const bf16x8*>(
                    slot[dz] + base + (ir * S::HX + xt * 16 + dx) * kVP);
    };
    auto mma = [&](int step, const bf16x8(&b)[NB]) {
        const int dz = step / 3, dx = step % 3;
        // dy outermost: successive updates of one accumulator are RW*NXT*2 MFMAs apart
#pragma unroll
        for (int dy = 0; dy < 3; ++dy)
#pragma unroll
            for (int r = 0; r < S::RW; ++r)
#pragma unroll
                for (int xt = 0; xt < S::NXT; ++xt)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        const int tap = dz * 9 + dy * 3 + dx;
                        floatx4& c = acc[r * S::NXT + xt][j];
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            wf[tap][j], b[(r + dy) * S::NXT + xt], c, 0, 0, 0);
                    }
    };
#pragma unroll
    for (int t = 0; t < S::NT; ++t)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[t][j] = floatx4{0.f, 0.f, 0.f, 0.f};
    bf16x8 b0[NB], b1[NB];
    load(0, b0);
    sfor<0, 9>([&](auto K) {
        constexpr int step = decltype(K)::value;
        // the scheduling fences keep the compiler from sinking the reads of
        // step k + 1 below the MFMAs of step k
        if constexpr (step % 2 == 0) {
            if constexpr (step + 1 < 9) load(step + 1, b1);
            __builtin_amdgcn_sched_barrier(0);
            mma(step, b0);
        } else {
            if constexpr (step + 1 < 9) load(step + 1, b0);
            __builtin_amdgcn_sched_barrier(0);
            mma(step, b1);
        }
        epi(K);
        __builtin_amdgcn_sched_barrier(0);
    });
}

// Epilogue of one output tile group GRP = (tile t, half j) of a finished plane.
// MODE 0: z = acc + bias -> Z (bf16), Y = silu(z) (bf16)
// MODE 1: Y = silu(acc + bias) only (inference)
// MODE 2: input gradient Y = acc * silu'(Zp), no bias
// lane holds output channels 16j + 4kq + i of voxel (tile base + li)
template <typename S, int MODE, int GRP>
DGFS_DEV void store_group(const floatx4 (&acc)[S::NT][2], size_t vrow0, int wave, int li, int kq,
                          const float (&bias)[2][4], const __bf16* __restrict__ Zp,
                          __bf16* __restrict__ Zout, __bf16* __restrict__ Y) {
    constexpr int G = S::G, t = GRP / 2, j = GRP % 2;
    const int yl = wave * S::RW + t / S::NXT;
    const int x = (t % S::NXT) * 16 + li;
    const size_t vo = (vrow0 + size_t(yl) * G + x) * kC;
    const int c0 = 16 * j + 4 * kq;
    __bf16 yo[4], zo[4];
    if (MODE == 2) {
        const uint2 zw = *reinterpret_cast<const uint2*>(&Zp[vo + c0]);
        __bf16 zh[4];
        __builtin_memcpy(zh, &zw, 8);
#pragma unroll
        for (int i = 0; i < 4; ++i) yo[i] = f2bf(acc[t][j][i] * silu_d(bf2f(zh[i])));
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const float zv = acc[t][j][i] + bias[j][i];
            zo[i] = f2bf(zv);
            yo[i] = f2bf(silu(zv));
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

// The plane sequence of one task with the epilogue of plane z - 1 interleaved
// into the MFMAs of plane z (two accumulator sets in turn).  advance(z) runs
// before plane z's MFMAs (prefetch issue) and finish(z) after them (ring
// update); both return nothing and see the same ring state as a plain loop.
template <typename S, int MODE, typename Pre, typename Post>
DGFS_DEV void mid_task(const bf16x8 (&wf)[kTaps][2], const __bf16* ring, int z0, size_t row0,
                       int wave, int li, int kq, const float (&bias)[2][4],
                       const __bf16* __restrict__ Zp, __bf16* __restrict__ Zout,
                       __bf16* __restrict__ Y, Pre&& pre, Post&& post) {
    static_assert(S::ZC % 2 == 0, "planes are processed in pairs");
    constexpr int NG = 2 * S::NT;
    constexpr size_t PL = size_t(S::G) * S::G;
    floatx4 accA[S::NT][2], accB[S::NT][2];
    auto pending = [&](const floatx4(&pa)[S::NT][2], size_t vrow0) {
        return [&pa, vrow0, wave, li, kq, &bias, Zp, Zout, Y](auto K) {
            constexpr int step = decltype(K)::value;
            sfor<0, NG>([&](auto Gi) {
                constexpr int g = decltype(Gi)::value;
                if constexpr (g % 9 == step)
                    store_group<S, MODE, g>(pa, vrow0, wave, li, kq, bias, Zp, Zout, Y);
            });
        };
    };
    auto none = [](auto) {};
    pre(z0);
    mid_plane<S>(accA, wf, ring, z0, wave, li, kq, none);
    post(z0);
    for (int k = 1; k < S::ZC; k += 2) {
        const int z = z0 + k;
        pre(z);
        mid_plane<S>(accB, wf, ring, z, wave, li, kq, pending(accA, row0 + (z - 1) * PL));
        post(z);
        if (k + 1 < S::ZC) {
            pre(z + 1);
            mid_plane<S>(accA, wf, ring, z + 1, wave, li, kq, pending(accB, row0 + z * PL));
            post(z + 1);
        }
    }
    // the last plane (z0 + ZC - 1) is in accB
    sfor<0, NG>([&](auto Gi) {
        store_group<S, MODE, decltype(Gi)::value>(accB, row0 + (z0 + S::ZC - 1) * PL, wave, li,
                                                  kq, bias, Zp, Zout, Y);
    });
}

DGFS_DEV void load_bias(float (&bias)[2][4], const float* __restrict__ b, int kq) {
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 4; ++i) bias[j][i] = b ? b[16 * j + 4 * kq + i] : 0.f;
}

// ---------------------------------------------- 32 -> 32 layer (z-streaming)
// Y = act(conv(X, W) + bias) over a [nb][G][G][G][32] bf16 tensor (MODE as in
// mid_store; for MODE 2, W is the tap-flipped transposed weight set and bias
// is unused).  Persistent: launch at most one work-group per CU.
template <int G, int MODE>
__global__ void __launch_bounds__(256)
k_mid_zs(int nb, const __bf16* __restrict__ X, const __bf16* __restrict__ W,
         const float* __restrict__ bias, const __bf16* __restrict__ Zp, __bf16* __restrict__ Zout,
         __bf16* __restrict__ Y) {
    using S = ZsMid<G>;
    __shared__ __attribute__((aligned(16))) __bf16 ring[3 * S::PS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
    const size_t n = size_t(G) * G * G;
    const int ntask = nb * S::NBLK;
    bf16x8 wf[kTaps][2];
    load_wfrag(wf, W, li, kq);
    float bv[2][4];
    load_bias(bv, MODE == 2 ? nullptr : bias, kq);
    uint4 pf[S::NPF];
    for (int round = 0;; ++round) {
        const int task = zs_task(round);
        if (task >= ntask) break;
        const ZsTile tl = zs_decode<S>(task);
        const __bf16* Xb = X + size_t(tl.b) * n * kC;
        if (round > 0) __syncthreads();                    // last task's ring reads are done
        for (int p = tl.z0 - 1; p <= tl.z0 + 1; ++p) {
            plane_fetch<S>(pf, Xb, p, tl.y0);
            plane_store<S>(ring + ((p + 3) % 3) * S::PS, pf);
        }
        __syncthreads();
        const int zend = tl.z0 + S::ZC;
        mid_task<S, MODE>(
            wf, ring, tl.z0, size_t(tl.b) * n + size_t(tl.y0) * G, wave, li, kq, bv, Zp, Zout, Y,
            [&](int z) {                                   // plane z + 2 feeds output z + 1
                if (z + 1 < zend) plane_fetch<S>(pf, Xb, z + 2, tl.y0);
            },
            [&](int z) {
                if (z + 1 < zend) {
                    __syncthreads();                       // slot (z - 1) mod 3 is free
                    plane_store<S>(ring + ((z + 2) % 3) * S::PS, pf);
                    __syncthreads();
                }
            });
    }
}

// ------------------------------------- layers 1 + 2 fused (z-streaming)
// Layer 1 (1 -> 32) runs on the matrix cores as well: its K dimension is the
// 27 taps (zero padded to 32) of the scaled input s f, rounded to bf16, times
// the bf16-rounded W1; FP32 accumulation, bias and SiLU in FP32.  Its output
// planes (with their halo rows, zero outside the grid) go straight into the
// LDS ring that layer 2 consumes: Y1 never round-trips through HBM.  The
// input f is staged in a 4-plane LDS ring of (RB + 4) x (G + 2) bf16 values.
// MODE 0 also stores Z1, Y1, Z2 of the tile (training); MODE 1 stores Y2 only.
// Persistent like k_mid_zs.
template <int G, int MODE, typename T>
__global__ void __launch_bounds__(256)
k_l12_zs(int nb, const T* __restrict__ u, const int* __restrict__ src,
         const float* __restrict__ scale, const float* __restrict__ W1,
         const float* __restrict__ b1, const __bf16* __restrict__ W2,
         const float* __restrict__ b2, __bf16* __restrict__ Z1, __bf16* __restrict__ Y1,
         __bf16* __restrict__ Z2, __bf16* __restrict__ Y2) {
    using S = ZsMid<G>;
    constexpr int FR = S::HR + 2, FX = G + 2, FP = FR * FX;   // f tile rows, columns, size
    constexpr int NFF = (FP + 255) / 256;
    constexpr int NY1 = S::HR * S::NXT;                       // layer-1 tiles per plane
    __shared__ __attribute__((aligned(16))) __bf16 ring[3 * S::PS];
    __shared__ __attribute__((aligned(16))) __bf16 fr[4 * FP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
    const size_t n = size_t(G) * G * G;
    const int ntask = nb * S::NBLK;
    // the halo columns of the ring are never written again: zero once
    for (int i = threadIdx.x; i < 3 * S::PS / 8; i += 256)
        reinterpret_cast<uint4*>(ring)[i] = make_uint4(0u, 0u, 0u, 0u);
    bf16x8 wf[kTaps][2];
    load_wfrag(wf, W2, li, kq);
    float bias2[2][4], bias1[2][4];
    load_bias(bias2, b2, kq);
    load_bias(bias1, b1, kq);
    // layer-1 operands: A[j] row li = out channel 16j + li, k = tap 8kq + i
    bf16x8 a1[2];
    int toff[8], tdz[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int tap = 8 * kq + i;
        int dz = 0, dy = 0, dx = 0;
        if (tap < kTaps) tap_offset(tap, dz, dy, dx);
        tdz[i] = dz;
        toff[i] = (1 + dy) * FX + (1 + dx);
#pragma unroll
        for (int j = 0; j < 2; ++j)
            a1[j][i] = tap < kTaps ? f2bf(W1[tap * kC + 16 * j + li]) : __bf16(0.0f);
    }

    for (int round = 0;; ++round) {
        const int task = zs_task(round);
        if (task >= ntask) break;
        const ZsTile tl = zs_decode<S>(task);
        const T* f = u + size_t(src[tl.b]) * n;
        const float s = scale[tl.b];

        // f plane p (rows y0-2 .. y0+RB+1, columns -1 .. G), scaled, bf16
        auto f_fetch = [&](float(&fv)[NFF], int p) {
#pragma unroll
            for (int i = 0; i < NFF; ++i) {
                const int k = threadIdx.x + 256 * i;
                const int y = tl.y0 - 2 + k / FX, x = k % FX - 1;
                float v = 0.f;
                if (k < FP && p >= 0 && p < G && y >= 0 && y < G && x >= 0 && x < G)
                    v = s * float(f[(size_t(p) * G + y) * G + x]);
                fv[i] = v;
            }
        };
        auto f_store = [&](const float(&fv)[NFF], int p) {
            __bf16* slot = fr + ((p + 8) & 3) * FP;
#pragma unroll
            for (int i = 0; i < NFF; ++i) {
                const int k = threadIdx.x + 256 * i;
                if (k < FP) slot[k] = f2bf(fv[i]);
            }
        };
        // layer-1 output plane p (halo rows, interior columns) -> ring slot p mod 3
        auto y1_plane = [&](int p) {
            __bf16* slot = ring + ((p + 3) % 3) * S::PS;
            const bool pin = p >= 0 && p < G;
            const bool own = MODE == 0 && p >= tl.z0 && p < tl.z0 + S::ZC;
            for (int t = wave; t < NY1; t += 4) {
                const int r = t / S::NXT, x = (t % S::NXT) * 16 + li;
                const int y = tl.y0 - 1 + r;
                // planes and rows outside the grid are zeros (the slot may hold an older plane)
                const bool in = pin && y >= 0 && y < G;
                bf16x8 bv;
#pragma unroll
                for (int i = 0; i < 8; ++i)
                    bv[i] = in && 8 * kq + i < kTaps
                                ? fr[((p + tdz[i] + 8) & 3) * FP + r * FX + x + toff[i]]
                                : __bf16(0.0f);
                const bool st = in && own && r >= 1 && r <= S::RB;
                const size_t vo = ((size_t(tl.b) * G + p) * G + y) * G * kC + size_t(x) * kC;
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const floatx4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                        a1[j], bv, floatx4{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
                    __bf16 zo[4], yo[4];
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const float zv = c[i] + bias1[j][i];
                        zo[i] = f2bf(zv);
                        yo[i] = in ? f2bf(silu(zv)) : __bf16(0.0f);
                    }
                    uint2 py;
                    __builtin_memcpy(&py, yo, 8);
                    *reinterpret_cast<uint2*>(&slot[(r * S::HX + x + 1) * kVP + 16 * j + 4 * kq]) =
                        py;
                    if (st) {
                        uint2 pz;
                        __builtin_memcpy(&pz, zo, 8);
                        *reinterpret_cast<uint2*>(&Z1[vo + 16 * j + 4 * kq]) = pz;
                        *reinterpret_cast<uint2*>(&Y1[vo + 16 * j + 4 * kq]) = py;
                    }
                }
            }
        };

        float fv[NFF];
        if (round > 0) __syncthreads();                    // last task's LDS reads are done
        for (int p = tl.z0 - 2; p <= tl.z0 + 1; ++p) {
            f_fetch(fv, p);
            f_store(fv, p);
        }
        __syncthreads();
        y1_plane(tl.z0 - 1);
        y1_plane(tl.z0);
        __syncthreads();                                   // f slot of z0 - 2 is free
        f_fetch(fv, tl.z0 + 2);
        f_store(fv, tl.z0 + 2);
        __syncthreads();
        y1_plane(tl.z0 + 1);
        __syncthreads();
        // invariant at the top of iteration z: the ring holds Y1 planes z-1,
        // z, z+1; the f ring holds planes z-1 .. z+2
        const int zend = tl.z0 + S::ZC;
        mid_task<S, MODE>(
            wf, ring, tl.z0, size_t(tl.b) * n + size_t(tl.y0) * G, wave, li, kq, bias2, nullptr, Z2,
            Y2,
            [&](int z) {
                if (z + 1 < zend) f_fetch(fv, z + 3);
            },
            [&](int z) {
                if (z + 1 < zend) {
                    f_store(fv, z + 3);                    // slot of plane z - 1, unused since z
                    __syncthreads();                       // ring slot (z - 1) mod 3 is free
                    y1_plane(z + 2);
                    __syncthreads();
                }
            });
    }
}

// --------------------------------------------- layer 4 (z-streaming, MFMA)
// out[v] = b4 + sum_tap P[tap][v + d(tap)] with P[tap][w] = W4[tap] . Y3[w]:
// P of one input plane (its RB + 2 rows) comes from two MFMAs per 16 voxels
// (A = W4 with the taps as rows, bf16, zero padded to 32; B = Y3 straight from
// HBM), goes to LDS as FP32, and each thread adds the 27 shifted values into
// three rolling output-plane accumulators (output plane q - 1 is complete
// after input plane q).  mode 0: pred[b][v] = out (FP32);
// mode 1: Q[dst[b]][v] += oscale[b] * out (solver dtype).
template <int G, typename T>
__global__ void __launch_bounds__(256)
k_l4_zs(int nb, const __bf16* __restrict__ Y3, const float* __restrict__ W4,
        const float* __restrict__ b4, int mode, float* __restrict__ pred,
        const int* __restrict__ dst, const float* __restrict__ oscale, T* __restrict__ Q) {
    using S = ZsOut<G>;
    constexpr int TS = S::HR * G + 4;          // tap stride: kq groups land on distinct banks
    constexpr int NY = S::HR * S::NXT;
    __shared__ float Ps[kTaps * TS];
    constexpr int NYW = (NY + 3) / 4;          // P tiles per wave and plane
    const ZsTile tl = zs_tile<S>(nb);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, li = lane & 15, kq = lane >> 4;
    const size_t n = size_t(G) * G * G;
    const __bf16* Yb = Y3 + size_t(tl.b) * n * kC;
    bf16x8 a4[2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int tap = 16 * j + li;
            a4[j][i] = tap < kTaps ? f2bf(W4[tap * kC + 8 * kq + i]) : __bf16(0.0f);
        }
    const int oy = threadIdx.x / G, ox = threadIdx.x % G;       // output voxel of this thread
    const bool owner = threadIdx.x < S::RB * G;
    float o_prev = 0.f, o_cur = 0.f, o_next = 0.f;              // planes q - 1, q, q + 1
    const float bias = b4[0];
    // B fragments of input plane q (zero outside the grid), one plane ahead
    auto fetch = [&](int q, bf16x8(&b)[NYW]) __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < NYW; ++k) {
            const int t = wave + 4 * k;
            const int r = t / S::NXT, x = (t % S::NXT) * 16 + li;
            const int y = tl.y0 - 1 + r;
            bf16x8 v;
#pragma unroll
            for (int e = 0; e < 8; ++e) v[e] = __bf16(0.0f);
            if (t < NY && q >= 0 && q < G && y >= 0 && y < G)
                v = *reinterpret_cast<const bf16x8*>(&Yb[((size_t(q) * G + y) * G + x) * kC + 8 * kq]);
            b[k] = v;
        }
    };
    bf16x8 bc[NYW], bn[NYW];
    fetch(tl.z0 - 1, bc);
    for (int q = tl.z0 - 1; q <= tl.z0 + S::ZC; ++q) {
        if (q < tl.z0 + S::ZC) fetch(q + 1, bn);
        if (q >= 0 && q < G) {
#pragma unroll
            for (int k = 0; k < NYW; ++k) {
                const int t = wave + 4 * k;
                if (t >= NY) break;
                const int r = t / S::NXT, x = (t % S::NXT) * 16 + li;
                floatx4 c[2] = {floatx4{0.f, 0.f, 0.f, 0.f}, floatx4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    c[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a4[j], bc[k], c[j], 0, 0, 0);
#pragma unroll
                for (int j = 0; j < 2; ++j)
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int tap = 16 * j + 4 * kq + i;
                        if (tap < kTaps) Ps[tap * TS + r * G + x] = c[j][i];
                    }
            }
            __syncthreads();
            if (owner) {
#pragma unroll
                for (int dz = -1; dz <= 1; ++dz) {
                    float sum = 0.f;
#pragma unroll
                    for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int tap = (dz + 1) * 9 + (dy + 1) * 3 + (dx + 1);
                            const int xx = ox + dx;
                            if (xx >= 0 && xx < G) sum += Ps[tap * TS + (oy + 1 + dy) * G + xx];
                        }
                    // input plane q feeds output plane q - dz
                    if (dz == 1) o_prev += sum;
                    else if (dz == 0) o_cur += sum;
                    else o_next += sum;
                }
            }
            __syncthreads();
        }
        const int zo = q - 1;                                    // complete now
        if (owner && zo >= tl.z0 && zo < tl.z0 + S::ZC) {
            const size_t v = (size_t(zo) * G + tl.y0 + oy) * G + ox;
            const float out = o_prev + bias;
            if (mode == 0) {
                pred[size_t(tl.b) * n + v] = out;
            } else {
                T* qp = Q + size_t(dst[tl.b]) * n + v;
                *qp = T(float(*qp) + oscale[tl.b] * out);
            }
        }
#pragma unroll
        for (int k = 0; k < NYW; ++k) bc[k] = bn[k];
        o_prev = o_cur;
        o_cur = o_next;
        o_next = 0.f;
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

// ------------------------------------------------ weight gradients
// ------------------------------------------- 32 -> 32 weight gradient, MFMA
// dW[tap][co][ci] = sum_v Gr[v][co] X[v + d(tap)][ci] and db[co] = sum_v Gr[v][co]
// over [nb][G^3][32] bf16 tensors.  A task is RB rows x G columns of one
// plane z of one slice; persistent work-groups stage, per task, the task's
// Gr tile and X planes z-1..z+1 (rows y0-1..y0+RB, zero halo) channel-major
// in LDS: GT[co][voxel] and XT[ci][plane][row][column].
// v_mfma_f32_16x16x32_bf16 with rows = co (16 jc + ..), columns = ci
// (16 ic + ..) and K = 32 voxels of the tile: A lane (li, kq) =
// GT[16 jc + li][k0 + 8kq .. +7] (one 16 B read, shared by every tap), B lane
// (li, kq) = X at the same 8 voxels shifted by d(tap), channel 16 ic + li
// (eight 2 B reads).  Wave w accumulates taps w, w+4, ... in FP32 over all of
// its tasks; each work-group writes one partial dW and db (reduced afterwards
// in a fixed order).
template <int G>
struct WgTile {
    static constexpr int RB = G == 16 ? 8 : 4;             // rows per task
    static constexpr int TV = RB * G;                     // tile voxels (multiple of 32)
    static constexpr int NK = TV / 32;                    // K chunks
    static constexpr int HR = RB + 2, HX = G + 2;
    static constexpr int GTP = TV + 8;                    // GT row pitch (bf16)
    static constexpr int XTP = 3 * HR * HX + 8;           // XT channel pitch (bf16)
    static constexpr int NTASK = G * (G / RB);            // tasks per slice
    static_assert(TV % 32 == 0 && G % RB == 0, "tile");
};

template <int G>
__global__ void __launch_bounds__(256)
k_wgrad_mma(int nb, const __bf16* __restrict__ Gr, const __bf16* __restrict__ X,
            float* __restrict__ part, float* __restrict__ bpart) {
    using W = WgTile<G>;
    constexpr int MT = (kTaps + 3) / 4;                   // taps per wave
    __shared__ __attribute__((aligned(16))) __bf16 GT[kC * W::GTP];
    __shared__ __attribute__((aligned(16))) __bf16 XT[kC * W::XTP];
    __shared__ float sb[256 * 8];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, li = lane & 15, kq = lane >> 4;
    const size_t n = size_t(G) * G * G;
    const int ntask = nb * W::NTASK;
    const int cg = tid & 3;                               // channel group of this thread's loads
    float bsum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    floatx4 acc[MT][2][2];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int jc = 0; jc < 2; ++jc)
#pragma unroll
            for (int ic = 0; ic < 2; ++ic) acc[m][jc][ic] = floatx4{0.f, 0.f, 0.f, 0.f};
    for (int task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int b = task / W::NTASK, rr = task - b * W::NTASK;
        const int z = rr / (G / W::RB), y0 = (rr % (G / W::RB)) * W::RB;
        const __bf16* Gb = Gr + size_t(b) * n * kC;
        const __bf16* Xb = X + size_t(b) * n * kC;
        __syncthreads();                                  // previous task's reads are done
        // Gr tile, transposed; bias partial sums (this thread's 8 channels)
        for (int k = tid; k < W::TV * 4; k += 256) {
            const int tv = k >> 2;
            const int y = y0 + tv / G, x = tv % G;
            const uint4 w = *reinterpret_cast<const uint4*>(
                &Gb[((size_t(z) * G + y) * G + x) * kC + 8 * cg]);
            __bf16 h[8];
            __builtin_memcpy(h, &w, 16);
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                GT[(8 * cg + e) * W::GTP + tv] = h[e];
                bsum[e] += bf2f(h[e]);
            }
        }
        // X planes z-1..z+1 with zero halo, transposed
        for (int k = tid; k < 3 * W::HR * W::HX * 4; k += 256) {
            const int vox = k >> 2;
            const int pz = vox / (W::HR * W::HX), rem = vox - pz * (W::HR * W::HX);
            const int r = rem / W::HX, c = rem - r * W::HX;
            const int zz = z + pz - 1, y = y0 + r - 1, x = c - 1;
            uint4 w = make_uint4(0u, 0u, 0u, 0u);
            if (zz >= 0 && zz < G && y >= 0 && y < G && x >= 0 && x < G)
                w = *reinterpret_cast<const uint4*>(
                    &Xb[((size_t(zz) * G + y) * G + x) * kC + 8 * cg]);
            __bf16 h[8];
            __builtin_memcpy(h, &w, 16);
#pragma unroll
            for (int e = 0; e < 8; ++e) XT[(8 * cg + e) * W::XTP + vox] = h[e];
        }
        __syncthreads();
        bf16x8 a[W::NK][2];
#pragma unroll
        for (int kc = 0; kc < W::NK; ++kc)
#pragma unroll
            for (int jc = 0; jc < 2; ++jc)
                a[kc][jc] = *reinterpret_cast<const bf16x8*>(
                    &GT[(16 * jc + li) * W::GTP + kc * 32 + 8 * kq]);
#pragma unroll
        for (int m = 0; m < MT; ++m) {
            const int tap = wave + 4 * m;
            if (tap >= kTaps) break;
            int dz, dy, dx;
            tap_offset(tap, dz, dy, dx);
#pragma unroll
            for (int kc = 0; kc < W::NK; ++kc) {
                // the lane's 8 voxels lie in one row: tile voxel kc*32 + 8kq + e
                const int tv = kc * 32 + 8 * kq;
                const int r = tv / G, x = tv % G;
                const int off = ((dz + 1) * W::HR + r + dy + 1) * W::HX + x + dx + 1;
#pragma unroll
                for (int ic = 0; ic < 2; ++ic) {
                    const __bf16* src = &XT[(16 * ic + li) * W::XTP + off];
                    bf16x8 bv;
#pragma unroll
                    for (int e = 0; e < 8; ++e) bv[e] = src[e];
#pragma unroll
                    for (int jc = 0; jc < 2; ++jc)
                        acc[m][jc][ic] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            a[kc][jc], bv, acc[m][jc][ic], 0, 0, 0);
                }
            }
        }
    }
    // partial dW of this work-group: lane holds co = 16 jc + 4 kq + i, ci = 16 ic + li
    float* p = part + size_t(blockIdx.x) * kTaps * kC * kC;
#pragma unroll
    for (int m = 0; m < MT; ++m) {
        const int tap = wave + 4 * m;
        if (tap >= kTaps) break;
#pragma unroll
        for (int jc = 0; jc < 2; ++jc)
#pragma unroll
            for (int ic = 0; ic < 2; ++ic)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    p[(size_t(tap) * kC + 16 * jc + 4 * kq + i) * kC + 16 * ic + li] =
                        acc[m][jc][ic][i];
    }
    // bias partial: threads with the same channel group, in a fixed order
#pragma unroll
    for (int e = 0; e < 8; ++e) sb[tid * 8 + e] = bsum[e];
    __syncthreads();
    if (tid < kC) {
        const int g = tid >> 3, e = tid & 7;
        float s = 0.f;
        for (int t = g; t < 256; t += 4) s += sb[t * 8 + e];
        bpart[size_t(blockIdx.x) * kC + tid] = s;
    }
}

// Thin layers (one side has one channel), tiled like k_wgrad_mma:
//   GC1 = true  (layer 4): out[tap][c] = sum_v s[v] T[v + d][c], s = dOut
//                          (FP32, the tile), T = Y3 (bf16, planes z-1..z+1)
//   GC1 = false (layer 1): out[tap][c] = sum_v T[v][c] s[v + d], T = dZ1
//                          (bf16, the tile), s = scaled input (FP32, halo)
// and the bias partials (sum of dOut, or of dZ1 per channel).  Thread
// (c = tid & 31, tq = tid >> 5) owns channel c and taps tq, tq+8, tq+16,
// tq+24; partials part[wg][tap][c], bpart[wg][32] (GC1: bpart[wg]).
template <int G, bool GC1>
__global__ void __launch_bounds__(256)
k_wgrad_thin(int nb, const float* __restrict__ s1, const __bf16* __restrict__ t32,
             float* __restrict__ part, float* __restrict__ bpart) {
    using W = WgTile<G>;
    constexpr int HV = 3 * W::HR * W::HX;                   // halo voxels
    constexpr int TP = GC1 ? HV + 8 : W::TV + 8;            // channel pitch of T (bf16)
    constexpr int SN = GC1 ? W::TV : HV;                    // scalar field size
    __shared__ __attribute__((aligned(16))) __bf16 Ts[kC * TP];
    __shared__ float Ss[SN];
    const int tid = threadIdx.x, c = tid & 31, tq = tid >> 5;
    const size_t n = size_t(G) * G * G;
    const int ntask = nb * W::NTASK;
    int toff[4];
    bool tin[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int tap = tq + 8 * k;
        tin[k] = tap < kTaps;
        int dz = 0, dy = 0, dx = 0;
        if (tin[k]) tap_offset(tap, dz, dy, dx);
        toff[k] = ((dz + 1) * W::HR + dy + 1) * W::HX + dx + 1;   // halo offset of (0, 0, 0) + d
    }
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    float bacc = 0.f;
    for (int task = blockIdx.x; task < ntask; task += gridDim.x) {
        const int b = task / W::NTASK, rr = task - b * W::NTASK;
        const int z = rr / (G / W::RB), y0 = (rr % (G / W::RB)) * W::RB;
        __syncthreads();
        if (GC1) {
            for (int k = tid; k < W::TV; k += 256) {
                const int y = y0 + k / G, x = k % G;
                Ss[k] = s1[size_t(b) * n + (size_t(z) * G + y) * G + x];
            }
            for (int k = tid; k < HV * 4; k += 256) {
                const int vox = k >> 2, cg = k & 3;
                const int pz = vox / (W::HR * W::HX), rem = vox - pz * (W::HR * W::HX);
                const int r = rem / W::HX, cc = rem - r * W::HX;
                const int zz = z + pz - 1, y = y0 + r - 1, x = cc - 1;
                uint4 w = make_uint4(0u, 0u, 0u, 0u);
                if (zz >= 0 && zz < G && y >= 0 && y < G && x >= 0 && x < G)
                    w = *reinterpret_cast<const uint4*>(
                        &t32[(size_t(b) * n + (size_t(zz) * G + y) * G + x) * kC + 8 * cg]);
                __bf16 h[8];
                __builtin_memcpy(h, &w, 16);
#pragma unroll
                for (int e = 0; e < 8; ++e) Ts[(8 * cg + e) * TP + vox] = h[e];
            }
        } else {
            for (int k = tid; k < HV; k += 256) {
                const int pz = k / (W::HR * W::HX), rem = k - pz * (W::HR * W::HX);
                const int r = rem / W::HX, cc = rem - r * W::HX;
                const int zz = z + pz - 1, y = y0 + r - 1, x = cc - 1;
                float v = 0.f;
                if (zz >= 0 && zz < G && y >= 0 && y < G && x >= 0 && x < G)
                    v = s1[size_t(b) * n + (size_t(zz) * G + y) * G + x];
                Ss[k] = v;
            }
            for (int k = tid; k < W::TV * 4; k += 256) {
                const int tv = k >> 2, cg = k & 3;
                const int y = y0 + tv / G, x = tv % G;
                const uint4 w = *reinterpret_cast<const uint4*>(
                    &t32[(size_t(b) * n + (size_t(z) * G + y) * G + x) * kC + 8 * cg]);
                __bf16 h[8];
                __builtin_memcpy(h, &w, 16);
#pragma unroll
                for (int e = 0; e < 8; ++e) Ts[(8 * cg + e) * TP + tv] = h[e];
            }
        }
        __syncthreads();
        for (int tv = 0; tv < W::TV; ++tv) {
            const int r = tv / G, x = tv % G;
            const int h0 = r * W::HX + x;                 // halo index of (r, x) minus the d offset
            const float a = GC1 ? Ss[tv] : bf2f(Ts[c * TP + tv]);
            if (tq == 0) bacc += a;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (tin[k]) acc[k] += a * (GC1 ? bf2f(Ts[c * TP + h0 + toff[k]]) : Ss[h0 + toff[k]]);
        }
    }
    float* p = part + size_t(blockIdx.x) * kTaps * kC;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (tin[k]) p[(tq + 8 * k) * kC + c] = acc[k];
    if (GC1) {
        if (tid == 0) bpart[blockIdx.x] = bacc;           // the same sum in every channel lane
    } else if (tq == 0) {
        bpart[size_t(blockIdx.x) * kC + c] = bacc;
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

// scaled input of layer 1, rounded to bf16 as the forward uses it, stored as
// FP32 (kept for the layer-1 weight gradient)
template <typename T>
__global__ void __launch_bounds__(256)
k_gather_scaled(int n, long nvox, const T* __restrict__ u, const int* __restrict__ src,
                const float* __restrict__ scale, float* __restrict__ out) {
    const long gid = long(blockIdx.x) * 256 + threadIdx.x;
    if (gid >= nvox) return;
    const int b = int(gid / n), v = int(gid - long(b) * n);
    out[gid] = bf2f(f2bf(scale[b] * float(u[size_t(src[b]) * n + v])));
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

constexpr int kWgradMaxGridDev = 512;   // == kWgradMaxGrid (work buffer sizing)

// compute units of the current device (persistent grids: one work-group per CU)
int cu_count() {
    static int cached[64] = {0};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cached[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            n <= 0)
            n = 256;
        cached[dev] = n;
    }
    return cached[dev];
}

template <int G>
dim3 persistent_grid(int nb) {
    const int ntask = nb * ZsMid<G>::NBLK, ncu = cu_count();
    return dim3(ntask < ncu ? ntask : ncu);
}

template <int G>
int launch_mid(int mode, int nb, const __bf16* x, const __bf16* w, const float* bias,
               const __bf16* zp, __bf16* zo, __bf16* y, hipStream_t st) {
    const dim3 grid = persistent_grid<G>(nb);
    if (mode == 0)
        hipLaunchKernelGGL((k_mid_zs<G, 0>), grid, dim3(256), 0, st, nb, x, w, bias, zp, zo, y);
    else if (mode == 1)
        hipLaunchKernelGGL((k_mid_zs<G, 1>), grid, dim3(256), 0, st, nb, x, w, bias, zp, zo, y);
    else
        hipLaunchKernelGGL((k_mid_zs<G, 2>), grid, dim3(256), 0, st, nb, x, w, bias, zp, zo, y);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

// 32 -> 32 layer; mode 0: store Z and Y = silu(Z), 1: Y only, 2: input gradient
int dispatch_mid(int G, int mode, int nb, const void* X, const void* W, const float* bias,
                 const void* Zp, void* Zout, void* Y, hipStream_t st) {
    const __bf16* x = (const __bf16*)X;
    const __bf16* w = (const __bf16*)W;
    const __bf16* zp = (const __bf16*)Zp;
    __bf16* zo = (__bf16*)Zout;
    __bf16* y = (__bf16*)Y;
    if (G == 16) return launch_mid<16>(mode, nb, x, w, bias, zp, zo, y, st);
    if (G == 32) return launch_mid<32>(mode, nb, x, w, bias, zp, zo, y, st);
    if (G == 48) return launch_mid<48>(mode, nb, x, w, bias, zp, zo, y, st);
    DGFS_REQUIRE(false, "closure: grid size must be 16, 32 or 48");
}

// layers 1 + 2 (mode 0 stores Z1 Y1 Z2 Y2, 1 stores Y2), then layer 3 and 4
template <int G, typename T>
int launch_forward(const dgfs_closure_t* net, int nb, const T* u, const int* src,
                   const float* scale, __bf16* A, size_t a, int mode, float* pred, const int* dst,
                   const float* oscale, T* Q, hipStream_t st) {
    const dim3 grid = persistent_grid<G>(nb);
    if (mode == 0)
        hipLaunchKernelGGL((k_l12_zs<G, 0, T>), grid, dim3(256), 0, st, nb, u, src, scale, net->W1,
                           net->b1, (const __bf16*)net->W2f, net->b2, A, A + a, A + 2 * a,
                           A + 3 * a);
    else
        hipLaunchKernelGGL((k_l12_zs<G, 1, T>), grid, dim3(256), 0, st, nb, u, src, scale, net->W1,
                           net->b1, (const __bf16*)net->W2f, net->b2, A, A + a, A + 2 * a,
                           A + 3 * a);
    DGFS_HIP_CHECK(hipGetLastError());
    const int rc = launch_mid<G>(mode, nb, A + 3 * a, (const __bf16*)net->W3f, net->b3, nullptr,
                                 A + 4 * a, A + 5 * a, st);
    if (rc) return rc;
    hipLaunchKernelGGL((k_l4_zs<G, T>), dim3(nb * ZsOut<G>::NBLK), dim3(256), 0, st, nb, A + 5 * a,
                       net->W4, net->b4,
                       mode, pred, dst, oscale, Q);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

// MFMA weight gradient of a 32 -> 32 layer into per-work-group partials at
// work (dW partials, then db partials); ng receives the number of partials
template <int G>
int launch_wgrad(int nb, const __bf16* g, const __bf16* x, float* work, int& ng, hipStream_t st) {
    const int ntask = nb * WgTile<G>::NTASK;
    int cap = 2 * cu_count();
    if (cap > kWgradMaxGridDev) cap = kWgradMaxGridDev;
    ng = ntask < cap ? ntask : cap;
    hipLaunchKernelGGL((k_wgrad_mma<G>), dim3(ng), dim3(256), 0, st, nb, g, x, work,
                       work + size_t(ng) * kTaps * kC * kC);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

template <int G, bool GC1>
int launch_wthin(int nb, const float* s1, const __bf16* t32, float* work, int& ng,
                 hipStream_t st) {
    const int ntask = nb * WgTile<G>::NTASK;
    int cap = 2 * cu_count();
    if (cap > kWgradMaxGridDev) cap = kWgradMaxGridDev;
    ng = ntask < cap ? ntask : cap;
    hipLaunchKernelGGL((k_wgrad_thin<G, GC1>), dim3(ng), dim3(256), 0, st, nb, s1, t32, work,
                       work + size_t(ng) * kTaps * kC);
    DGFS_HIP_CHECK(hipGetLastError());
    return 0;
}

// thin-layer weight gradient partials (dW then db) at work; ng partials
int wgrad_thin(int G, bool gc1, int nb, const float* s1, const void* t32, float* work, int& ng,
               hipStream_t st) {
    const __bf16* t = (const __bf16*)t32;
#define DGFS_WT(GG)                                                             \
    if (G == GG)                                                                \
        return gc1 ? launch_wthin<GG, true>(nb, s1, t, work, ng, st)            \
                   : launch_wthin<GG, false>(nb, s1, t, work, ng, st);
    DGFS_WT(16)
    DGFS_WT(32)
    DGFS_WT(48)
#undef DGFS_WT
    DGFS_REQUIRE(false, "closure: grid size must be 16, 32 or 48");
}

int wgrad_mid(int G, int nb, const void* g, const void* x, float* work, int& ng, hipStream_t st) {
    if (G == 16) return launch_wgrad<16>(nb, (const __bf16*)g, (const __bf16*)x, work, ng, st);
    if (G == 32) return launch_wgrad<32>(nb, (const __bf16*)g, (const __bf16*)x, work, ng, st);
    if (G == 48) return launch_wgrad<48>(nb, (const __bf16*)g, (const __bf16*)x, work, ng, st);
    DGFS_REQUIRE(false, "closure: grid size must be 16, 32 or 48");
}

template <typename T>
int dispatch_forward(int G, const dgfs_closure_t* net, int nb, const T* u, const int* src,
                     const float* scale, __bf16* A, size_t a, int mode, float* pred,
                     const int* dst, const float* oscale, T* Q, hipStream_t st) {
    if (G == 16)
        return launch_forward<16, T>(net, nb, u, src, scale, A, a, mode, pred, dst, oscale, Q, st);
    if (G == 32)
        return launch_forward<32, T>(net, nb, u, src, scale, A, a, mode, pred, dst, oscale, Q, st);
    if (G == 48)
        return launch_forward<48, T>(net, nb, u, src, scale, A, a, mode, pred, dst, oscale, Q, st);
    DGFS_REQUIRE(false, "closure: grid size must be 16, 32 or 48");
}

}  // namespace

// most work-groups of the MFMA weight-gradient kernel (one partial each)
constexpr int kWgradMaxGrid = kWgradMaxGridDev;

extern "C" size_t dgfs_closure_work_bytes(int G, int nb, int chunk) {
    const long total = long(nb) * G * G * G;
    long nchunk = (total + chunk - 1) / chunk;
    if (nchunk < kWgradMaxGrid) nchunk = kWgradMaxGrid;
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
    DGFS_REQUIRE(long(nb) * 72 < 2147483647L, "closure_forward: batch too large");
    const size_t a = size_t(nvox) * kC;         // one activation tensor (bf16 elements)
    __bf16* A = (__bf16*)act;                   // [Z1 Y1 Z2 Y2 Z3 Y3], 6 tensors
    if (dtype == DGFS_F64)
        return dispatch_forward<double>(G, net, nb, (const double*)u, src, scale, A, a, mode, pred,
                                        dst, oscale, (double*)Q, st);
    return dispatch_forward<float>(G, net, nb, (const float*)u, src, scale, A, a, mode, pred, dst,
                                   oscale, (float*)Q, st);
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

    const size_t a = size_t(nvox) * kC;
    const __bf16* A = (const __bf16*)act;       // Z1 Y1 Z2 Y2 Z3 Y3
    __bf16* GA = (__bf16*)gact;                 // two gradient tensors (ping-pong)
    const dim3 gv((nvox + 255) / 256);
    // layer 4: dW4, db4 and dZ3
    int ng = 0;
    int rc0 = wgrad_thin(G, true, nb, dOut, A + 5 * a, work, ng, st);
    if (rc0) return rc0;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC + 255) / 256), dim3(256), 0, st, ng,
                       kTaps * kC, work, gr->W4);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, ng, 1,
                       work + size_t(ng) * kTaps * kC, gr->b4);
    DGFS_HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_out_bwd, gv, dim3(256), 0, st, G, nvox, dOut, net->W4, A + 4 * a, GA);
    DGFS_HIP_CHECK(hipGetLastError());
    // layer 3: dW3 = dZ3 x Y2, dZ2 = silu'(Z2) convT(dZ3, W3)
    rc0 = wgrad_mid(G, nb, GA, A + 3 * a, work, ng, st);
    if (rc0) return rc0;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC * kC + 255) / 256), dim3(256), 0, st, ng,
                       kTaps * kC * kC, work, gr->W3);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, ng, kC,
                       work + size_t(ng) * kTaps * kC * kC, gr->b3);
    DGFS_HIP_CHECK(hipGetLastError());
    int rc = dispatch_mid(G, 2, nb, GA, net->W3t, nullptr, A + 2 * a, nullptr, GA + a, st);
    if (rc) return rc;
    // layer 2: dW2 = dZ2 x Y1, dZ1 = silu'(Z1) convT(dZ2, W2)
    rc0 = wgrad_mid(G, nb, GA + a, A + a, work, ng, st);
    if (rc0) return rc0;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC * kC + 255) / 256), dim3(256), 0, st, ng,
                       kTaps * kC * kC, work, gr->W2);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, ng, kC,
                       work + size_t(ng) * kTaps * kC * kC, gr->b2);
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
    rc0 = wgrad_thin(G, false, nb, fin, GA, work, ng, st);
    if (rc0) return rc0;
    hipLaunchKernelGGL(k_reduce_chunks, dim3((kTaps * kC + 255) / 256), dim3(256), 0, st, ng,
                       kTaps * kC, work, gr->W1);
    hipLaunchKernelGGL(k_reduce_chunks, dim3(1), dim3(256), 0, st, ng, kC,
                       work + size_t(ng) * kTaps * kC, gr->b1);
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
