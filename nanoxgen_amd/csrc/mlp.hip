// Attribute surrogate: a 10 -> 64 -> 64 -> 64 -> 64 -> 4 SiLU MLP on gfx950
// MFMA (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32 accumulation), with
// its training step (forward, backward, deterministic gradient reduction,
// Adam on FP32 master weights), the fused root-feature + inference kernel,
// and the sampled exact check that sends failing tiles of roots back to the
// exact expression runtime.  No reference counterpart: the reference only
// has the exact runtime (src/xgen_expression.cpp:984-1148), which stays the
// fallback and the arbiter.
//
// Numerics.  Weights are BF16 (rounded from the FP32 masters; training sees
// the rounded weights, so it learns around their rounding).  Inputs and
// hidden activations are carried as a pair of BF16 terms, hi = bf16(h) and
// lo = bf16(h - hi), and every product is two MFMAs, W.hi + W.lo, into one
// FP32 accumulator: the activations keep about 16 significant bits instead
// of 8 (Henry, Tang & Heinecke 2019, "Leveraging the bfloat16 artificial
// intelligence datatype for higher-precision computations").  Biases are
// FP32.
//
// Inference (surrogate_kernel, mlp_forward_kernel).  The network runs
// transposed, Z^T = W^T H^T: the weights are the MFMA A operand, 16 output
// neurons by 32 inputs per fragment, pre-swizzled into the per-lane order of
// the instruction and staged once per workgroup in LDS; 16 roots are the
// columns.  A lane of a 16 x 16 result holds 4 consecutive neurons of one
// root; the next layer's B operand wants 8 consecutive inputs of one root in
// the same lane, so the inputs of every layer after the first are numbered
// in the order the previous layer leaves them (the weight fragments carry
// that permutation) and activations never leave registers.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "../../include/nanoxgen_amd.h"
#include "internal.h"

namespace nxg {
namespace {

constexpr int kIn = NXG_MLP_INPUTS, kHid = NXG_MLP_HIDDEN, kOut = NXG_MLP_OUTPUTS;
constexpr int kInPad = 16, kOutPad = 16, kLayers = NXG_MLP_HIDDEN_LAYERS + 1;
constexpr int kTile = 64;

// Padded, input-major weight shapes and the flat parameter layout
// [W0 b0 W1 b1 ... W4 b4].
__host__ __device__ constexpr int rows_of(int l) { return l == 0 ? kInPad : kHid; }
__host__ __device__ constexpr int cols_of(int l) { return l == kLayers - 1 ? kOutPad : kHid; }
__host__ __device__ constexpr int w_off(int l) {
    int o = 0;
    for (int i = 0; i < l; ++i) o += rows_of(i) * cols_of(i) + cols_of(i);
    return o;
}
__host__ __device__ constexpr int b_off(int l) { return w_off(l) + rows_of(l) * cols_of(l); }
constexpr int kParams = w_off(kLayers);
static_assert(kParams == 14608, "parameter count");
static_assert(NXG_MLP_MAX_BATCH % 64u == 0u, "batch tiles");

// ---- MFMA fragments --------------------------------------------------------
// v_mfma_f32_16x16x32_bf16, lane L, g = L / 16, r = L % 16:
//   A (16 x 32): A[r][8g + j], j = 0..7;  B (32 x 16): B[8g + j][r];
//   D (16 x 16): D[4g + j][r], j = 0..3.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4f mfma(const bf16x8 &a, const bf16x8 &b, const v4f &c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Fragment (layer l, output tile tm, input chunk kc): 16 output neurons by 32
// inputs, 64 lanes x 8 BF16.  Layer 0 has one chunk (16 padded inputs, the
// upper 16 slots zero); the output layer one tile (16 padded outputs).
__host__ __device__ constexpr int frag_index(int l, int tm, int kc) {
    return l == 0 ? tm : (l < kLayers - 1 ? 4 + (l - 1) * 8 + tm * 2 + kc : 28 + kc);
}
constexpr int kFrags = 30;
constexpr int kFragElems = kFrags * 64 * 8;  // BF16 elements
static_assert(frag_index(kLayers - 1, 0, 1) == kFrags - 1, "fragment count");

// Where W_l[k][n] (input-major master) sits among the fragments.  Input slot
// of a hidden or output layer: the previous layer leaves neuron
// 16 T + 4 g + jj (T = output tile, jj = 0..3) in lane group g, so chunk kc
// slot j of lane group g holds neuron 16 (2 kc + j / 4) + 4 g + j % 4.
__host__ __device__ inline int frag_position(int l, int k, int n) {
    const int tm = n / 16, r = n % 16;
    int g, j, kc;
    if (l == 0) {
        kc = 0;
        g = k / 8;
        j = k % 8;
    } else {
        const int T = k / 16;
        g = (k % 16) / 4;
        kc = T / 2;
        j = (T % 2) * 4 + k % 4;
    }
    return (frag_index(l, tm, kc) * 64 + g * 16 + r) * 8 + j;
}

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16u) & 1u);
    return static_cast<uint16_t>(u >> 16u);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float(uint32_t{b} << 16u); }

// SiLU with the hardware exp2 and reciprocal (used alike by training and
// inference); its derivative for the backward pass.
__device__ __forceinline__ float sigmoid(float z) {
    return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.4426950408889634f));
}
__device__ __forceinline__ float silu(float z) { return z * sigmoid(z); }

// The two BF16 terms of 8 FP32 values h as MFMA operands, from nh = -h:
// hi = bf16(h); the dot product of the hi pair with a unit vector plus nh is
// hi - h exactly (one rounding of an exactly representable sum), so
// lo = bf16(h - hi) is its negation -- two instructions per pair instead of
// widening hi back to FP32 and subtracting.  The unit vectors are held in
// scalar registers: written as constants, the compiler encodes (1, 0) as the
// inline constant 1.0, which the hardware reads as the pair (0, 1).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split8_neg(const float (&nh)[8], bf16x8 &hi, bf16x8 &lo) {
    uint32_t u0 = 0x00003f80u, u1 = 0x3f800000u;  // (1, 0) and (0, 1)
    asm volatile("" : "+s"(u0));
    asm volatile("" : "+s"(u1));
    const bf16x2 e0 = __builtin_bit_cast(bf16x2, u0);
    const bf16x2 e1 = __builtin_bit_cast(bf16x2, u1);
#pragma unroll
    for (int j = 0; j < 8; j += 2) {
        bf16x2 p;
        p[0] = static_cast<__bf16>(-nh[j]);
        p[1] = static_cast<__bf16>(-nh[j + 1]);
        const float d0 = __builtin_amdgcn_fdot2_f32_bf16(p, e0, nh[j], false);
        const float d1 = __builtin_amdgcn_fdot2_f32_bf16(p, e1, nh[j + 1], false);
        hi[j] = p[0];
        hi[j + 1] = p[1];
        lo[j] = static_cast<__bf16>(-d0);
        lo[j + 1] = static_cast<__bf16>(-d1);
    }
}

// Orders this wavefront's LDS traffic: every lane's earlier LDS writes are
// visible to every lane's later reads (the feature staging is per
// wavefront, and wavefronts of a workgroup run different numbers of tiles,
// so no workgroup barrier is possible there).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct MlpState {
    float *params;          // FP32 master parameters
    float *adam_m, *adam_v;
    uint16_t *params_bf16;  // BF16 copy of the flat layout (training)
    uint16_t *frags;        // BF16 weight fragments (inference)
    float *norm;            // in_mean[10], in_inv_std[10], out_scale[4], out_offset[4]
    float *partials;        // [max_blocks][kParams + 1] per-block gradients and loss
    float *loss;            // mean squared error of the last step (normalised targets)
};

// ---- inference --------------------------------------------------------------

// Workgroup shape of the inference kernels: kInferWaves wavefronts share one
// staged copy of the network; each runs the network kInferBlocks column
// blocks of 16 roots at a time.
#ifndef NXG_INFER_WAVES
#define NXG_INFER_WAVES 4
#endif
#ifndef NXG_INFER_BLOCKS
#define NXG_INFER_BLOCKS 2
#endif
constexpr int kInferWaves = NXG_INFER_WAVES, kInferBlocks = NXG_INFER_BLOCKS;
constexpr int kFeatRow = 12;  // floats per staged feature row (10 features, zero padding)
static_assert(kTile % (16 * kInferBlocks) == 0, "column blocks per tile");

// What a workgroup stages once: the weight fragments, the biases and the
// normalisation; and per wavefront the normalised features of its tile.
struct InferShared {
    bf16x8 frag[kFrags * 64];
    float bias[NXG_MLP_HIDDEN_LAYERS * kHid + kOutPad];
    float norm[28];
    float feat[kInferWaves][kTile][kFeatRow];
};

__device__ void stage_network(const MlpState &s, InferShared &sh) {
    const uint4 *src = reinterpret_cast<const uint4 *>(s.frags);
    uint4 *dst = reinterpret_cast<uint4 *>(sh.frag);
    for (int i = threadIdx.x; i < kFrags * 64; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < NXG_MLP_HIDDEN_LAYERS * kHid + kOutPad; i += blockDim.x) {
        const int l = i / kHid, n = i % kHid;
        sh.bias[i] = s.params[b_off(l) + n];
    }
    for (int i = threadIdx.x; i < 28; i += blockDim.x) sh.norm[i] = s.norm[i];
    for (int i = threadIdx.x; i < kInferWaves * kTile; i += blockDim.x)
        for (int k = kIn; k < kFeatRow; ++k) sh.feat[i / kTile][i % kTile][k] = 0.0f;
    __syncthreads();
}

// A weight fragment from LDS.
__device__ __forceinline__ bf16x8 frag_at(const InferShared &sh, int index, int lane) {
    return sh.frag[index * 64 + lane];
}

// A compiler-only barrier at each layer: it keeps a layer's fragment reads
// in that layer.  Hoisted out of the tile loop, the 30 fragments would hold
// 120 registers and leave one wavefront per SIMD.
__device__ __forceinline__ void layer_fence() { asm volatile("" ::: "memory"); }

// The network over CB column blocks of 16 roots whose normalised features
// sit in f[root][0..9] (LDS rows of kFeatRow floats, zero from 10 on);
// returns in lanes 0..15 of y[cb] the 4 normalised outputs of root
// cb*16 + lane.
template <int CB>
__device__ __forceinline__ void network(const InferShared &sh, const float (*f)[kFeatRow], v4f (&y)[CB]) {
    const int lane = threadIdx.x & 63, g = lane >> 4, r = lane & 15;
    bf16x8 hi[CB][2], lo[CB][2];  // [column block][input chunk]
    // layer 0: inputs 8g..8g+7 of root r -- features 0..7 for lane group 0,
    // 8..9 and zeros for group 1, zeros for groups 2 and 3
    layer_fence();
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) {
        const float4 *row = reinterpret_cast<const float4 *>(f[cb * 16 + r]);
        const float4 p = row[g & 1 ? 2 : 0], q = row[1];
        const float kp = g < 2 ? 1.0f : 0.0f, kq = g == 0 ? 1.0f : 0.0f;
        const float nx[8] = {-p.x * kp, -p.y * kp, -p.z * kp, -p.w * kp,
                             -q.x * kq, -q.y * kq, -q.z * kq, -q.w * kq};
        split8_neg(nx, hi[cb][0], lo[cb][0]);
    }
    v4f acc[CB][4];
#pragma unroll
    for (int tm = 0; tm < 4; ++tm) {
        const bf16x8 a = frag_at(sh, frag_index(0, tm, 0), lane);
        const float4 b4 = *reinterpret_cast<const float4 *>(&sh.bias[16 * tm + 4 * g]);
        const v4f c = {b4.x, b4.y, b4.z, b4.w};
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) acc[cb][tm] = mfma(a, lo[cb][0], mfma(a, hi[cb][0], c));
    }
#pragma unroll
    for (int l = 1; l < kLayers; ++l) {
        layer_fence();
        // SiLU of the previous layer, in the next layer's input order
#pragma unroll
        for (int cb = 0; cb < CB; ++cb) {
#pragma unroll
            for (int kc = 0; kc < 2; ++kc) {
                float nh[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const float z = acc[cb][2 * kc + j / 4][j % 4];
                    nh[j] = -z * sigmoid(z);
                }
                split8_neg(nh, hi[cb][kc], lo[cb][kc]);
            }
        }
        const int tiles = l < kLayers - 1 ? 4 : 1;
#pragma unroll
        for (int tm = 0; tm < tiles; ++tm) {
            const float4 b4 = *reinterpret_cast<const float4 *>(&sh.bias[l * kHid + 16 * tm + 4 * g]);
            const v4f c = {b4.x, b4.y, b4.z, b4.w};
            const bf16x8 a0 = frag_at(sh, frag_index(l, tm, 0), lane);
            const bf16x8 a1 = frag_at(sh, frag_index(l, tm, 1), lane);
#pragma unroll
            for (int cb = 0; cb < CB; ++cb)
                acc[cb][tm] = mfma(a1, lo[cb][1], mfma(a1, hi[cb][1], mfma(a0, lo[cb][0], mfma(a0, hi[cb][0], c))));
        }
    }
#pragma unroll
    for (int cb = 0; cb < CB; ++cb) y[cb] = acc[cb][0];
}

// Outputs of a 64-root tile: lanes 0..15 of each column block hold the 4
// normalised outputs of one root.
template <class Store>
__device__ __forceinline__ void network_tile(const InferShared &sh, const float (*f)[kFeatRow], const Store &store) {
    const int lane = threadIdx.x & 63;
    constexpr int kRoots = 16 * kInferBlocks;
#pragma unroll 1
    for (int pass = 0; pass < kTile / kRoots; ++pass) {
        v4f y[kInferBlocks];
        network<kInferBlocks>(sh, f + pass * kRoots, y);
        if (lane < 16) {
#pragma unroll
            for (int cb = 0; cb < kInferBlocks; ++cb) {
                float o[4];
#pragma unroll
                for (int j = 0; j < 4; ++j) o[j] = y[cb][j] * sh.norm[2 * kIn + j] + sh.norm[2 * kIn + kOut + j];
                store(pass * kRoots + cb * 16 + lane, make_float4(o[0], o[1], o[2], o[3]));
            }
        }
    }
    wave_lds_sync();  // the staged features may be overwritten from here on
}

// Inference over a feature array: outputs are attributes (denormalised).
// Each wavefront takes 64-row tiles in turn.
__global__ void __launch_bounds__(64 * kInferWaves) mlp_forward_kernel(MlpState s, const float *features,
                                                                       float *outputs, uint64_t count) {
    extern __shared__ uint8_t smem[];
    InferShared &sh = *reinterpret_cast<InferShared *>(smem);
    stage_network(s, sh);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t tiles = (count + kTile - 1) / kTile;
    for (uint64_t t = uint64_t{blockIdx.x} * kInferWaves + wave; t < tiles; t += uint64_t{gridDim.x} * kInferWaves) {
        const uint64_t first = t * kTile;
        const uint64_t row = first + lane < count ? first + lane : count - 1;
        for (int k = 0; k < kIn; ++k)
            sh.feat[wave][lane][k] = (features[row * kIn + k] - sh.norm[k]) * sh.norm[kIn + k];
        wave_lds_sync();
        network_tile(sh, sh.feat[wave], [&](int m, float4 o) {
            if (first + m < count) {
                float *dst = outputs + (first + m) * kOut;
                dst[0] = o.x;
                dst[1] = o.y;
                dst[2] = o.z;
                dst[3] = o.w;
            }
        });
    }
}

// ---- training ---------------------------------------------------------------

// One 16 x 16 output tile of C = A(M x K) * B(K x N), K a multiple of 32,
// with operands gathered element by element: a(m, k) and b(k, n) return
// BF16 bits.  acc[j] = C[tm*16 + 4*(lane/16) + j][tn*16 + lane%16].
template <class FA, class FB>
__device__ __forceinline__ v4f mfma_gather(int tm, int tn, int K, const FA &a, const FB &b, int lane, v4f acc) {
    const int r = lane & 15, q = lane >> 4;
    for (int kk = 0; kk < K; kk += 32) {
        bf16x8 av, bv;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            av[j] = __builtin_bit_cast(__bf16, a(tm * 16 + r, kk + 8 * q + j));
            bv[j] = __builtin_bit_cast(__bf16, b(kk + 8 * q + j, tn * 16 + r));
        }
        acc = mfma(av, bv, acc);
    }
    return acc;
}

// A 64-sample training tile in LDS.  The forward keeps every layer's
// pre-activation z (FP32); the hidden activations are recomputed from z when
// the backward pass needs them (the same function of the same value, so the
// same BF16 terms).
struct TrainBuffers {
    uint16_t x_hi[kTile * kInPad], x_lo[kTile * kInPad];  // normalised inputs
    uint16_t h_hi[2][kTile * kHid], h_lo[2][kTile * kHid];  // activations (forward ping-pong; backward input)
    float z[NXG_MLP_HIDDEN_LAYERS][kTile * kHid];
    float y[kTile * kOutPad];
    uint16_t dz[kTile * kHid];  // BF16 gradient of the current layer's pre-activation
    float dh[kTile * kHid];     // FP32 gradient of the previous layer's output
};

__device__ __forceinline__ void split_store(float h, uint16_t &hi, uint16_t &lo) {
    hi = f2bf(h);
    lo = f2bf(h - bf2f(hi));
}

// One training step over samples [first, first + batch) (wrapping modulo
// count): forward, loss, backward; per-block gradient partials (no atomics:
// the reduction order is fixed).
__global__ void __launch_bounds__(256) mlp_train_kernel(MlpState s, const float *features, const float *targets,
                                                        uint64_t count, uint64_t first, uint32_t batch) {
    extern __shared__ uint8_t smem[];
    TrainBuffers &t = *reinterpret_cast<TrainBuffers *>(smem);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint64_t tile_first = first + uint64_t{blockIdx.x} * kTile;
    const uint32_t valid = batch - blockIdx.x * kTile < static_cast<uint32_t>(kTile) ? batch - blockIdx.x * kTile : kTile;
    for (int i = threadIdx.x; i < kTile * kInPad; i += blockDim.x) {
        const int m = i / kInPad, k = i % kInPad;
        const uint64_t sidx = (tile_first + m) % count;
        float v = 0.0f;
        if (k < kIn && static_cast<uint32_t>(m) < valid) v = (features[sidx * kIn + k] - s.norm[k]) * s.norm[kIn + k];
        split_store(v, t.x_hi[i], t.x_lo[i]);
    }
    __syncthreads();
    const uint16_t *W = s.params_bf16;
    // forward
    for (int l = 0; l < kLayers; ++l) {
        const int K = rows_of(l), N = cols_of(l), ldin = l == 0 ? kInPad : kHid;
        const uint16_t *Wl = W + w_off(l);
        const float *bl = s.params + b_off(l);
        const uint16_t *in_hi = l == 0 ? t.x_hi : t.h_hi[(l - 1) & 1];
        const uint16_t *in_lo = l == 0 ? t.x_lo : t.h_lo[(l - 1) & 1];
        const int tiles = (kTile / 16) * (N / 16);
        for (int tt = wave; tt < tiles; tt += 4) {
            const int tm = tt / (N / 16), tn = tt % (N / 16);
            auto wb = [&](int k, int n) -> uint16_t { return k < K ? Wl[k * N + n] : uint16_t{0}; };
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma_gather(tm, tn, 32 * ((K + 31) / 32),
                              [&](int m, int k) -> uint16_t { return k < K ? in_hi[m * ldin + k] : uint16_t{0}; },
                              wb, lane, acc);
            acc = mfma_gather(tm, tn, 32 * ((K + 31) / 32),
                              [&](int m, int k) -> uint16_t { return k < K ? in_lo[m * ldin + k] : uint16_t{0}; },
                              wb, lane, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = tm * 16 + 4 * (lane >> 4) + j, n = tn * 16 + (lane & 15);
                const float z = acc[j] + bl[n];
                if (l < kLayers - 1) {
                    t.z[l][m * kHid + n] = z;
                    split_store(silu(z), t.h_hi[l & 1][m * kHid + n], t.h_lo[l & 1][m * kHid + n]);
                } else {
                    t.y[m * kOutPad + n] = z;
                }
            }
        }
        __syncthreads();
    }
    float *part = s.partials + uint64_t{blockIdx.x} * (kParams + 1);
    // output gradient dy = 2 (y - t) / (batch * 4) on valid rows; loss partial
    float loss = 0.0f;
    for (int i = threadIdx.x; i < kTile * kOutPad; i += blockDim.x) {
        const int m = i / kOutPad, n = i % kOutPad;
        float d = 0.0f;
        if (n < kOut && static_cast<uint32_t>(m) < valid) {
            const uint64_t sidx = (tile_first + m) % count;
            const float tn = (targets[sidx * kOut + n] - s.norm[2 * kIn + kOut + n]) / s.norm[2 * kIn + n];
            const float diff = t.y[i] - tn;
            loss += diff * diff;
            d = 2.0f * diff / (static_cast<float>(batch) * kOut);
        }
        t.dh[m * kHid + n] = d;  // dh holds the FP32 output gradient
    }
    // block loss reduction (fixed order: per-wave shuffle, then wave 0)
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    __shared__ float wave_loss[4];
    if (lane == 0) wave_loss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) part[kParams] = wave_loss[0] + wave_loss[1] + wave_loss[2] + wave_loss[3];

    for (int l = kLayers - 1; l >= 0; --l) {
        const int K = rows_of(l), N = cols_of(l), ldin = l == 0 ? kInPad : kHid;
        // dz = dh * silu'(z) (identity on the output layer), FP32 -> BF16;
        // the layer's input activations (hi, lo) recomputed from z into
        // ping-pong slot 0
        for (int i = threadIdx.x; i < kTile * kHid; i += blockDim.x) {
            const int m = i / kHid, n = i % kHid;
            if (n < N) {
                float d = t.dh[m * kHid + n];
                if (l < kLayers - 1) {
                    const float z = t.z[l][m * kHid + n];
                    const float sg = sigmoid(z);
                    d *= sg * (1.0f + z * (1.0f - sg));
                }
                t.dz[m * kHid + n] = f2bf(d);
                t.dh[m * kHid + n] = d;  // keep the FP32 dz for the bias gradient
            }
            if (l > 0) split_store(silu(t.z[l - 1][i]), t.h_hi[0][i], t.h_lo[0][i]);
        }
        __syncthreads();
        // bias gradient: column sums of FP32 dz, fixed order
        for (int n = threadIdx.x; n < N; n += blockDim.x) {
            float acc = 0.0f;
            for (int m = 0; m < kTile; ++m) acc += t.dh[m * kHid + n];
            part[b_off(l) + n] = acc;
        }
        // weight gradient dW[k][n] = sum_m (hi + lo)[m][k] dz[m][n]
        const uint16_t *in_hi = l == 0 ? t.x_hi : t.h_hi[0];
        const uint16_t *in_lo = l == 0 ? t.x_lo : t.h_lo[0];
        const int wt = (K / 16) * (N / 16);
        for (int tt = wave; tt < wt; tt += 4) {
            const int tk = tt / (N / 16), tn = tt % (N / 16);
            auto dzb = [&](int m, int n) -> uint16_t { return t.dz[m * kHid + n]; };
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma_gather(tk, tn, kTile, [&](int k, int m) -> uint16_t { return in_hi[m * ldin + k]; }, dzb,
                              lane, acc);
            acc = mfma_gather(tk, tn, kTile, [&](int k, int m) -> uint16_t { return in_lo[m * ldin + k]; }, dzb,
                              lane, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = tk * 16 + 4 * (lane >> 4) + j, n = tn * 16 + (lane & 15);
                part[w_off(l) + k * N + n] = acc[j];
            }
        }
        __syncthreads();
        if (l == 0) break;
        // dh_prev[m][k] = sum_n dz[m][n] W[k][n]  (64 x K, contraction over N)
        const uint16_t *Wl = W + w_off(l);
        const int ht = (kTile / 16) * (K / 16);
        for (int tt = wave; tt < ht; tt += 4) {
            const int tm = tt / (K / 16), tk = tt % (K / 16);
            v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mfma_gather(
                tm, tk, 32 * ((N + 31) / 32),
                [&](int m, int n) -> uint16_t { return n < N ? t.dz[m * kHid + n] : uint16_t{0}; },
                [&](int n, int k) -> uint16_t { return n < N ? Wl[k * N + n] : uint16_t{0}; }, lane, acc);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = tm * 16 + 4 * (lane >> 4) + j, k = tk * 16 + (lane & 15);
                t.dh[m * kHid + k] = acc[j];
            }
        }
        __syncthreads();
    }
}

// The BF16 copies of parameter i: the flat copy and, for a weight, its slot
// among the inference fragments.
__device__ __forceinline__ void store_bf16(const MlpState &s, int i, float p) {
    const uint16_t b = f2bf(p);
    s.params_bf16[i] = b;
    for (int l = 0; l < kLayers; ++l) {
        if (i >= w_off(l) && i < b_off(l)) {
            const int e = i - w_off(l), k = e / cols_of(l), n = e % cols_of(l);
            s.frags[frag_position(l, k, n)] = b;
        }
    }
}

// Reduce the per-block partials in block order, then Adam on the FP32
// master parameters and refresh of the BF16 copies.  A block owns 64
// parameters; its four wavefronts sum interleaved quarters of the partials
// (independent loads in flight), then the quarters are added in order.
__global__ void __launch_bounds__(256) mlp_adam_kernel(MlpState s, uint32_t blocks, uint32_t batch, uint32_t step,
                                                       float lr, float beta1, float beta2, float eps) {
    __shared__ float quarter[4][64];
    __shared__ float loss_quarter[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int i = blockIdx.x * 64 + lane;
    const uint64_t stride = kParams + 1;
    float g = 0.0f, lq = 0.0f;
#pragma unroll 8
    for (uint32_t b = wave; b < blocks; b += 4) {
        if (i < kParams) g += s.partials[uint64_t{b} * stride + i];
        if (blockIdx.x == 0 && lane == 0) lq += s.partials[uint64_t{b} * stride + kParams];
    }
    quarter[wave][lane] = g;
    if (lane == 0) loss_quarter[wave] = lq;
    __syncthreads();
    if (wave != 0) return;
    if (blockIdx.x == 0 && lane == 0)
        s.loss[0] = (((loss_quarter[0] + loss_quarter[1]) + loss_quarter[2]) + loss_quarter[3]) /
                    (static_cast<float>(batch) * kOut);
    if (i >= kParams) return;
    g = ((quarter[0][lane] + quarter[1][lane]) + quarter[2][lane]) + quarter[3][lane];
    const float m = beta1 * s.adam_m[i] + (1.0f - beta1) * g;
    const float v = beta2 * s.adam_v[i] + (1.0f - beta2) * g * g;
    s.adam_m[i] = m;
    s.adam_v[i] = v;
    const float mh = m / (1.0f - powf(beta1, static_cast<float>(step)));
    const float vh = v / (1.0f - powf(beta2, static_cast<float>(step)));
    const float p = s.params[i] - lr * mh / (sqrtf(vh) + eps);
    s.params[i] = p;
    store_bf16(s, i, p);
}

// Both BF16 copies from the masters; also zeroes the fragment slots no
// parameter maps to (layer 0's inputs 16..31).
__global__ void mlp_sync_kernel(MlpState s) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kParams) store_bf16(s, i, s.params[i]);
    if (i < 4 * 32 * 8) {  // tile, lane group 2..3 x 16 lanes, slot
        const int tm = i / 256, lane = 32 + (i / 8) % 32, j = i % 8;
        s.frags[(frag_index(0, tm, 0) * 64 + lane) * 8 + j] = 0u;
    }
}

}  // namespace

// ---- root features (shared with the fused surrogate kernel) ---------------

struct FeatureArgs {
    const nxg_vec3 *positions;
    const nxg_vec3 *normals;
    const nxg_vec2 *texcoords;
    const nxg_uint3 *triangles;
    const nxg_alias_entry *alias;
    nxg_texel_map map;
    uint32_t strand_count, seed, triangle_count, has_normals;
};

__device__ __forceinline__ uint32_t f_hash32(uint32_t x) {  // generate.h:72-79
    x ^= x >> 16u;
    x *= 0x7feb352du;
    x ^= x >> 15u;
    x *= 0x846ca68bu;
    x ^= x >> 16u;
    return x;
}
__device__ __forceinline__ float f_random01(uint32_t seed, uint32_t strand, uint32_t dim) {
    const uint32_t h = f_hash32(seed ^ f_hash32(strand + 0x9e3779b9u * (dim + 1u)));
    return static_cast<float>(h >> 8u) * (1.0f / 16777216.0f);
}

__device__ __forceinline__ float texel_at(const nxg_texel_map &m, float face, float u, float v) {
    const float R = static_cast<float>(m.resolution);
    auto clampf = [](float x, float lo, float hi) { return x < lo ? lo : (hi < x ? hi : x); };
    const float ff = clampf(floorf(face), 0.0f, static_cast<float>(m.face_count - 1u));
    const float fu = clampf(floorf(u * R), 0.0f, R - 1.0f);
    const float fv = clampf(floorf(v * R), 0.0f, R - 1.0f);
    return m.texels[(static_cast<uint64_t>(ff) * m.resolution + static_cast<uint32_t>(fv)) * m.resolution +
                    static_cast<uint32_t>(fu)];
}

// The 10 surrogate features of a strand's root: position xyz, normal xyz,
// texcoord uv, texel(face, b1, b2) and texel(face, b2, b1) of map 0 -- the
// same sample_root (generate.h:101-135) and point samples the exact
// programs read.
__device__ void root_features(const FeatureArgs &a, uint32_t strand, float *f) {
    const float choose = f_random01(a.seed, strand, 0u) * static_cast<float>(a.triangle_count);
    uint32_t bucket = static_cast<uint32_t>(choose);
    if (bucket >= a.triangle_count) bucket = a.triangle_count - 1u;
    const nxg_alias_entry e = a.alias[bucket];
    uint32_t tri = f_random01(a.seed, strand, 1u) < e.probability ? bucket : e.alias;
    if (tri >= a.triangle_count) tri = bucket;
    const float r1 = f_random01(a.seed, strand, 2u), r2 = f_random01(a.seed, strand, 3u);
    const float s = sqrtf(r1), b0 = 1.0f - s, b1 = s * (1.0f - r2), b2 = s * r2;
    const nxg_uint3 t = a.triangles[tri];
    const nxg_vec3 p0 = a.positions[t.x], p1 = a.positions[t.y], p2 = a.positions[t.z];
    f[0] = p0.x * b0 + p1.x * b1 + p2.x * b2;
    f[1] = p0.y * b0 + p1.y * b1 + p2.y * b2;
    f[2] = p0.z * b0 + p1.z * b1 + p2.z * b2;
    float nx, ny, nz;
    if (a.has_normals) {
        const nxg_vec3 n0 = a.normals[t.x], n1 = a.normals[t.y], n2 = a.normals[t.z];
        nx = n0.x * b0 + n1.x * b1 + n2.x * b2;
        ny = n0.y * b0 + n1.y * b1 + n2.y * b2;
        nz = n0.z * b0 + n1.z * b1 + n2.z * b2;
    } else {
        const float ex = p1.x - p0.x, ey = p1.y - p0.y, ez = p1.z - p0.z;
        const float fx = p2.x - p0.x, fy = p2.y - p0.y, fz = p2.z - p0.z;
        nx = ey * fz - ez * fy;
        ny = ez * fx - ex * fz;
        nz = ex * fy - ey * fx;
    }
    const float l2 = nx * nx + ny * ny + nz * nz;
    if (l2 > 1.0e-20f) {
        const float l = sqrtf(l2);
        nx = nx / l;
        ny = ny / l;
        nz = nz / l;
    } else {
        nx = 0.0f;
        ny = 1.0f;
        nz = 0.0f;
    }
    f[3] = nx;
    f[4] = ny;
    f[5] = nz;
    f[6] = 0.0f;
    f[7] = 0.0f;
    if (a.texcoords) {
        const nxg_vec2 u0 = a.texcoords[t.x], u1 = a.texcoords[t.y], u2 = a.texcoords[t.z];
        f[6] = u0.x * b0 + u1.x * b1 + u2.x * b2;
        f[7] = u0.y * b0 + u1.y * b1 + u2.y * b2;
    }
    const float face = static_cast<float>(tri);
    f[8] = texel_at(a.map, face, b1, b2);
    f[9] = texel_at(a.map, face, b2, b1);
}

namespace {

__global__ void __launch_bounds__(256) features_kernel(FeatureArgs a, float *out) {
    const uint64_t s = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x;
    if (s >= a.strand_count) return;
    float f[kIn];
    root_features(a, static_cast<uint32_t>(s), f);
    for (int k = 0; k < kIn; ++k) out[s * kIn + k] = f[k];
}

// Fused: root features of a wavefront's 64 roots (one per lane), their
// normalisation, the network, attributes.  Each wavefront takes 64-root
// tiles in turn; the workgroup stages the network once.  With `order`
// (strand indices grouped by root triangle) position p of the processing
// order is strand order[p], and its attributes go to that strand's slot.
__global__ void __launch_bounds__(64 * kInferWaves) surrogate_kernel(FeatureArgs a, MlpState s, const uint32_t *order,
                                                                     float4 *attributes) {
    extern __shared__ uint8_t smem[];
    InferShared &sh = *reinterpret_cast<InferShared *>(smem);
    __shared__ uint32_t strand_of[kInferWaves][kTile];
    stage_network(s, sh);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t tiles = (a.strand_count + kTile - 1) / kTile;
    for (uint32_t t = blockIdx.x * kInferWaves + wave; t < tiles; t += gridDim.x * kInferWaves) {
        const uint32_t first = t * kTile;
        const uint32_t p = first + lane < a.strand_count ? first + lane : a.strand_count - 1u;
        const uint32_t strand = order ? order[p] : p;
        float f[kIn];
        root_features(a, strand, f);
#pragma unroll
        for (int k = 0; k < kIn; ++k) sh.feat[wave][lane][k] = (f[k] - sh.norm[k]) * sh.norm[kIn + k];
        strand_of[wave][lane] = strand;
        wave_lds_sync();
        network_tile(sh, sh.feat[wave], [&](int m, float4 o) {
            if (first + m < a.strand_count) attributes[strand_of[wave][m]] = o;
        });
    }
}

// Workgroups of a persistent inference launch: enough to fill every CU
// (as many as fit at once), at most one per kInferWaves tiles.
uint32_t infer_grid(const void *kernel, uint64_t tiles) {
    static int ctas = [] {
        int device = 0, cus = 256;
        if (hipGetDevice(&device) == hipSuccess)
            (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
        return cus;
    }();
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, 64 * kInferWaves, sizeof(InferShared)) !=
            hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    const uint64_t want = (tiles + kInferWaves - 1u) / kInferWaves;
    const uint64_t most = uint64_t{static_cast<uint32_t>(ctas)} * static_cast<uint32_t>(per_cu);
    return static_cast<uint32_t>(want < most ? want : most);
}

}  // namespace

// The training tile needs more than the default 64 KB of dynamic LDS
// (gfx950 has 160 KB per workgroup).
bool allow_train_lds() {
    static const bool ok = hipFuncSetAttribute(reinterpret_cast<const void *>(&mlp_train_kernel),
                                               hipFuncAttributeMaxDynamicSharedMemorySize,
                                               static_cast<int>(sizeof(TrainBuffers))) == hipSuccess;
    return ok;
}

MlpState mlp_state_of(void *base) {
    nxg_mlp_layout L;
    nxg_mlp_get_layout(&L);
    uint8_t *b = static_cast<uint8_t *>(base);
    MlpState s;
    s.params = reinterpret_cast<float *>(b + L.params_offset);
    s.adam_m = reinterpret_cast<float *>(b + L.adam_m_offset);
    s.adam_v = reinterpret_cast<float *>(b + L.adam_v_offset);
    s.params_bf16 = reinterpret_cast<uint16_t *>(b + L.params_bf16_offset);
    s.frags = reinterpret_cast<uint16_t *>(b + L.fragments_offset);
    s.norm = reinterpret_cast<float *>(b + L.norm_offset);
    s.partials = reinterpret_cast<float *>(b + L.partials_offset);
    s.loss = reinterpret_cast<float *>(b + L.loss_offset);
    return s;
}

FeatureArgs feature_args(const nxg_device_asset &asset, const nxg_generation_params &p, const nxg_texel_map &map) {
    const nxg_asset_header &h = asset.header;
    auto at = [&](uint64_t off) { return off ? static_cast<const uint8_t *>(asset.data) + off : nullptr; };
    FeatureArgs a{};
    a.positions = reinterpret_cast<const nxg_vec3 *>(at(h.positions_offset));
    a.normals = reinterpret_cast<const nxg_vec3 *>(at(h.normals_offset));
    a.texcoords = (h.flags & NXG_ASSET_HAS_TEXCOORDS) ? reinterpret_cast<const nxg_vec2 *>(at(h.texcoords_offset)) : nullptr;
    a.triangles = reinterpret_cast<const nxg_uint3 *>(at(h.triangles_offset));
    a.alias = reinterpret_cast<const nxg_alias_entry *>(at(h.alias_table_offset));
    a.map = map;
    a.strand_count = p.strand_count;
    a.seed = p.seed;
    a.triangle_count = h.triangle_count;
    a.has_normals = (h.flags & NXG_ASSET_HAS_NORMALS) != 0u;
    return a;
}

}  // namespace nxg

extern "C" {

void nxg_mlp_get_layout(nxg_mlp_layout *out) {
    if (!out) return;
    auto align = [](uint64_t v) { return (v + 255u) & ~uint64_t{255u}; };
    const uint64_t P = nxg::kParams;
    nxg_mlp_layout L{};
    L.param_count = P;
    L.params_offset = 0;
    L.adam_m_offset = align(L.params_offset + P * 4u);
    L.adam_v_offset = align(L.adam_m_offset + P * 4u);
    L.params_bf16_offset = align(L.adam_v_offset + P * 4u);
    L.fragments_offset = align(L.params_bf16_offset + P * 2u);
    L.norm_offset = align(L.fragments_offset + uint64_t{nxg::kFragElems} * 2u);
    L.loss_offset = align(L.norm_offset + 28u * 4u);
    L.partials_offset = align(L.loss_offset + 4u);
    L.max_batch = NXG_MLP_MAX_BATCH;
    L.total_bytes = align(L.partials_offset + uint64_t{NXG_MLP_MAX_BATCH / 64u} * (P + 1u) * 4u);
    *out = L;
}

uint32_t nxg_mlp_sync(void *state, void *stream) {
    if (!state) return NXG_ERROR_NULL_ARGUMENT;
    hipLaunchKernelGGL(nxg::mlp_sync_kernel, dim3((nxg::kParams + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), nxg::mlp_state_of(state));
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t nxg_mlp_train_step(void *state, const float *features, const float *targets, uint64_t sample_count,
                            uint64_t batch_first, uint32_t batch_size, float learning_rate, uint32_t step,
                            void *stream) {
    if (!state || !features || !targets) return NXG_ERROR_NULL_ARGUMENT;
    if (sample_count == 0u || batch_size == 0u || batch_size > NXG_MLP_MAX_BATCH || step == 0u ||
        !(learning_rate > 0.0f))
        return NXG_ERROR_INVALID_MLP_REQUEST;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const nxg::MlpState st = nxg::mlp_state_of(state);
    const uint32_t blocks = (batch_size + 63u) / 64u;
    if (!nxg::allow_train_lds()) return NXG_ERROR_HIP;
    hipLaunchKernelGGL(nxg::mlp_train_kernel, dim3(blocks), dim3(256), sizeof(nxg::TrainBuffers), s, st, features,
                       targets, sample_count, batch_first % sample_count, batch_size);
    if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    hipLaunchKernelGGL(nxg::mlp_adam_kernel, dim3((nxg::kParams + 63) / 64), dim3(256), 0, s, st, blocks,
                       batch_size, step, learning_rate, 0.9f, 0.999f, 1.0e-8f);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t nxg_mlp_gradients(void *state, const float *features, const float *targets, uint64_t sample_count,
                           uint64_t batch_first, uint32_t batch_size, void *stream) {
    if (!state || !features || !targets) return NXG_ERROR_NULL_ARGUMENT;
    if (sample_count == 0u || batch_size == 0u || batch_size > NXG_MLP_MAX_BATCH) return NXG_ERROR_INVALID_MLP_REQUEST;
    const uint32_t blocks = (batch_size + 63u) / 64u;
    if (!nxg::allow_train_lds()) return NXG_ERROR_HIP;
    hipLaunchKernelGGL(nxg::mlp_train_kernel, dim3(blocks), dim3(256), sizeof(nxg::TrainBuffers),
                       static_cast<hipStream_t>(stream), nxg::mlp_state_of(state), features, targets, sample_count,
                       batch_first % sample_count, batch_size);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t nxg_mlp_forward(const void *state, const float *features, float *outputs, uint64_t count, void *stream) {
    if (!state || !features || !outputs) return NXG_ERROR_NULL_ARGUMENT;
    if (count == 0u) return NXG_OK;
    const uint64_t tiles = (count + 63u) / 64u;
    hipLaunchKernelGGL(nxg::mlp_forward_kernel,
                       dim3(nxg::infer_grid(reinterpret_cast<const void *>(&nxg::mlp_forward_kernel), tiles)),
                       dim3(64 * nxg::kInferWaves), sizeof(nxg::InferShared), static_cast<hipStream_t>(stream),
                       nxg::mlp_state_of(const_cast<void *>(state)), features, outputs, count);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t nxg_root_features(const nxg_device_asset *asset, const nxg_generation_params *params,
                           const nxg_texel_map *map, float *features, uint64_t capacity, void *stream) {
    if (uint32_t s = nxg::validate_common(asset, nullptr, params, nullptr); s != NXG_OK) return s;
    if (!map || !map->texels || map->face_count == 0u || map->resolution == 0u || !features)
        return NXG_ERROR_NULL_ARGUMENT;
    if (capacity < uint64_t{params->strand_count} * NXG_MLP_INPUTS) return NXG_ERROR_BUFFER_TOO_SMALL;
    hipLaunchKernelGGL(nxg::features_kernel, dim3((params->strand_count + 255u) / 256u), dim3(256), 0,
                       static_cast<hipStream_t>(stream), nxg::feature_args(*asset, *params, *map), features);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t nxg_surrogate_attributes(const nxg_device_asset *asset, const nxg_generation_params *params,
                                  const void *state, const nxg_texel_map *map, float *attributes,
                                  uint64_t attribute_capacity, const void *prepared_workspace,
                                  uint64_t workspace_bytes, void *stream) {
    if (uint32_t s = nxg::validate_common(asset, nullptr, params, nullptr); s != NXG_OK) return s;
    if (!state || !map || !map->texels || map->face_count == 0u || map->resolution == 0u)
        return NXG_ERROR_NULL_ARGUMENT;
    if (!attributes || attribute_capacity < uint64_t{params->strand_count} * NXG_ATTR_COUNT)
        return NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL;
    const uint32_t *order = nullptr;
    if (prepared_workspace) {
        const nxg::WorkspaceLayout L = nxg::workspace_layout(asset->header.guide_count, params->cvs_per_strand,
                                                             params->strand_count, asset->header.triangle_count);
        if (workspace_bytes < L.total) return NXG_ERROR_WORKSPACE_TOO_SMALL;
        order = reinterpret_cast<const uint32_t *>(static_cast<const uint8_t *>(prepared_workspace) + L.order_offset);
    }
    const uint64_t tiles = (uint64_t{params->strand_count} + 63u) / 64u;
    hipLaunchKernelGGL(nxg::surrogate_kernel,
                       dim3(nxg::infer_grid(reinterpret_cast<const void *>(&nxg::surrogate_kernel), tiles)),
                       dim3(64 * nxg::kInferWaves), sizeof(nxg::InferShared), static_cast<hipStream_t>(stream),
                       nxg::feature_args(*asset, *params, *map), nxg::mlp_state_of(const_cast<void *>(state)),
                       order, reinterpret_cast<float4 *>(attributes));
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

}  // extern "C"
