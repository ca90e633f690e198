// Attribute surrogate: a 10 -> 64 -> 64 -> 64 -> 64 -> 4 SiLU MLP on gfx950
// MFMA (v_mfma_f32_16x16x16_bf16: BF16 operands, FP32 accumulation), with
// its training step (forward, backward, deterministic gradient reduction,
// Adam on FP32 master weights), the fused root-feature + inference kernel,
// and the sampled exact check that sends failing tiles of roots back to the
// exact expression runtime.  No reference counterpart: the reference only
// has the exact runtime (src/xgen_expression.cpp:984-1148), which stays the
// fallback and the arbiter.
//
// A block of 256 lanes (four wavefronts) owns a tile of 64 samples; every
// layer is a 64 x N x K product split into 16 x 16 output tiles, four per
// wavefront.
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

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16u) & 1u);
    return static_cast<uint16_t>(u >> 16u);
}
__device__ __forceinline__ float sigmoid(float z) { return 1.0f / (1.0f + __expf(-z)); }

// One 16 x 16 output tile of C = A(64 x K) * B(K x N): a(m, k) and b(k, n)
// return BF16 bits.  Result: acc[j] = C[tm*16 + 4*(lane/16) + j][tn*16 + lane%16].
template <class FA, class FB>
__device__ __forceinline__ v4f mfma_tile(int tm, int tn, int K, const FA &a, const FB &b, int lane) {
    v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
    const int r = lane & 15, q = lane >> 4;
    for (int kk = 0; kk < K; kk += 16) {
        v4s av, bv;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            av[j] = static_cast<short>(a(tm * 16 + r, kk + 4 * q + j));
            bv[j] = static_cast<short>(b(kk + 4 * q + j, tn * 16 + r));
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
    }
    return acc;
}

struct MlpState {
    float *params;          // FP32 master parameters
    float *adam_m, *adam_v;
    uint16_t *params_bf16;  // BF16 copy used by the kernels
    float *norm;            // in_mean[10], in_inv_std[10], out_scale[4], out_offset[4]
    float *partials;        // [max_blocks][kParams + 1] per-block gradients and loss
    float *loss;            // mean squared error of the last step (normalised targets)
};

// Forward pass of one 64-sample tile held in LDS; h0: [64][kInPad] BF16
// normalised inputs, y: [64][kOutPad] normalised outputs.  The training form
// keeps every hidden layer's z (FP32, for SiLU') and h (BF16); the
// inference form ping-pongs two activation buffers.
struct TileBuffers {
    uint16_t h0[kTile * kInPad];
    uint16_t h[NXG_MLP_HIDDEN_LAYERS][kTile * kHid];
    float z[NXG_MLP_HIDDEN_LAYERS][kTile * kHid];
    float y[kTile * kOutPad];
};
struct InferBuffers {
    uint16_t h0[kTile * kInPad];
    uint16_t h[2][kTile * kHid];
    float y[kTile * kOutPad];
};

template <bool Train, class Buf>
__device__ void tile_forward(const uint16_t *W, const float *P, Buf &t) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int l = 0; l < kLayers; ++l) {
        const int K = rows_of(l), N = cols_of(l);
        const uint16_t *Wl = W + w_off(l);
        const float *bl = P + b_off(l);
        const uint16_t *hin = l == 0 ? t.h0 : t.h[Train ? l - 1 : (l - 1) & 1];
        uint16_t *hout = l < kLayers - 1 ? t.h[Train ? l : l & 1] : nullptr;
        const int ldin = l == 0 ? kInPad : kHid;
        const int tiles = (kTile / 16) * (N / 16);
        for (int tt = wave; tt < tiles; tt += 4) {
            const int tm = tt / (N / 16), tn = tt % (N / 16);
            const v4f acc = mfma_tile(
                tm, tn, K, [&](int m, int k) { return hin[m * ldin + k]; },
                [&](int k, int n) { return Wl[k * N + n]; }, lane);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = tm * 16 + 4 * (lane >> 4) + j, n = tn * 16 + (lane & 15);
                const float z = acc[j] + bl[n];
                if (l < kLayers - 1) {
                    if constexpr (Train) t.z[l][m * kHid + n] = z;
                    hout[m * kHid + n] = f2bf(z * sigmoid(z));
                } else {
                    t.y[m * kOutPad + n] = z;
                }
            }
        }
        __syncthreads();
    }
}

struct TrainBuffers {
    TileBuffers f;
    uint16_t dz[kTile * kHid];  // BF16 gradient of the current layer's pre-activation
    float dh[kTile * kHid];     // FP32 gradient of the previous layer's output
};

// Normalised features of sample rows into h0 (BF16), zero padding.
__device__ void load_inputs(const float *features, uint64_t first, uint64_t count, const float *norm,
                            uint16_t *h0) {
    for (int i = threadIdx.x; i < kTile * kInPad; i += blockDim.x) {
        const int m = i / kInPad, k = i % kInPad;
        const uint64_t s = first + m;
        float v = 0.0f;
        if (k < kIn && s < count) v = (features[s * kIn + k] - norm[k]) * norm[kIn + k];
        h0[i] = f2bf(v);
    }
    __syncthreads();
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
    // rows of this tile (sample indices wrap around the data set)
    for (int i = threadIdx.x; i < kTile * kInPad; i += blockDim.x) {
        const int m = i / kInPad, k = i % kInPad;
        const uint64_t sidx = (tile_first + m) % count;
        float v = 0.0f;
        if (k < kIn && static_cast<uint32_t>(m) < valid) v = (features[sidx * kIn + k] - s.norm[k]) * s.norm[kIn + k];
        t.f.h0[i] = f2bf(v);
    }
    __syncthreads();
    tile_forward<true>(s.params_bf16, s.params, t.f);
    float *part = s.partials + uint64_t{blockIdx.x} * (kParams + 1);
    // output gradient dy = 2 (y - t) / (batch * 4) on valid rows; loss partial
    float loss = 0.0f;
    for (int i = threadIdx.x; i < kTile * kOutPad; i += blockDim.x) {
        const int m = i / kOutPad, n = i % kOutPad;
        float d = 0.0f;
        if (n < kOut && static_cast<uint32_t>(m) < valid) {
            const uint64_t sidx = (tile_first + m) % count;
            const float tn = (targets[sidx * kOut + n] - s.norm[2 * kIn + kOut + n]) / s.norm[2 * kIn + n];
            const float diff = t.f.y[i] - tn;
            loss += diff * diff;
            d = 2.0f * diff / (static_cast<float>(batch) * kOut);
        }
        t.dh[m * kHid + n] = d;  // reuse dh as the FP32 output gradient
    }
    // block loss reduction (fixed order: per-wave shuffle, then wave 0)
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
    __shared__ float wave_loss[4];
    if (lane == 0) wave_loss[wave] = loss;
    __syncthreads();
    if (threadIdx.x == 0) part[kParams] = wave_loss[0] + wave_loss[1] + wave_loss[2] + wave_loss[3];

    for (int l = kLayers - 1; l >= 0; --l) {
        const int K = rows_of(l), N = cols_of(l);
        // dz = dh * silu'(z) (identity on the output layer), FP32 -> BF16
        for (int i = threadIdx.x; i < kTile * N; i += blockDim.x) {
            const int m = i / N, n = i % N;
            float d = t.dh[m * kHid + n];
            if (l < kLayers - 1) {
                const float z = t.f.z[l][m * kHid + n];
                const float sg = sigmoid(z);
                d *= sg * (1.0f + z * (1.0f - sg));
            }
            t.dz[m * kHid + n] = f2bf(d);
            t.dh[m * kHid + n] = d;  // keep the FP32 dz for the bias gradient
        }
        __syncthreads();
        // bias gradient: column sums of FP32 dz, fixed order
        for (int n = threadIdx.x; n < N; n += blockDim.x) {
            float acc = 0.0f;
            for (int m = 0; m < kTile; ++m) acc += t.dh[m * kHid + n];
            part[b_off(l) + n] = acc;
        }
        // weight gradient dW[k][n] = sum_m h_in[m][k] dz[m][n]  (K x N, contraction over samples)
        const uint16_t *hin = l == 0 ? t.f.h0 : t.f.h[l - 1];
        const int ldin = l == 0 ? kInPad : kHid;
        const int wt = (K / 16) * (N / 16);
        for (int tt = wave; tt < wt; tt += 4) {
            const int tk = tt / (N / 16), tn = tt % (N / 16);
            // A(row=k, col=m) = h_in[m][k]; B(row=m, col=n) = dz[m][n]; contraction length 64
            const v4f acc = mfma_tile(
                tk, tn, kTile, [&](int k, int m) { return hin[m * ldin + k]; },
                [&](int m, int n) { return t.dz[m * kHid + n]; }, lane);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int k = tk * 16 + 4 * (lane >> 4) + j, n = tn * 16 + (lane & 15);
                part[w_off(l) + k * N + n] = acc[j];
            }
        }
        __syncthreads();
        if (l == 0) break;
        // dh_prev[m][k] = sum_n dz[m][n] W[k][n]  (64 x K, contraction over N)
        const uint16_t *Wl = s.params_bf16 + w_off(l);
        const int ht = (kTile / 16) * (K / 16);
        for (int tt = wave; tt < ht; tt += 4) {
            const int tm = tt / (K / 16), tk = tt % (K / 16);
            const v4f acc = mfma_tile(
                tm, tk, N, [&](int m, int n) { return t.dz[m * kHid + n]; },
                [&](int n, int k) { return Wl[k * N + n]; }, lane);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int m = tm * 16 + 4 * (lane >> 4) + j, k = tk * 16 + (lane & 15);
                t.dh[m * kHid + k] = acc[j];
            }
        }
        __syncthreads();
    }
}

// Reduce the per-block partials in block order, then Adam on the FP32
// master parameters and refresh of the BF16 copy.
__global__ void __launch_bounds__(256) mlp_adam_kernel(MlpState s, uint32_t blocks, uint32_t batch, uint32_t step,
                                                       float lr, float beta1, float beta2, float eps) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        float loss = 0.0f;
        for (uint32_t b = 0; b < blocks; ++b) loss += s.partials[uint64_t{b} * (kParams + 1) + kParams];
        s.loss[0] = loss / (static_cast<float>(batch) * kOut);
    }
    if (i >= static_cast<uint32_t>(kParams)) return;
    float g = 0.0f;
    for (uint32_t b = 0; b < blocks; ++b) g += s.partials[uint64_t{b} * (kParams + 1) + i];
    const float m = beta1 * s.adam_m[i] + (1.0f - beta1) * g;
    const float v = beta2 * s.adam_v[i] + (1.0f - beta2) * g * g;
    s.adam_m[i] = m;
    s.adam_v[i] = v;
    const float mh = m / (1.0f - powf(beta1, static_cast<float>(step)));
    const float vh = v / (1.0f - powf(beta2, static_cast<float>(step)));
    const float p = s.params[i] - lr * mh / (sqrtf(vh) + eps);
    s.params[i] = p;
    s.params_bf16[i] = f2bf(p);
}

__global__ void mlp_sync_kernel(MlpState s) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < static_cast<uint32_t>(kParams)) s.params_bf16[i] = f2bf(s.params[i]);
}

// Inference over a feature array: outputs are attributes (denormalised).
__global__ void __launch_bounds__(256) mlp_forward_kernel(MlpState s, const float *features, float *outputs,
                                                          uint64_t count) {
    extern __shared__ uint8_t smem[];
    InferBuffers &t = *reinterpret_cast<InferBuffers *>(smem);
    const uint64_t first = uint64_t{blockIdx.x} * kTile;
    load_inputs(features, first, count, s.norm, t.h0);
    tile_forward<false>(s.params_bf16, s.params, t);
    for (int i = threadIdx.x; i < kTile * kOut; i += blockDim.x) {
        const int m = i / kOut, n = i % kOut;
        if (first + m < count)
            outputs[(first + m) * kOut + n] = t.y[m * kOutPad + n] * s.norm[2 * kIn + n] + s.norm[2 * kIn + kOut + n];
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

// Fused: features of 64 roots, normalisation, MLP, attributes.
__global__ void __launch_bounds__(256) surrogate_kernel(FeatureArgs a, MlpState s, float4 *attributes) {
    extern __shared__ uint8_t smem[];
    InferBuffers &t = *reinterpret_cast<InferBuffers *>(smem);
    const uint64_t first = uint64_t{blockIdx.x} * kTile;
    if (threadIdx.x < kTile) {
        const uint64_t sidx = first + threadIdx.x;
        float f[kIn];
        root_features(a, static_cast<uint32_t>(sidx < a.strand_count ? sidx : a.strand_count - 1u), f);
        for (int k = 0; k < kInPad; ++k)
            t.h0[threadIdx.x * kInPad + k] = k < kIn ? f2bf((f[k] - s.norm[k]) * s.norm[kIn + k]) : 0;
    }
    __syncthreads();
    tile_forward<false>(s.params_bf16, s.params, t);
    if (threadIdx.x < kTile && first + threadIdx.x < a.strand_count) {
        const int m = threadIdx.x;
        float o[kOut];
        for (int n = 0; n < kOut; ++n) o[n] = t.y[m * kOutPad + n] * s.norm[2 * kIn + n] + s.norm[2 * kIn + kOut + n];
        attributes[first + m] = make_float4(o[0], o[1], o[2], o[3]);
    }
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
    L.norm_offset = align(L.params_bf16_offset + P * 2u);
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
    hipLaunchKernelGGL(nxg::mlp_adam_kernel, dim3((nxg::kParams + 255) / 256), dim3(256), 0, s, st, blocks,
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
    hipLaunchKernelGGL(nxg::mlp_forward_kernel, dim3(static_cast<uint32_t>((count + 63u) / 64u)), dim3(256),
                       sizeof(nxg::InferBuffers), static_cast<hipStream_t>(stream),
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
                                  uint64_t attribute_capacity, void *stream) {
    if (uint32_t s = nxg::validate_common(asset, nullptr, params, nullptr); s != NXG_OK) return s;
    if (!state || !map || !map->texels || map->face_count == 0u || map->resolution == 0u)
        return NXG_ERROR_NULL_ARGUMENT;
    if (!attributes || attribute_capacity < uint64_t{params->strand_count} * NXG_ATTR_COUNT)
        return NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL;
    hipLaunchKernelGGL(nxg::surrogate_kernel, dim3((params->strand_count + 63u) / 64u), dim3(256),
                       sizeof(nxg::InferBuffers), static_cast<hipStream_t>(stream),
                       nxg::feature_args(*asset, *params, *map), nxg::mlp_state_of(const_cast<void *>(state)),
                       reinterpret_cast<float4 *>(attributes));
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

}  // extern "C"
