// Classic CutFX on generated curves (gfx950): cLength, culling, cut from the
// tip, uniform rebuild and stable compaction, in the operation order of the
// reference's CPU float runtime (src/xgen_classic_runtime.cpp:587-663 spline
// evaluation and length, 692-743 cut_from_tip_and_rebuild, 1648-1678 and
// 1806-1821 the per-strand chain, 1893-1918 compaction), so positions are
// bit-identical to it (no FP contraction, correctly rounded div/sqrt).
//
// Five launches on the caller's stream, none of them allocating:
//   cut_table_kernel    the spline samples every strand shares: all strands
//                       of a call have n CVs, so cLength's 2n+3 parameters
//                       k/(2n+4) and the cut walk's parameters 1 - j/(2n+4)
//                       (each step subtracted from the last, as the walk
//                       does), their spans and B-spline weights are the same
//                       for every strand; one block writes them to the
//                       workspace, and the search reads them with scalar
//                       loads instead of recomputing them in every lane.
//   cut_search_kernel   one 64-lane wavefront per tile of 64 strands.  The
//                       tile's points are one contiguous run of 64 x n
//                       float4s: the wave reads it coalesced into LDS,
//                       CV-major with row c rotated by c lanes, so the
//                       staging stores are bank-conflict-free, and so are
//                       the reads (every lane reads the same row).  Each lane
//                       measures its spline (2n+4 chords), culls, walks
//                       back from the tip and stores its cut parameter
//                       (-1: culled); lane 0 stores the tile's kept count
//                       and the error bits the tile raised.
//   cut_scan_tiles      exclusive scan of the tile counts, kScanBlock tiles
//                       per block, each block's total into the group array.
//   cut_scan_groups     one block: exclusive scan of the group totals, the
//                       kept count and the call's status word.
//   cut_write_kernel    the tile again through LDS; each kept lane resamples
//                       its n CVs on [0, cut] (its own parameters, so no
//                       shared table) and writes them, with its point
//                       count, root and root uv, at its compacted slot
//                       (group prefix + tile prefix + rank in the wave).
//                       Up to 32 CVs (cut_write_held_kernel) the resampled
//                       strand stays in registers until the whole wave is
//                       done reading, then goes back through the same LDS in
//                       output order, so the wave's stores are one
//                       contiguous run.
// Above 64 CVs a tile no longer fits 64 KB of LDS and the lanes read their
// CVs from memory instead (same arithmetic).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <type_traits>

#include "../../include/nanoxgen_amd.h"
#include "block_scan.h"
#include "expr_device.h"
#include "internal.h"

namespace nxg {
namespace {

constexpr uint32_t kWave = kClassicTile;  // strands per tile = lanes per wavefront
constexpr uint32_t kStagedCvs = 64u;      // 64 strands x 64 CVs x 16 B = 64 KB of LDS
constexpr uint32_t kErrShift = 30u;
constexpr uint32_t kCountMask = (1u << kErrShift) - 1u;

struct F3 {
    float x, y, z;
};
__device__ __forceinline__ F3 xyz(float4 v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ F3 operator*(F3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ float length_squared(F3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
__device__ __forceinline__ float max_std(float a, float b) { return a < b ? b : a; }  // std::max

// One spline sample shared by every strand of a call (all have n CVs): the
// parameter and, for interior samples, the blended span and its weights, as
// xgen_curve_eval computes them (src/xgen_classic_runtime.cpp:614-642).
struct Sample {
    float b0, b1, b2, b3;
    float t;
    float previous_t;  // cut walk: the parameter one step nearer the tip
    uint32_t span;     // CVs span-1 .. span+2 are blended
    uint32_t kind;     // kAtRoot, kAtTip, or the kReflect* flags of an interior sample
};
static_assert(sizeof(Sample) == 32, "one 32-byte scalar load per sample");
constexpr uint32_t kAtRoot = 1u, kAtTip = 2u, kReflectLow = 4u, kReflectHigh = 8u;

// One ClumpingFX module's geometry binding; its guides own entries
// guide_base .. guide_base + guide_count - 1 of the per-guide workspace arrays.
struct ClumpModule {
    uint32_t guide_count, guide_base;
    const uint32_t *strand_guides;
    const nxg_vec3 *axes, *normals, *tangents, *reference;
    const nxg_vec2 *uvs;
    const uint32_t *prefixes;
    const nxg_vec3 *local_axes;  // optional translation-free guide CVs (their spline lengths)
    uint32_t local_frame;        // root_relative with local_axes: render and noise axes from them
};

struct CutArgs {
    const float4 *points;
    const uint32_t *counts;
    const nxg_root_sample *roots;
    const nxg_vec2 *uvs;
    uint32_t strands, cvs;
    uint32_t div_magic;  // i / cvs == __umulhi(i, div_magic) for i < 2^24
    float scale, offset, radius_scale;
    const float *multiplier;
    uint32_t multiplier_stride;
    float *cut;
    uint32_t *tiles, *groups, *status_word;
    uint32_t tile_count, group_count;
    Sample *length_samples;  // 2n+3 entries: k/(2n+4), k = 1..2n+3
    Sample *walk_samples;    // walk_count entries, the last at parameter <= 1e-10
    uint32_t walk_count;
    float4 *out_points;
    uint32_t *out_counts;
    nxg_root_sample *out_roots;
    nxg_vec2 *out_uvs;
    uint32_t *kept, *status;
    // Classic primitive expressions (nxg_classic_runtime_apply only)
    const uint8_t *image;        // packed plan, device
    uint32_t program_offset[NXG_CLASSIC_EXPR_COUNT];
    uint64_t present;            // bit k: program k given
    uint32_t has_cut, final_length, profile;  // has_cut: the write resamples (a trailing CutFX)
    const uint32_t *ids, *prefixes;
    const float *columns;
    uint32_t column_count;
    float description_id, name_seed;
    uint32_t uniform;            // programs whose value depends on t alone
    float2 *uniform_values;      // {value, ok} per uniform program (uniform_slot)
    struct RuntimeState *state;  // per strand, from the search to the write
    uint32_t *err_bits;          // the NXG_CLASSIC_* bits of every failing strand
    // the effects in order: type << 8 | module (NXG_CLASSIC_EFFECT_*)
    uint32_t effect_count;
    uint32_t effects[NXG_CLASSIC_MAX_EFFECTS];
    uint32_t rewrites;           // some effect other than a trailing CutFX changes the CVs
    uint32_t root_relative;      // clump goals about each strand's root
    uint32_t clump_count;
    const nxg_vec3 *surface_tangents, *noise_domain;
    uint32_t clump_length;       // bit m: module m's amount or clumpScale reads cLength
    uint32_t defer_length;       // the first effect is a ClumpingFX that re-measures cLength, and no
                                 // program before its re-measure reads cLength (see runtime_strand)
    ClumpModule clumps[NXG_CLASSIC_MAX_CLUMPS];
    float4 *clumped;             // [strands x cvs] every strand after the last module
    float4 *guide_render;        // [all modules' guides x cvs] render axes
    float4 *guide_noise;         // [all modules' guides x cvs] unit clump-noise axes
    float2 *guide_info;          // [all modules' guides] {spline length, 1 if its noise frequency/correlation evaluated}
};

// A lane's CVs: staged in LDS (CV-major; ROTATED: row c turned by c lanes),
// or read from memory.  Rotated rows make the coalesced staging stores
// conflict-free; plain rows make reads of different rows by different lanes
// conflict-free.  The search reads one row per step for all lanes, the
// rebuild reads per-lane rows.
template <bool ROTATED>
__device__ __forceinline__ uint32_t staged_index(uint32_t c, uint32_t lane) {
    return ROTATED ? c * kWave + ((lane + c) & (kWave - 1u)) : c * kWave + lane;
}
template <bool ROTATED>
struct StagedCurve {
    float4 *tile;
    uint32_t lane;
    __device__ __forceinline__ float4 operator[](uint32_t c) const { return tile[staged_index<ROTATED>(c, lane)]; }
};
struct MemoryCurve {
    const float4 *strand;
    __device__ __forceinline__ float4 operator[](uint32_t c) const { return strand[c]; }
};

// Loads in flight per lane while a tile is staged: a loop that stores each
// load before issuing the next waits one memory round trip per 64 CVs.
constexpr uint32_t kStageBatch = 8u;

// The tile's `count` strands, read coalesced, into LDS.
template <bool ROTATED>
__device__ __forceinline__ void stage_tile(const CutArgs &a, float4 *tile, uint32_t first, uint32_t count) {
    const float4 *src = a.points + uint64_t{first} * a.cvs;
    const uint32_t total = count * a.cvs;
    for (uint32_t base = threadIdx.x; base < total; base += kStageBatch * kWave) {
        // no branches, so every load issues before the first store: a lane
        // past the end loads and stores the last CV again (the same value
        // to the same place)
        float4 v[kStageBatch];
#pragma unroll
        for (uint32_t j = 0u; j < kStageBatch; ++j) v[j] = src[min(base + j * kWave, total - 1u)];
#pragma unroll
        for (uint32_t j = 0u; j < kStageBatch; ++j) {
            const uint32_t i = min(base + j * kWave, total - 1u);
            const uint32_t l = __umulhi(i, a.div_magic);
            tile[staged_index<ROTATED>(i - l * a.cvs, l)] = v[j];
        }
    }
    __syncthreads();
}

// xgen_curve_eval's parameter-only half: span and weights at t.
__device__ __forceinline__ Sample make_sample(uint32_t n, float t, float previous_t) {
    Sample s{};
    s.t = t;
    s.previous_t = previous_t;
    if (t < 1.0e-7f) {
        s.kind = kAtRoot;
        return s;
    }
    if (t > 1.0f - 1.0e-7f) {
        s.kind = kAtTip;
        return s;
    }
    const uint32_t spans = n - 1u;
    const float scaled = t * static_cast<float>(spans);
    const uint32_t raw = static_cast<uint32_t>(scaled);
    const float f = scaled - static_cast<float>(raw);
    s.span = raw < n - 2u ? raw : n - 2u;  // t < 1 keeps raw <= n - 2; the clamp guards memory only
    const float f2 = f * f;
    const float f3 = f2 * f;
    const float omf = 1.0f - f;
    s.b0 = omf * omf * omf;
    s.b1 = 3.0f * f3 - 6.0f * f2 + 4.0f;
    s.b2 = -3.0f * f3 + 3.0f * f2 + 3.0f * f + 1.0f;
    s.b3 = f3;
    s.kind = (s.span == 0u ? kReflectLow : 0u) | (s.span + 2u >= n ? kReflectHigh : 0u);
    return s;
}

// xgen_curve_eval's point half.  Sample fields that are uniform across the
// wave (a shared table entry) make every branch here a scalar one.
template <class Curve>
__device__ __forceinline__ F3 blend(const Curve &p, uint32_t n, const Sample &s) {
    if (s.kind & kAtRoot) return xyz(p[0]);
    if (s.kind & kAtTip) return xyz(p[n - 1u]);
    const F3 p1 = xyz(p[s.span]);
    const F3 p2 = xyz(p[s.span + 1u]);
    const F3 p0 = (s.kind & kReflectLow) ? p1 * 2.0f - p2 : xyz(p[s.span - 1u]);
    const F3 p3 = (s.kind & kReflectHigh) ? p2 * 2.0f - p1 : xyz(p[s.span + 2u]);
    return (p0 * s.b0 + p1 * s.b1 + p2 * s.b2 + p3 * s.b3) * (1.0f / 6.0f);
}

// xgen_curve_eval (src/xgen_classic_runtime.cpp:614-642) at a per-lane t.
template <class Curve>
__device__ __forceinline__ F3 curve_eval(const Curve &p, uint32_t n, float t) {
    if (t < 1.0e-7f) return xyz(p[0]);
    if (t > 1.0f - 1.0e-7f) return xyz(p[n - 1u]);
    const Sample s = make_sample(n, t, 0.0f);
    const F3 p1 = xyz(p[s.span]);
    const F3 p2 = xyz(p[s.span + 1u]);
    const F3 q0 = xyz(p[s.span == 0u ? 0u : s.span - 1u]);
    const F3 q3 = xyz(p[s.span + 2u >= n ? n - 1u : s.span + 2u]);
    const F3 p0 = (s.kind & kReflectLow) ? p1 * 2.0f - p2 : q0;
    const F3 p3 = (s.kind & kReflectHigh) ? p2 * 2.0f - p1 : q3;
    return (p0 * s.b0 + p1 * s.b1 + p2 * s.b2 + p3 * s.b3) * (1.0f / 6.0f);
}

// A shared sample through the scalar cache: the index is uniform, so the
// entry lands in scalar registers and the blend's branches are scalar.
typedef const __attribute__((address_space(4))) uint32_t *ConstWords;
__device__ __forceinline__ Sample load_sample(const Sample *table, uint32_t k) {
    ConstWords w = (ConstWords)(table + k);
    uint32_t v[8];
#pragma unroll
    for (uint32_t j = 0; j < 8u; ++j) v[j] = w[j];
    Sample s;
    __builtin_memcpy(&s, v, sizeof(s));
    return s;
}

// curve_spline_length (src/xgen_classic_runtime.cpp:644-663) over the shared samples.
template <class Curve>
__device__ __forceinline__ float spline_length(const Sample *samples, const Curve &p, uint32_t n) {
    F3 previous = xyz(p[0]);
    float result = 0.0f;
#pragma unroll 4
    for (uint32_t k = 0u; k < 2u * n + 3u; ++k) {
        const F3 current = blend(p, n, load_sample(samples, k));
        result += sqrtf(length_squared(current - previous));
        previous = current;
    }
    result += sqrtf(length_squared(xyz(p[n - 1u]) - previous));
    return result;
}

// The search half of cut_from_tip_and_rebuild (src/xgen_classic_runtime.cpp:
// 701-728) over the shared walk samples: the rebuild parameter for a finite
// amount >= 0 (0: nothing left).  The last walk sample has parameter
// <= 1e-10, where every lane still walking stops.
template <class Curve>
__device__ __forceinline__ float cut_parameter(const Sample *walk, uint32_t walk_count, const Curve &p, uint32_t n,
                                               float amount) {
    if (!(amount >= 1.0e-10f)) return 1.0f;
    F3 previous = xyz(p[n - 1u]);
    float accumulated = 0.0f;
#pragma unroll 4
    for (uint32_t j = 0u; j < walk_count; ++j) {
        const Sample s = load_sample(walk, j);
        const F3 current = blend(p, n, s);
        const float segment = sqrtf(length_squared(previous - current));
        accumulated += segment;
        if (accumulated >= amount && segment > 0.0f)
            return s.t + ((accumulated - amount) / segment) * (s.previous_t - s.t);
        if (s.t <= 1.0e-10f) return 0.0f;
        previous = current;
    }
    return 0.0f;
}

// One strand of the chain (src/xgen_classic_runtime.cpp:1650-1678, 1806-1821):
// the cut parameter of a kept strand, -1 for a culled one; *err gets the
// NXG_CLASSIC_NONFINITE_* bit where the reference throws.
template <class Curve>
__device__ __forceinline__ float strand_cut(const CutArgs &a, const Sample *samples, const Curve &p, uint32_t s,
                                            uint32_t *err) {
    const uint32_t n = a.cvs;
    const float c_length = spline_length(samples, p, n);
    const float radius = p[0].w;
    const float width = a.radius_scale > 0.0f ? 2.0f * radius / a.radius_scale : 0.0f;
    if (!isfinite(width)) {
        *err = NXG_CLASSIC_NONFINITE_WIDTH;
        return -1.0f;
    }
    if (c_length < 1.0e-4f || width < 1.0e-4f) return -1.0f;
    float amount = a.scale * c_length + a.offset;
    if (a.multiplier) amount = amount * a.multiplier[uint64_t{s} * a.multiplier_stride];
    if (!isfinite(amount)) {
        *err = NXG_CLASSIC_NONFINITE_AMOUNT;
        return -1.0f;
    }
    const float cut = cut_parameter(a.walk_samples, a.walk_count, p, n, amount < 0.0f ? 0.0f : amount);
    return cut < 1.0e-4f ? -1.0f : cut;
}

// The shared samples: cLength's k/(2n+4) in parallel, the walk's parameters
// by the walk's own repeated subtraction (one lane), then their weights.
__global__ void __launch_bounds__(256) cut_table_kernel(CutArgs a) {
    __shared__ float walk_t[kClassicTableEntries];
    const uint32_t n = a.cvs;
    const uint32_t intervals = n * 2u + 4u;
    const float step = 1.0f / static_cast<float>(intervals);
    for (uint32_t k = threadIdx.x + 1u; k < intervals; k += blockDim.x)
        a.length_samples[k - 1u] = make_sample(n, step * static_cast<float>(k), 0.0f);
    if (threadIdx.x == 0u) {
        const float walk_step = 1.0f / (2.0f * static_cast<float>(n) + 4.0f);
        float previous = 1.0f;
        for (uint32_t j = 0u; j < a.walk_count; ++j) {
            const float lowered = previous - walk_step;
            walk_t[j] = lowered < 0.0f ? 0.0f : lowered;
            previous = walk_t[j];
        }
    }
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < a.walk_count; j += blockDim.x)
        a.walk_samples[j] = make_sample(n, walk_t[j], j ? walk_t[j - 1u] : 1.0f);
}

__global__ void clear_word_kernel(uint32_t *word) { *word = 0u; }

template <bool STAGED>
__global__ void __launch_bounds__(kWave) cut_search_kernel(CutArgs a) {
    extern __shared__ float4 tile[];
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    if constexpr (STAGED) stage_tile<true>(a, tile, first, count);
    const uint32_t s = first + lane;
    float cut = -1.0f;
    uint32_t err = 0u;
    if (lane < count) {
        if constexpr (STAGED)
            cut = strand_cut(a, a.length_samples, StagedCurve<true>{tile, lane}, s, &err);
        else
            cut = strand_cut(a, a.length_samples, MemoryCurve{a.points + uint64_t{s} * n}, s, &err);
        a.cut[s] = cut;
    }
    const uint64_t kept = __ballot(cut >= 0.0f);
    const uint32_t bits = (__ballot(err & NXG_CLASSIC_NONFINITE_AMOUNT) ? NXG_CLASSIC_NONFINITE_AMOUNT : 0u) |
                          (__ballot(err & NXG_CLASSIC_NONFINITE_WIDTH) ? NXG_CLASSIC_NONFINITE_WIDTH : 0u);
    if (lane == 0u) a.tiles[t] = static_cast<uint32_t>(__popcll(kept)) | (bits << kErrShift);
}

__global__ void __launch_bounds__(kScanThreads) cut_scan_tiles_kernel(CutArgs a) {
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * 4u;
    uint32_t v[4], sum = 0u, flags = 0u;
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        const uint32_t w = base + j < a.tile_count ? a.tiles[base + j] : 0u;
        v[j] = w & kCountMask;
        flags |= w >> kErrShift;
        sum += v[j];
    }
    uint32_t total, all_flags;
    uint32_t run = block_scan(sum, flags, total, all_flags);
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        if (base + j < a.tile_count) a.tiles[base + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 0u) a.groups[blockIdx.x] = total | (all_flags << kErrShift);
}

__global__ void __launch_bounds__(kScanThreads) cut_scan_groups_kernel(CutArgs a) {
    uint32_t carry = 0u, flags = 0u;
    for (uint32_t first = 0u; first < a.group_count; first += kScanThreads) {
        const uint32_t i = first + threadIdx.x;
        const uint32_t w = i < a.group_count ? a.groups[i] : 0u;
        uint32_t total, all_flags;
        const uint32_t ex = block_scan(w & kCountMask, w >> kErrShift, total, all_flags);
        if (i < a.group_count) a.groups[i] = carry + ex;
        carry += total;
        flags |= all_flags;
    }
    if (threadIdx.x == 0u) {
        // the runtime path marks failing tiles with one flag and keeps the
        // bits themselves in err_bits
        const uint32_t bits = flags && a.err_bits ? *a.err_bits : flags;
        *a.status_word = bits;
        if (a.status) *a.status = bits;
        *a.kept = bits ? 0u : carry;
    }
}

template <bool STAGED>
__global__ void __launch_bounds__(kWave) cut_write_kernel(CutArgs a) {
    extern __shared__ float4 tile[];
    if (*a.status_word != 0u) return;  // the reference throws: nothing is written
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    const uint32_t s = first + lane;
    const float cut = lane < count ? a.cut[s] : -1.0f;
    const uint64_t kept = __ballot(cut >= 0.0f);
    if (kept == 0u) return;  // the block is this one wavefront
    if constexpr (STAGED) stage_tile<false>(a, tile, first, count);
    if (!(cut >= 0.0f)) return;
    const uint32_t rank = static_cast<uint32_t>(__popcll(kept & ((uint64_t{1} << lane) - 1u)));
    const uint32_t slot = a.groups[t / kScanBlock] + a.tiles[t] + rank;
    float4 *dst = a.out_points + uint64_t{slot} * n;
    auto rebuild = [&](const auto &p) {  // src/xgen_classic_runtime.cpp:729-742
        const float step = cut / static_cast<float>(n - 1u);
        dst[0] = p[0];
        for (uint32_t i = 1u; i < n; ++i) {
            const F3 q = curve_eval(p, n, step * static_cast<float>(i));
            dst[i] = make_float4(q.x, q.y, q.z, p[i].w);
        }
    };
    if constexpr (STAGED)
        rebuild(StagedCurve<false>{tile, lane});
    else
        rebuild(MemoryCurve{a.points + uint64_t{s} * n});
    if (a.out_counts) a.out_counts[slot] = a.counts ? a.counts[s] : n;
    if (a.out_roots) a.out_roots[slot] = a.roots[s];
    if (a.out_uvs) a.out_uvs[slot] = a.uvs[s];
}

// The rebuild with each strand held in registers (n <= MAXN): once every
// lane has read its CVs, the tile's LDS is reused to lay the kept strands
// out in output order (rows of n + 1 float4s, bank-skewed), and the wave
// stores them as one contiguous run, so the point stores are coalesced.
template <uint32_t MAXN>
__global__ void __launch_bounds__(kWave) cut_write_held_kernel(CutArgs a) {
    extern __shared__ float4 tile[];
    if (*a.status_word != 0u) return;  // the reference throws: nothing is written
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    const uint32_t s = first + lane;
    const float cut = lane < count ? a.cut[s] : -1.0f;
    const bool keep = cut >= 0.0f;
    const uint64_t kept = __ballot(keep);
    if (kept == 0u) return;  // the block is this one wavefront
    stage_tile<false>(a, tile, first, count);
    const uint32_t rank = static_cast<uint32_t>(__popcll(kept & ((uint64_t{1} << lane) - 1u)));
    const uint32_t slot0 = a.groups[t / kScanBlock] + a.tiles[t];
    float4 held[MAXN];
    if (keep) {  // src/xgen_classic_runtime.cpp:729-742
        const StagedCurve<false> p{tile, lane};
        const float step = cut / static_cast<float>(n - 1u);
        held[0] = p[0];
#pragma unroll
        for (uint32_t i = 1u; i < MAXN; ++i) {
            if (i < n) {
                const F3 q = curve_eval(p, n, step * static_cast<float>(i));
                held[i] = make_float4(q.x, q.y, q.z, p[i].w);
            }
        }
    }
    __syncthreads();
    const uint32_t stride = n + 1u;
    if (keep) {
#pragma unroll
        for (uint32_t i = 0u; i < MAXN; ++i)
            if (i < n) tile[rank * stride + i] = held[i];
    }
    __syncthreads();
    const uint32_t total = static_cast<uint32_t>(__popcll(kept)) * n;
    float4 *dst = a.out_points + uint64_t{slot0} * n;
    for (uint32_t e = lane; e < total; e += kWave) {
        const uint32_t r = __umulhi(e, a.div_magic);
        dst[e] = tile[r * stride + (e - r * n)];
    }
    if (!keep) return;
    const uint32_t slot = slot0 + rank;
    if (a.out_counts) a.out_counts[slot] = a.counts ? a.counts[s] : n;
    if (a.out_roots) a.out_roots[slot] = a.roots[s];
    if (a.out_uvs) a.out_uvs[slot] = a.uvs[s];
}

// add_xgen_classic_renderer_endpoints (src/xgen_classic_runtime.cpp:
// 1921-1956): one 256-thread block per 64-strand tile, each thread one
// output point of the tile's contiguous 64 x (n + 2) run (input and output
// both read and written in order).
struct EndpointArgs {
    const float4 *points;
    float4 *out;
    uint32_t *counts;
    const uint32_t *device_count;
    uint32_t strands, cvs, magic;  // e / (cvs + 2) == __umulhi(e, magic) for e < 2^24
};

__global__ void __launch_bounds__(256) endpoints_kernel(EndpointArgs a) {
    const uint32_t strands = a.device_count ? min(*a.device_count, a.strands) : a.strands;
    const uint32_t first = blockIdx.x * kWave;
    if (first >= strands) return;
    const uint32_t count = strands - first < kWave ? strands - first : kWave;
    const uint32_t n = a.cvs, m = n + 2u;
    const float4 *src = a.points + uint64_t{first} * n;
    float4 *dst = a.out + uint64_t{first} * m;
    // output k of a strand: its CV k - 1, or at the ends 2 p0 - p1 with
    // (p0, p1) = (CV 0, CV 1) or (CV n-1, CV n-2).  Five outputs per thread
    // in flight, branch-free (every lane loads a p1, its own CV again inside
    // the strand; a lane past the end redoes the last output), so all loads
    // issue before the first store.
    constexpr uint32_t kBatch = 5u;
    const uint32_t total = count * m;
    for (uint32_t base = threadIdx.x; base < total; base += kBatch * 256u) {
        float4 p0[kBatch], p1[kBatch];
#pragma unroll
        for (uint32_t j = 0u; j < kBatch; ++j) {
            const uint32_t e = min(base + j * 256u, total - 1u);
            const uint32_t r = __umulhi(e, a.magic);
            const uint32_t k = e - r * m;
            const uint32_t i0 = k == 0u ? 0u : k == m - 1u ? n - 1u : k - 1u;
            const uint32_t i1 = k == 0u ? 1u : k == m - 1u ? n - 2u : i0;
            p0[j] = src[r * n + i0];
            p1[j] = src[r * n + i1];
        }
#pragma unroll
        for (uint32_t j = 0u; j < kBatch; ++j) {
            const uint32_t e = min(base + j * 256u, total - 1u);
            const uint32_t k = e - __umulhi(e, a.magic) * m;
            const float4 a0 = p0[j], a1 = p1[j];
            dst[e] = k == 0u || k == m - 1u
                         ? make_float4(2.0f * a0.x - a1.x, 2.0f * a0.y - a1.y, 2.0f * a0.z - a1.z, a0.w)
                         : a0;
        }
    }
    if (a.counts && threadIdx.x < count) a.counts[first + threadIdx.x] = m;
}

// ---- ClumpingFX ---------------------------------------------------------

__constant__ float4 kClumpGradients[256] = {
#define NXG_GRADIENT(x, y, z) {x, y, z, 0.0f},
#include "seexpr_gradients.inc"
#undef NXG_GRADIENT
};

struct ClumpArgs {
    const float4 *points;
    float4 *out;
    const uint32_t *strand_guides;
    const nxg_vec3 *axes, *normals, *tangents, *reference;
    uint32_t strands, cvs, guides, div_magic;
    float mask;        // clamped to [0, 1]
    float amount, clump_scale;
    float magnitude;   // mask * noise * max(noise_scale, 0); 0 when the noise axis is off
    float frequency;   // max(noise_frequency, 0)
    float correlation; // clamp(noise_correlation * 0.01, 0, 1)
    const float *amount_multiplier;
    uint32_t amount_stride;
    float4 *render;    // [guides x cvs] guide render axes
    float4 *noise;     // [guides x cvs] clump-noise axes, already scaled
    float *lengths;    // [guides] guide spline lengths
    const Sample *length_samples;
    uint32_t *status;
};

struct Vec3Curve {
    const nxg_vec3 *p;
    __device__ __forceinline__ float4 operator[](uint32_t c) const { return make_float4(p[c].x, p[c].y, p[c].z, 0.0f); }
};

__device__ __forceinline__ float dot3(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ F3 cross3(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ F3 normalize3(F3 v) {  // generate.h:67-70
    const float l2 = length_squared(v);
    return l2 > 1.0e-20f ? F3{v.x / sqrtf(l2), v.y / sqrtf(l2), v.z / sqrtf(l2)} : F3{0.0f, 1.0f, 0.0f};
}

// xgen_classic_noise_float (src/xgen_classic_runtime.cpp:802-837)
__device__ float classic_noise(float sx, float sy, float sz) {
    const float fx = floorf(sx), fy = floorf(sy), fz = floorf(sz);
    const int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy), iz = static_cast<int32_t>(fz);
    const float wx = sx - fx, wy = sy - fy, wz = sz - fz;
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        uint32_t seed = 0u;  // noise_hash (generate.h:353-364)
        seed = seed * 1664525u + static_cast<uint32_t>(ix + ox) + 1013904223u;
        seed = seed * 1664525u + static_cast<uint32_t>(iy + oy) + 1013904223u;
        seed = seed * 1664525u + static_cast<uint32_t>(iz + oz) + 1013904223u;
        seed ^= seed >> 11u;
        seed = xor_and(seed, seed << 7u, 0x9d2c5680u);
        seed = xor_and(seed, seed << 15u, 0xefc60000u);
        seed ^= seed >> 18u;
        const float4 g = kClumpGradients[(((seed & 0x00ff0000u) >> 4u) + (seed & 0xffu)) & 0xffu];
        v[c] = g.x * (wx - static_cast<float>(ox)) + g.y * (wy - static_cast<float>(oy)) +
               g.z * (wz - static_cast<float>(oz));
    }
    const float ax = wx * wx * wx * (wx * (wx * 6.0f - 15.0f) + 10.0f);
    const float ay = wy * wy * wy * (wy * (wy * 6.0f - 15.0f) + 10.0f);
    const float az = wz * wz * wz * (wz * (wz * 6.0f - 15.0f) + 10.0f);
    v[0] = (1.0f - ax) * v[0] + ax * v[1];
    v[2] = (1.0f - ax) * v[2] + ax * v[3];
    v[4] = (1.0f - ax) * v[4] + ax * v[5];
    v[6] = (1.0f - ax) * v[6] + ax * v[7];
    v[0] = (1.0f - ay) * v[0] + ay * v[2];
    v[4] = (1.0f - ay) * v[4] + ay * v[6];
    v[0] = (1.0f - az) * v[0] + az * v[4];
    return 0.5f * v[0] + 0.5f;
}

// rotate_between (src/xgen_classic_runtime.cpp:1109-1122)
__device__ F3 rotate_between(F3 value, F3 from, F3 to, float fraction) {
    F3 axis = cross3(from, to);
    const float als = length_squared(axis);
    if (!(als >= 1.0e-12f)) return value;
    axis = axis * (1.0f / sqrtf(als));
    const float angle = acosf(clamp_std(dot3(from, to), -1.0f, 1.0f)) * fraction;
    const float c = cosf(angle), s = sinf(angle);
    return value * c + cross3(axis, value) * s + axis * (dot3(axis, value) * (1.0f - c));
}

// One thread per guide: prepare_xgen_classic_clump_runtime_data's render
// axis and spline length (src/xgen_classic_runtime.cpp:839-875), then the
// guide's clump-noise axis (build_clump_noise_axis, 1041-1186), shared by
// every strand the guide clumps since the module's expressions are constants.
// One guide's render axis (CVs at cv / (n - 1) of its spline), spline
// length and clump-noise axis (build_clump_noise_axis, src/xgen_classic_
// runtime.cpp:1041-1186) scaled by `magnitude` (1: the unit axis, which a
// caller scales per strand in the reference's order); magnitude 0 writes
// the zero axis without building the frame.
// curve_spline_length (src/xgen_classic_runtime.cpp:665-690) of n guide CVs
__device__ float axes_spline_length(const nxg_vec3 *axes, uint32_t n) {
    const Vec3Curve src{axes};
    const uint32_t intervals = n * 2u + 4u;
    const float step = 1.0f / static_cast<float>(intervals);
    F3 previous = xyz(src[0]);
    float result = 0.0f;
    for (uint32_t k = 1u; k < intervals; ++k) {
        const F3 current = curve_eval(src, n, step * static_cast<float>(k));
        result += sqrtf(length_squared(current - previous));
        previous = current;
    }
    return result + sqrtf(length_squared(xyz(src[n - 1u]) - previous));
}

__device__ void clump_guide_axes(const nxg_vec3 *axes, uint32_t n, nxg_vec3 domain_point, nxg_vec3 raw_normal,
                                 nxg_vec3 raw_tangent, float frequency, float correlation, float magnitude,
                                 float4 *render, float4 *noise, float *length) {
    const Vec3Curve src{axes};
    render[0] = src[0];
    for (uint32_t cv = 1u; cv + 1u < n; ++cv) {
        const F3 q = curve_eval(src, n, static_cast<float>(cv) / static_cast<float>(n - 1u));
        render[cv] = make_float4(q.x, q.y, q.z, 0.0f);
    }
    render[n - 1u] = src[n - 1u];
    *length = axes_spline_length(axes, n);
    for (uint32_t cv = 0u; cv < n; ++cv) noise[cv] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (magnitude == 0.0f) return;  // the noise or mask test failed: a zero axis
    float polyline = 0.0f;
    for (uint32_t cv = 1u; cv < n; ++cv) polyline += sqrtf(length_squared(xyz(src[cv]) - xyz(src[cv - 1u])));
    const float effective = polyline > 0.0f ? max_std(0.5f / polyline, frequency) : frequency;
    const float decorrelation = 1.0f - correlation;
    const float domain_scale = 100.0f * decorrelation * decorrelation;
    const F3 domain = (F3{domain_point.x, domain_point.y, domain_point.z} + F3{0.419276f, 0.184247f, 0.805721f}) *
                      domain_scale;
    F3 normal{raw_normal.x, raw_normal.y, raw_normal.z}, tangent{raw_tangent.x, raw_tangent.y, raw_tangent.z};
    normal = !(length_squared(normal) > 1.0e-20f) ? F3{0.0f, 1.0f, 0.0f} : normalize3(normal);
    if (!(length_squared(tangent) > 1.0e-20f)) {  // fallback_surface_u (generate.h:312-316)
        const F3 axis = normal.z > -0.999f && normal.z < 0.999f ? F3{0.0f, 0.0f, 1.0f} : F3{0.0f, 1.0f, 0.0f};
        tangent = normalize3(cross3(normal, axis));
    } else {
        tangent = normalize3(tangent);
    }
    F3 tu = tangent, tv = cross3(normal, tangent), current = normal;
    const F3 first_segment = xyz(render[1]) - xyz(render[0]);
    if (length_squared(first_segment) > 1.0e-20f) current = normalize3(first_segment);
    if (dot3(normal, current) < 0.99999f) {
        tu = rotate_between(tu, normal, current, 1.0f);
        tv = rotate_between(tv, normal, current, 1.0f);
    }
    float travelled = 0.0f;
    for (uint32_t cv = 1u; cv < n; ++cv) {
        travelled += sqrtf(length_squared(xyz(src[cv]) - xyz(src[cv - 1u])));
        F3 su = tu, sv = tv, next = current;
        bool advance = false;
        if (cv + 1u < n) {
            const F3 segment = xyz(render[cv + 1u]) - xyz(render[cv]);
            if (length_squared(segment) > 1.0e-20f) {
                next = normalize3(segment);
                if (fabsf(dot3(current, next)) < 0.99999f) {
                    su = rotate_between(tu, current, next, 0.5f);
                    sv = rotate_between(tv, current, next, 0.5f);
                    advance = true;
                } else {
                    current = next;
                }
            }
        }
        const float distance = travelled * effective;
        const float first = classic_noise(domain.x + distance, domain.y, domain.z) - 0.5f;
        const float second = classic_noise(domain.x, domain.y, domain.z + distance) - 0.5f;
        const F3 d = (su * first + sv * second) * magnitude;
        noise[cv] = make_float4(d.x, d.y, d.z, 0.0f);
        if (advance) {
            tu = rotate_between(tu, current, next, 1.0f);
            tv = rotate_between(tv, current, next, 1.0f);
            current = next;
        }
    }
}

__global__ void __launch_bounds__(64) clump_guide_kernel(ClumpArgs a) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= a.guides) return;
    const uint32_t n = a.cvs;
    clump_guide_axes(a.axes + uint64_t{g} * n, n, a.reference ? a.reference[g] : a.axes[uint64_t{g} * n],
                     a.normals[g], a.tangents[g], a.frequency, a.correlation, a.magnitude,
                     a.render + uint64_t{g} * n, a.noise + uint64_t{g} * n, a.lengths + g);
}

// The search half of cut_to_length_and_rebuild (src/xgen_classic_runtime.cpp:
// 750-799) over the shared cLength samples: the rebuild parameter that keeps
// `target` of spline length measured from the root (1 when the strand is
// shorter), -1 when every CV collapses onto the root.
template <class Curve>
__device__ __forceinline__ float cut_to_length(const Sample *samples_table, const Curve &p, uint32_t n, float target) {
    if (target <= 1.0e-10f) return -1.0f;
    float previous_t = 0.0f, accumulated = 0.0f;
    F3 previous = xyz(p[0]);
    const uint32_t samples = 2u * n + 3u;
#pragma unroll 4
    for (uint32_t k = 0u; k <= samples; ++k) {
        float t;
        F3 current;
        if (k < samples) {
            const Sample smp = load_sample(samples_table, k);
            t = smp.t;
            current = blend(p, n, smp);
        } else {
            t = 1.0f;
            current = xyz(p[n - 1u]);
        }
        const float segment = sqrtf(length_squared(current - previous));
        accumulated += segment;
        if (accumulated >= target && segment > 0.0f) return t - ((accumulated - target) / segment) * (t - previous_t);
        previous_t = t;
        previous = current;
    }
    return 1.0f;
}

// The clump step of one strand (src/xgen_classic_runtime.cpp:1752-1792),
// CVs read from the staged tile, point i handed to out(i, value); false
// leaves the strand alone.  MAXN > 0 unrolls the CV loop (i < MAXN), so a
// caller holding the strand in registers indexes them statically.
template <uint32_t MAXN, class Out>
__device__ __forceinline__ bool clump_strand(const ClumpArgs &a, const StagedCurve<false> &p, uint32_t s, Out &&out,
                                             uint32_t *err) {
    const uint32_t n = a.cvs;
    const uint32_t g = a.strand_guides[s];
    if (g == NXG_INVALID_INDEX || !(a.mask > 0.0f)) return false;
    if (g >= a.guides) {
        *err |= NXG_CLASSIC_INVALID_GUIDE;
        return false;
    }
    float raw = a.amount;
    if (a.amount_multiplier) raw = raw * a.amount_multiplier[uint64_t{s} * a.amount_stride];
    if (!isfinite(raw)) {
        *err |= NXG_CLASSIC_NONFINITE_AMOUNT;
        return false;
    }
    const float amount = clamp_std(raw, 0.0f, 1.0f);
    const float cut = cut_to_length(a.length_samples, p, n, a.lengths[g]);
    const float factor = a.mask * amount * (1.0f - 2.0f * a.clump_scale);
    const float4 *render = a.render + uint64_t{g} * n;
    const float4 *noise = a.noise + uint64_t{g} * n;
    const float step = cut / static_cast<float>(n - 1u);
    auto point = [&](uint32_t i) {
        F3 q = cut < 0.0f ? xyz(p[0]) : curve_eval(p, n, step * static_cast<float>(i));
        q = q + (xyz(render[i]) - q) * factor;
        q = q + xyz(noise[i]);
        return make_float4(q.x, q.y, q.z, p[i].w);
    };
    out(0u, p[0]);
    if constexpr (MAXN > 0u) {
#pragma unroll
        for (uint32_t i = 1u; i < MAXN; ++i)
            if (i < n) out(i, point(i));
    } else {
        for (uint32_t i = 1u; i < n; ++i) out(i, point(i));
    }
    return true;
}

__device__ __forceinline__ void clump_report(const ClumpArgs &a, uint32_t err) {
    const uint32_t bits = (__ballot(err & NXG_CLASSIC_INVALID_GUIDE) ? NXG_CLASSIC_INVALID_GUIDE : 0u) |
                          (__ballot(err & NXG_CLASSIC_NONFINITE_AMOUNT) ? NXG_CLASSIC_NONFINITE_AMOUNT : 0u);
    if (bits && threadIdx.x == 0u && a.status) atomicOr(a.status, bits);
}

// One wavefront per 64-strand tile.  MAXN > 0: each strand's CVs are held
// in registers, then laid out in LDS in output order and stored as one
// contiguous run; MAXN == 0: per-lane stores (the source stays in LDS, so
// the call may run in place).
template <uint32_t MAXN>
__global__ void __launch_bounds__(kWave) clump_strand_kernel(ClumpArgs a) {
    extern __shared__ float4 tile[];
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    {
        CutArgs c{};
        c.points = a.points;
        c.cvs = n;
        c.div_magic = a.div_magic;
        stage_tile<false>(c, tile, first, count);
    }
    const StagedCurve<false> p{tile, lane};
    const uint32_t s = first + lane;
    uint32_t err = 0u;
    if constexpr (MAXN == 0u) {
        if (lane < count) {
            float4 *dst = a.out + uint64_t{s} * n;
            if (!clump_strand<0>(a, p, s, [&](uint32_t i, float4 v) { dst[i] = v; }, &err))
                for (uint32_t i = 0u; i < n; ++i) dst[i] = p[i];
        }
    } else {
        float4 held[MAXN];
        bool done = false;
        if (lane < count) {
            done = clump_strand<MAXN>(a, p, s, [&](uint32_t i, float4 v) { held[i] = v; }, &err);
            if (!done) {
#pragma unroll
                for (uint32_t j = 0u; j < MAXN; ++j)
                    if (j < n) held[j] = p[j];
            }
        }
        __syncthreads();
        const uint32_t stride = n + 1u;
        if (lane < count) {
#pragma unroll
            for (uint32_t j = 0u; j < MAXN; ++j)
                if (j < n) tile[lane * stride + j] = held[j];
        }
        __syncthreads();
        float4 *dst = a.out + uint64_t{first} * n;
        for (uint32_t e = lane; e < count * n; e += kWave) {
            const uint32_t r = __umulhi(e, a.div_magic);
            dst[e] = tile[r * stride + (e - r * n)];
        }
    }
    clump_report(a, err);
}

// ---- Classic primitive expressions ---------------------------------------
// The plan's programs run per strand in the search (length, width, the cut
// amount, taper, taperStart, and widthRamp at every CV for its error
// checks) and widthRamp again per CV in the write, which produces the radii.

// What the search hands the write for a strand.
struct RuntimeState {
    float length_scale;  // the length expression's value (1 without one)
    float c_length;      // cLength after the effects, as the post-effect programs read it
    float c_width;
    float taper, taper_start;
    float reserved[3];
};

constexpr uint32_t kClassicSlots = NXG_CLASSIC_MAX_PROGRAM_SLOTS;

// The per-strand context of the Classic binding (input_value and the float
// context, src/xgen_classic_runtime.cpp:504-574, 1628-1640).
struct ClassicContext {
    float id;
    ExprContext expr;
    const float *row;  // the strand's input row (columns)
};

__device__ __forceinline__ ClassicContext classic_context(const CutArgs &a, uint32_t s) {
    ClassicContext c;
    c.id = static_cast<float>(a.ids ? a.ids[s] : s);
    const nxg_root_sample r = a.roots[s];
    const uint32_t face = r.surface_face_id == NXG_INVALID_INDEX ? r.triangle_index : r.surface_face_id;
    c.expr.u = r.uv.x;
    c.expr.v = r.uv.y;
    // xgen_runtime_face_seed (src/xgen_expression.cpp:724-736)
    c.expr.face_seed =
        hash_to_unit(hash_step(hash_step(hash_step(0u, a.description_id), a.name_seed), static_cast<float>(face)));
    c.expr.t = 0.0f;
    c.expr.random_prefix = a.prefixes ? a.prefixes[s] : 0u;
    c.expr.has_random_prefix = a.prefixes != nullptr;
    c.row = a.columns ? a.columns + uint64_t{s} * a.column_count : nullptr;
    return c;
}

// Program k of the plan for one strand (evaluate_runtime_expression,
// src/xgen_classic_runtime.cpp:555-574): false where the reference throws
// (a non-finite context, input or value).  Input i sits in value slot i; the
// slots are a register array indexed with wave-uniform slot numbers.
struct Evaluated {
    float value;
    bool ok;
};
__device__ Evaluated classic_program(const uint8_t *plan_image, uint32_t program_offset, ClassicContext c,
                                     float c_length, float c_width, float t) {
    const Image image = (Image)plan_image;
    ImageOf(PackedProgram) prog = (ImageOf(PackedProgram))(image + program_offset);
    ImageOf(PackedInstruction) ins = (ImageOf(PackedInstruction))(image + prog->instructions_offset);
    ImageOf(uint32_t) operand_slots = (ImageOf(uint32_t))(image + prog->operands_offset);
    ImageOf(nxg_expr_ramp) ramps = (ImageOf(nxg_expr_ramp))(image + prog->ramps_offset);
    ImageOf(nxg_expr_ramp_point) points = (ImageOf(nxg_expr_ramp_point))(image + prog->ramp_points_offset);
    ExprContext ctx = c.expr;
    ctx.t = t;
    bool bad = !isfinite(ctx.u) || !isfinite(ctx.v) || !isfinite(ctx.face_seed) || !isfinite(t);
    float regs[kClassicSlots];
#define SLOT_GET(x) regs[__builtin_amdgcn_readfirstlane(x) & (kClassicSlots - 1u)]
    const uint32_t inputs = prog->input_count;
    for (uint32_t i = 0; i < inputs; ++i) {
        const uint32_t b = prog->input_bindings[i];
        const float v = b == NXG_CLASSIC_VAR_ID        ? c.id
                        : b == NXG_CLASSIC_VAR_CLENGTH ? c_length
                        : b == NXG_CLASSIC_VAR_CWIDTH  ? c_width
                                                       : c.row[b - NXG_CLASSIC_VAR_COLUMN0];
        if (!isfinite(v)) bad = true;
        SLOT_GET(i) = v;
    }
    const uint32_t n = prog->instruction_count;
    PackedInstruction next = load_instruction(ins);
    for (uint32_t i = 0; i < n; ++i) {
        const PackedInstruction in = next;
        next = load_instruction(ins + i + 1u);
#define OPERAND(e) (((e) & kImmediate) ? (((e) & 1u) ? in.immediate1 : in.immediate) : SLOT_GET(e))
#define EXPR_OPV(j) OPERAND(uint32_t{in.slot[j]})
#define EXPR_OPA(j) OPERAND(operand_slots[in.operand_offset + (j)])
#define EXPR_NOISE3(x, y, z) __builtin_nanf("")  // no extension ops in a Classic plan
#define EXPR_TEXEL(map, f, u, v) __builtin_nanf("")
        float value = 0.0f;
#include "expr_ops.inc"
#undef EXPR_OPV
#undef EXPR_OPA
#undef EXPR_NOISE3
#undef EXPR_TEXEL
#undef OPERAND
        if (!isfinite(value)) bad = true;
        SLOT_GET(uint32_t{in.dest}) = value;
    }
    const float out = SLOT_GET(prog->result);
#undef SLOT_GET
    return {out, !bad};
}

__device__ __forceinline__ bool classic_eval(const CutArgs &a, uint32_t k, const ClassicContext &c, float c_length,
                                             float c_width, float t, float &out) {
    const Evaluated e = classic_program(a.image, a.program_offset[k], c, c_length, c_width, t);
    out = e.value;
    return e.ok;
}

// The parameter of CV cv (src/xgen_classic_runtime.cpp:1840-1841).
__device__ __forceinline__ float cv_parameter(uint32_t cv, uint32_t n) {
    return static_cast<float>(cv) / static_cast<float>(n - 1u);
}

// The programs evaluated per CV (at t = cv / (n - 1)): widthRamp (table 0)
// and each clump module's clumpScale and noiseScale (tables 1 + 2m, 2 + 2m);
// the others run at t = 0.
__device__ __forceinline__ int cv_table(uint32_t k) {
    if (k == NXG_CLASSIC_EXPR_WIDTH_RAMP) return 0;
    if (k < NXG_CLASSIC_EXPR_CLUMP_MASK || k >= NXG_CLASSIC_EXPR_NOISE_MASK) return -1;
    const uint32_t m = (k - NXG_CLASSIC_EXPR_CLUMP_MASK) / NXG_CLASSIC_CLUMP_EXPRS;
    const uint32_t role = k - NXG_CLASSIC_CLUMP_EXPRS * m;
    return role == NXG_CLASSIC_EXPR_CLUMP_SCALE         ? static_cast<int>(1u + 2u * m)
           : role == NXG_CLASSIC_EXPR_CLUMP_NOISE_SCALE ? static_cast<int>(2u + 2u * m)
                                                        : -1;
}
// The program behind CV table `table`.
__device__ __forceinline__ uint32_t cv_table_program(uint32_t table) {
    if (table == 0u) return NXG_CLASSIC_EXPR_WIDTH_RAMP;
    const uint32_t m = (table - 1u) / 2u;
    return NXG_CLASSIC_EXPR_CLUMP_OF(m, (table - 1u) % 2u == 0u ? NXG_CLASSIC_EXPR_CLUMP_SCALE
                                                                : NXG_CLASSIC_EXPR_CLUMP_NOISE_SCALE);
}
// The programs a clump leaves t at 1 for: the reference's context keeps the
// t of the clump's last CV (src/xgen_classic_runtime.cpp:1761-1764), which
// the cut amount, taper and taperStart then read.  Their values at t = 1
// take the table's spare scalar slots.
__device__ __forceinline__ int after_clump_slot(uint32_t k) {
    return k == NXG_CLASSIC_EXPR_CUT_AMOUNT ? 0 : k == NXG_CLASSIC_EXPR_TAPER ? 1
         : k == NXG_CLASSIC_EXPR_TAPER_START ? 2 : -1;
}
static_assert(NXG_CLASSIC_EXPR_COUNT + 3u <= kClassicScalarSlots, "spare scalar slots for t = 1");

// Where a uniform program's value lives in the call's table.
__device__ __forceinline__ uint32_t uniform_slot(uint32_t k, uint32_t cv, bool at_one) {
    const int table = cv_table(k);
    if (table >= 0) return kClassicScalarSlots + static_cast<uint32_t>(table) * NXG_CLASSIC_MAX_CVS + cv;
    return at_one ? NXG_CLASSIC_EXPR_COUNT + static_cast<uint32_t>(after_clump_slot(k)) : k;
}

// Program k for one strand (per-CV programs at CV cv, the others at t = 0,
// or at t = 1 after a clump when at_one): a uniform program's value comes
// from the call's table, after the context check the reference makes at
// every evaluation (xgen_expression.cpp:991-994).
__device__ __forceinline__ bool classic_value(const CutArgs &a, uint32_t k, const ClassicContext &c, float c_length,
                                              float c_width, uint32_t cv, float &out, bool at_one = false) {
    if (k < 32u && ((a.uniform >> k) & 1u)) {
        // the slot is wave-uniform and the table constant during the call: a scalar load
        ImageOf(float) table = (ImageOf(float))a.uniform_values;
        const uint32_t slot = uniform_slot(k, cv, at_one);
        out = table[2u * slot];
        return table[2u * slot + 1u] != 0.0f && isfinite(c.expr.u) && isfinite(c.expr.v) &&
               isfinite(c.expr.face_seed);
    }
    const float t = cv_table(k) >= 0 ? cv_parameter(cv, a.cvs) : (at_one ? 1.0f : 0.0f);
    return classic_eval(a, k, c, c_length, c_width, t, out);
}

// The plan's uniform programs, once per call: entry j < kClassicScalarSlots
// is program j at t = 0 (and the cut amount, taper and taperStart at t = 1
// in the spare slots), then each per-CV program at every CV.
__global__ void __launch_bounds__(256) classic_uniform_kernel(CutArgs a) {
    const uint32_t entries = kClassicScalarSlots + kClassicCvTables * NXG_CLASSIC_MAX_CVS;
    for (uint32_t j = threadIdx.x; j < entries; j += blockDim.x) {
        uint32_t k, cv = 0u;
        float t = 0.0f;
        if (j < NXG_CLASSIC_EXPR_COUNT) {
            k = j;
            if (cv_table(k) >= 0) continue;
        } else if (j < NXG_CLASSIC_EXPR_COUNT + 3u) {
            const uint32_t spare = j - NXG_CLASSIC_EXPR_COUNT;
            k = spare == 0u ? NXG_CLASSIC_EXPR_CUT_AMOUNT : spare == 1u ? NXG_CLASSIC_EXPR_TAPER
                                                                        : NXG_CLASSIC_EXPR_TAPER_START;
            t = 1.0f;
        } else if (j < kClassicScalarSlots) {
            continue;
        } else {
            const uint32_t table = (j - kClassicScalarSlots) / NXG_CLASSIC_MAX_CVS;
            cv = (j - kClassicScalarSlots) % NXG_CLASSIC_MAX_CVS;
            if (cv >= a.cvs) continue;
            k = cv_table_program(table);
            t = cv_parameter(cv, a.cvs);
        }
        if (k >= 32u || !((a.uniform >> k) & 1u)) continue;  // NoiseFX programs are never uniform
        ClassicContext c{};
        float v = 0.0f;
        const bool ok = classic_eval(a, k, c, 0.0f, 0.0f, t, v);
        a.uniform_values[j] = make_float2(v, ok ? 1.0f : 0.0f);
    }
}

// The context module m's noise expressions see (build_clump_noise_axis,
// src/xgen_classic_runtime.cpp:1730-1737, 1057-1059): the strand's, with the
// guide's uv and its SeExpr prefix.  The face seed feeds only a rand()
// without a prefix, which this context never evaluates.
__device__ __forceinline__ ClassicContext guide_context(const ClumpModule &cm, const ClassicContext &strand,
                                                        uint32_t g) {
    ClassicContext c = strand;
    const nxg_vec2 uv = cm.uvs[g];
    c.expr.u = uv.x;
    c.expr.v = uv.y;
    c.expr.face_seed = 0.0f;
    c.expr.random_prefix = cm.prefixes[g];
    c.expr.has_random_prefix = true;
    return c;
}

// Per guide of module m, once per call: render axis, spline length and the
// unit clump-noise axis with the guide's noise frequency and correlation
// (programs that read no input, so one value per guide).
__global__ void __launch_bounds__(64) runtime_guide_kernel(CutArgs a, uint32_t m) {
    const ClumpModule &cm = a.clumps[m];
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= cm.guide_count) return;
    const uint32_t n = a.cvs;
    const uint64_t slot = uint64_t{cm.guide_base} + g;
    const ClassicContext gc = guide_context(cm, ClassicContext{}, g);
    float raw_frequency = 0.0f, raw_correlation = 0.0f;
    const bool ok =
        classic_value(a, NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_NOISE_FREQUENCY), gc, 0.0f, 0.0f, 0u,
                      raw_frequency) &&
        classic_value(a, NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_NOISE_CORRELATION), gc, 0.0f, 0.0f, 0u,
                      raw_correlation);
    const float frequency = max_std(raw_frequency, 0.0f);
    const float correlation = clamp_std(raw_correlation * 0.01f, 0.0f, 1.0f);
    const nxg_vec3 *axes = cm.axes + uint64_t{g} * n;
    const nxg_vec3 *local = cm.local_axes ? cm.local_axes + uint64_t{g} * n : nullptr;
    // with root_relative and local axes the render and noise axes come from
    // the local copy; the noise domain stays the world root or reference
    // position, and the spline length is the local copy's whenever there is
    // one (prepare_xgen_classic_clump_runtime_data, :839-890, 1697-1721)
    float length = 0.0f;
    clump_guide_axes(cm.local_frame ? local : axes, n, cm.reference ? cm.reference[g] : axes[0], cm.normals[g],
                     cm.tangents[g], frequency, correlation, ok ? 1.0f : 0.0f, a.guide_render + slot * n,
                     a.guide_noise + slot * n, &length);
    if (local && !cm.local_frame) length = axes_spline_length(local, n);
    a.guide_info[slot] = make_float2(length, ok ? 1.0f : 0.0f);
}

// scale_curve (src/xgen_classic_runtime.cpp:576-585) as a view: positions
// root + (p - root) * scale when `on` (a length expression ran), radii kept.
template <class Curve>
struct ScaledCurve {
    Curve base;
    F3 root;
    float scale;
    bool on;
    __device__ __forceinline__ float4 operator[](uint32_t c) const {
        const float4 v = base[c];
        if (!on) return v;
        const F3 q = root + (xyz(v) - root) * scale;
        return make_float4(q.x, q.y, q.z, v.w);
    }
};

// The curve cut_from_tip_and_rebuild leaves (src/xgen_classic_runtime.cpp:
// 729-742) as a view: CV i at parameter i cut / (n - 1), every CV at the
// root when the walk removed the whole strand (cut 0).
template <class Curve>
struct RebuiltCurve {
    Curve source;
    uint32_t n;
    float cut;
    __device__ __forceinline__ float4 operator[](uint32_t i) const {
        if (cut <= 0.0f || i == 0u) return source[0];
        const F3 q = curve_eval(source, n, cut / static_cast<float>(n - 1u) * static_cast<float>(i));
        return make_float4(q.x, q.y, q.z, 0.0f);
    }
};

__device__ __forceinline__ bool has_program(const CutArgs &a, uint32_t k) { return (a.present >> k) & 1u; }

// scale_curve applied once to a lane's staged CVs, in place (each lane owns
// its strand's LDS entries); the same operations as ScaledCurve per read,
// CV 0 included (a -0 root coordinate becomes +0 there).  Strands read from
// memory keep the view.
template <bool ROTATED>
__device__ __forceinline__ bool scale_in_place(const StagedCurve<ROTATED> &p, uint32_t n, float scale) {
    const F3 root = xyz(p[0]);
    for (uint32_t c = 0u; c < n; ++c) {
        const float4 v = p[c];
        const F3 q = root + (xyz(v) - root) * scale;
        p.tile[staged_index<ROTATED>(c, p.lane)] = make_float4(q.x, q.y, q.z, v.w);
    }
    return true;
}
__device__ __forceinline__ bool scale_in_place(const MemoryCurve &, uint32_t, float) { return false; }

// A lane's strand in a runtime-chain tile: positions only, CV-major, CV c
// of lane L at float c * kRowStride + 3 L.  A 16-CV tile of 64 strands is
// 12 KB, so 13 tiles share a CU's LDS (16-byte rows would allow 10); no
// effect changes a radius, so radii come from the input points.  The row
// padding puts the 16 CVs one lane stages in 16 different banks, and a
// wave-uniform CV index costs one address add per access (a rotated layout
// would cost three).
constexpr uint32_t kRowStride = 3u * kWave + 1u;
__device__ __forceinline__ uint32_t row_index(uint32_t c, uint32_t lane) { return c * kRowStride + 3u * lane; }
struct StagedRows {
    float *tile;
    uint32_t lane;
    __device__ __forceinline__ float4 operator[](uint32_t c) const {
        const uint32_t i = row_index(c, lane);
        return make_float4(tile[i], tile[i + 1u], tile[i + 2u], 0.0f);
    }
    __device__ __forceinline__ void store(uint32_t c, F3 q) const {
        const uint32_t i = row_index(c, lane);
        tile[i] = q.x;
        tile[i + 1u] = q.y;
        tile[i + 2u] = q.z;
    }
};

// The tile's `count` strands' positions, read coalesced, into LDS rows.
__device__ __forceinline__ void stage_rows(const CutArgs &a, float *tile, uint32_t first, uint32_t count) {
    const float4 *src = a.points + uint64_t{first} * a.cvs;
    const uint32_t total = count * a.cvs;
    for (uint32_t base = threadIdx.x; base < total; base += kStageBatch * kWave) {
        float4 v[kStageBatch];  // no branches, as in stage_tile
#pragma unroll
        for (uint32_t k = 0u; k < kStageBatch; ++k) v[k] = src[min(base + k * kWave, total - 1u)];
#pragma unroll
        for (uint32_t k = 0u; k < kStageBatch; ++k) {
            const uint32_t i = min(base + k * kWave, total - 1u);
            const uint32_t l = __umulhi(i, a.div_magic);
            const uint32_t j = row_index(i - l * a.cvs, l);
            tile[j] = v[k].x;
            tile[j + 1u] = v[k].y;
            tile[j + 2u] = v[k].z;
        }
    }
    __syncthreads();
}

__device__ __forceinline__ bool scale_in_place(const StagedRows &p, uint32_t n, float scale) {
    const F3 root = xyz(p[0]);
    for (uint32_t c = 0u; c < n; ++c) p.store(c, root + (xyz(p[c]) - root) * scale);
    return true;
}

// The taper part of the width profile at t (src/xgen_classic_runtime.cpp:
// 1842-1848).
__device__ __forceinline__ float taper_scale(float t, float taper, float taper_start) {
    float scale = 1.0f;
    if (t > taper_start && taper_start < 1.0f) scale *= 1.0f - taper * ((t - taper_start) / (1.0f - taper_start));
    return scale;
}

// ClumpingFX module m on one strand (src/xgen_classic_runtime.cpp:1680-1793)
// over its staged CVs, rewritten in the lane's LDS row (the later effects
// read it there; the search copies every row to a.clumped at the end);
// c_length follows the reference's context.  at_one tracks the
// context's t: a module with a valid guide sets it to 0 for its mask and,
// when it clumps, leaves the t of its last CV, 1.  False (with *err) where
// the reference throws.
template <class Staged>
__device__ __forceinline__ bool runtime_clump(const CutArgs &a, uint32_t m, const ClassicContext &c,
                                              const Staged &p, uint32_t s, float &c_length,
                                              float c_width, bool &at_one, uint32_t *err) {
    const uint32_t n = a.cvs;
    const ClumpModule &cm = a.clumps[m];
    const uint32_t g = cm.strand_guides[s];
    if (g == NXG_INVALID_INDEX) return true;
    if (g >= cm.guide_count) {
        *err = NXG_CLASSIC_INVALID_GUIDE;
        return false;
    }
    at_one = false;
    float raw_mask;
    if (!classic_value(a, NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_MASK), c, c_length, c_width, 0u,
                       raw_mask)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return false;
    }
    const float mask = clamp_std(raw_mask, 0.0f, 1.0f);
    if (!(mask > 0.0f)) return true;  // the strand passes through
    // the noise axis's parameters, in the guide's context (:1041-1186)
    const ClassicContext gc = guide_context(cm, c, g);
    const uint32_t noise_scale = NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_NOISE_SCALE);
    float raw_noise;
    if (!classic_value(a, NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_NOISE), gc, c_length, c_width, 0u,
                       raw_noise)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return false;
    }
    const float noise = max_std(raw_noise, 0.0f);
    const bool noisy = noise > 1.0e-5f && mask > 1.0e-4f;
    const uint64_t slot = uint64_t{cm.guide_base} + g;
    const float2 info = a.guide_info[slot];
    if (noisy) {
        if (info.y == 0.0f) {  // the guide's noise frequency or correlation failed
            *err = NXG_CLASSIC_EXPRESSION_ERROR;
            return false;
        }
        for (uint32_t cv = 1u; cv < n; ++cv) {
            float raw_scale;
            if (!classic_value(a, noise_scale, gc, c_length, c_width, cv, raw_scale)) {
                *err = NXG_CLASSIC_EXPRESSION_ERROR;
                return false;
            }
            if (!isfinite(mask * noise * max_std(raw_scale, 0.0f))) {
                *err = NXG_CLASSIC_NONFINITE_CLUMP;
                return false;
            }
        }
    }
    // cut_to_length_and_rebuild to the guide's length, then the amount
    const float cut = cut_to_length(a.length_samples, p, n, info.x);
    const float noise_length = c_length;  // the noise expressions' cLength
    if ((a.clump_length >> m) & 1u)
        c_length = spline_length(a.length_samples, RebuiltCurve<Staged>{p, n, cut}, n);
    float raw_amount;
    if (!classic_value(a, NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP), c, c_length, c_width, 0u,
                       raw_amount)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return false;
    }
    const float amount = clamp_std(raw_amount, 0.0f, 1.0f);
    const float4 *render = a.guide_render + slot * n;
    const float4 *unit = a.guide_noise + slot * n;
    // the goal's offset (:1627-1628, 1773-1776): render - origin, or for a
    // local guide render + (guide root - origin); origin is the strand's
    // root with root_relative, else 0 (render - 0 is render, -0 included)
    F3 shift{0.0f, 0.0f, 0.0f};
    const bool offset = a.root_relative || cm.local_frame;
    if (offset) {
        const nxg_vec3 r = a.roots[s].position;
        const F3 origin = a.root_relative ? F3{r.x, r.y, r.z} : F3{0.0f, 0.0f, 0.0f};
        if (cm.local_frame) {
            const nxg_vec3 gr = cm.axes[uint64_t{g} * n];
            shift = F3{gr.x, gr.y, gr.z} - origin;
        } else {
            shift = origin;
        }
    }
    const float step = cut / static_cast<float>(n - 1u);
    // the clumped CVs go back into the lane's row from the tip down: output
    // cv reads input CVs up to cv + 2 (its spline position cut * cv is at
    // most cv), so slot cv + 2 is free once output cv is computed; the two
    // outputs above wait in registers.  CV 0 stays.
    F3 pending1 = xyz(p[0]), pending2 = pending1;  // outputs cv + 1 and cv + 2
    for (uint32_t cv = n - 1u; cv >= 1u; --cv) {
        F3 q = cut < 0.0f ? xyz(p[0]) : curve_eval(p, n, step * static_cast<float>(cv));
        float scale;
        if (!classic_value(a, NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_SCALE), c, c_length, c_width, cv,
                           scale)) {
            *err = NXG_CLASSIC_EXPRESSION_ERROR;
            return false;
        }
        F3 goal = xyz(render[cv]);
        if (offset) goal = cm.local_frame ? goal + shift : goal - shift;
        q = q + (goal - q) * (mask * amount * (1.0f - 2.0f * scale));
        // the noise axis is added to every CV, zero when it is off (so -0 becomes +0)
        F3 d{0.0f, 0.0f, 0.0f};
        if (noisy) {
            float raw_scale = 0.0f;
            classic_value(a, noise_scale, gc, noise_length, c_width, cv, raw_scale);
            d = xyz(unit[cv]) * (mask * noise * max_std(raw_scale, 0.0f));
        }
        q = q + d;
        if (cv + 2u < n) p.store(cv + 2u, pending2);
        pending2 = pending1;
        pending1 = q;
    }
    if (n > 2u) p.store(2u, pending2);
    p.store(1u, pending1);
    c_length = spline_length(a.length_samples, p, n);
    at_one = true;
    return true;
}

// fallback_surface_u (generate.h:312-316)
__device__ __forceinline__ F3 fallback_surface_u(F3 normal) {
    const F3 axis = normal.z > -0.999f && normal.z < 0.999f ? F3{0.0f, 0.0f, 1.0f} : F3{0.0f, 1.0f, 0.0f};
    return normalize3(cross3(normal, axis));
}

// The NoiseFX step (apply_noise, src/xgen_classic_runtime.cpp:897-1039) on
// one strand's staged CVs: mask, frequency, correlation and preserveLength
// at the context's t; the frame transported from the surface normal along
// the segments (the rotation's cosine and sine from the clamped dot product); per CV
// the magnitude at its t and the classic noise of the domain moved along
// each axis; the preserve-length rescale.  The result is rewritten in the
// lane's LDS row; c_length follows.  False (with *err) where the reference
// throws.
template <class Staged>
__device__ __forceinline__ bool runtime_noise(const CutArgs &a, uint32_t module, const ClassicContext &c,
                                              const Staged &p, uint32_t s, float &c_length, float c_width,
                                              bool &at_one, uint32_t *err) {
    const uint32_t n = a.cvs;
#define NOISE_PROGRAM(k) NXG_CLASSIC_EXPR_NOISE_OF(module, NXG_CLASSIC_EXPR_NOISE_##k)
    const float t0 = at_one ? 1.0f : 0.0f;
    float raw_mask, raw_frequency, raw_correlation, raw_preserve;
    if (!classic_eval(a, NOISE_PROGRAM(MASK), c, c_length, c_width, t0, raw_mask)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return false;
    }
    const float mask = clamp_std(raw_mask, 0.0f, 1.0f);
    if (!(mask > 1.0e-6f)) {
        // the curve is unchanged, but cLength is re-measured after the module
        // whatever its mask (:1805), so a length expression's scaled value
        // gives way to the spline length of the scaled curve
        c_length = spline_length(a.length_samples, p, n);
        return true;
    }
    if (!classic_eval(a, NOISE_PROGRAM(FREQUENCY), c, c_length, c_width, t0, raw_frequency) ||
        !classic_eval(a, NOISE_PROGRAM(CORRELATION), c, c_length, c_width, t0, raw_correlation) ||
        !classic_eval(a, NOISE_PROGRAM(PRESERVE_LENGTH), c, c_length, c_width, t0, raw_preserve)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return false;
    }
    const float frequency = max_std(raw_frequency, 0.0f);
    const float correlation = clamp_std(raw_correlation * 0.01f, 0.0f, 1.0f);
    const float preserve = clamp_std(raw_preserve * 0.01f, 0.0f, 1.0f);
    if (!isfinite(frequency) || !isfinite(correlation) || !isfinite(preserve)) {
        *err = NXG_CLASSIC_NONFINITE_NOISE_PARAMETER;
        return false;
    }
    float original = 0.0f;  // curve_polyline_length
    for (uint32_t cv = 1u; cv < n; ++cv) original += sqrtf(length_squared(xyz(p[cv]) - xyz(p[cv - 1u])));
    const float effective = original > 0.0f ? max_std(0.5f / original, frequency) : frequency;
    const float decorrelation = 1.0f - correlation;
    const float domain_scale = 100.0f * decorrelation * decorrelation;
    const F3 root = xyz(p[0]);
    F3 domain = root;
    if (a.noise_domain) domain = F3{a.noise_domain[s].x, a.noise_domain[s].y, a.noise_domain[s].z};
    domain = (domain + F3{0.419276f, 0.184247f, 0.805721f}) * domain_scale;
    const nxg_vec3 rn = a.roots[s].normal, rt = a.surface_tangents[s];
    const F3 surface_normal{rn.x, rn.y, rn.z};
    F3 surface_tangent{rt.x, rt.y, rt.z};
    surface_tangent = !(length_squared(surface_tangent) >= 1.0e-12f) ? fallback_surface_u(surface_normal)
                                                                     : normalize3(surface_tangent);
    F3 previous_base = root, current_base = root, next_base = xyz(p[1]);
    F3 prior = length_squared(surface_normal) >= 1.0e-12f ? normalize3(surface_normal) : F3{0.0f, 1.0f, 0.0f};
    F3 transported = surface_tangent;
    float travelled = 0.0f;
    // each output goes into its own slot at once: the loop holds input CVs
    // cv and cv + 1 in registers and reads cv + 2 from the row
    for (uint32_t cv = 0u; cv < n; ++cv) {
        if (cv > 0u) travelled += sqrtf(length_squared(current_base - previous_base));
        F3 next = prior;
        if (cv + 1u < n) {
            const F3 segment = next_base - current_base;
            if (length_squared(segment) >= 1.0e-12f) next = normalize3(segment);
        }
        const F3 axis = cross3(prior, next);
        const float als = length_squared(axis);
        const float tangent_dot = clamp_std(dot3(prior, next), -1.0f, 1.0f);
        if (als >= 1.0e-12f) {
            const F3 ra = axis * (1.0f / sqrtf(als));
            // cos(acos c) = c and sin(acos c) = sqrt((1 - c)(1 + c)) for c in
            // [-1, 1]: the reference's acos, cos and sin of the same angle
            // (src/xgen_classic_runtime.cpp:968-970) to a few ulps, without
            // three library calls per CV
            const float cosine = tangent_dot;
            const float sine = sqrtf((1.0f - tangent_dot) * (1.0f + tangent_dot));
            const F3 rotated = transported * cosine + cross3(ra, transported) * sine +
                               ra * (dot3(ra, transported) * (1.0f - cosine));
            transported = length_squared(rotated) >= 1.0e-12f ? normalize3(rotated) : F3{0.0f, 0.0f, 0.0f};
        }
        const F3 normal = transported;
        F3 binormal = cross3(normal, next);
        binormal = length_squared(binormal) >= 1.0e-12f ? normalize3(binormal) : F3{0.0f, 0.0f, 0.0f};
        const F3 tangent = cross3(binormal, normal);
        const float t = cv_parameter(cv, n);
        float raw_magnitude, raw_scale;
        if (!classic_eval(a, NOISE_PROGRAM(MAGNITUDE), c, c_length, c_width, t, raw_magnitude) ||
            !classic_eval(a, NOISE_PROGRAM(MAGNITUDE_SCALE), c, c_length, c_width, t, raw_scale)) {
            *err = NXG_CLASSIC_EXPRESSION_ERROR;
            return false;
        }
        const float magnitude = max_std(raw_magnitude, 0.0f) * max_std(raw_scale, 0.0f) * mask;
        if (!isfinite(magnitude)) {
            *err = NXG_CLASSIC_NONFINITE_NOISE;
            return false;
        }
        if (cv > 0u) {
            const float distance = travelled * effective;
            const F3 local{(classic_noise(domain.x + distance, domain.y, domain.z) - 0.5f) * magnitude,
                           (classic_noise(domain.x, domain.y + distance, domain.z) - 0.5f) * magnitude,
                           (classic_noise(domain.x, domain.y, domain.z + distance) - 0.5f) * magnitude};
            p.store(cv, current_base + normal * local.x + binormal * local.y + tangent * local.z);
        }
        prior = next;
        previous_base = current_base;
        if (cv + 1u < n) {
            current_base = next_base;
            if (cv + 2u < n) next_base = xyz(p[cv + 2u]);
        }
    }
    if (preserve > 0.001f) {
        float noisy = 0.0f;
        for (uint32_t cv = 1u; cv < n; ++cv) noisy += sqrtf(length_squared(xyz(p[cv]) - xyz(p[cv - 1u])));
        if (noisy > 0.0f) {
            const float target = original * preserve + noisy * (1.0f - preserve);
            if (fabsf(noisy - target) >= 0.0001f) {  // scale_curve (:576-585)
                const float scale = target / noisy;
                const F3 r0 = xyz(p[0]);
                for (uint32_t cv = 0u; cv < n; ++cv) p.store(cv, r0 + (xyz(p[cv]) - r0) * scale);
            }
        }
    }
    c_length = spline_length(a.length_samples, p, n);
    at_one = true;
    return true;
#undef NOISE_PROGRAM
}

// The rebuild half of cut_from_tip_and_rebuild (src/xgen_classic_runtime.cpp:
// 729-742) for a CutFX that later modules follow: n CVs uniformly on
// [0, cut] (all at the root when cut is 0), radii kept, in the lane's LDS
// row.
template <class Staged>
__device__ __forceinline__ void rebuild_staged(const CutArgs &a, const Staged &p, float cut) {
    const uint32_t n = a.cvs;
    const float step = cut / static_cast<float>(n - 1u);
    const F3 root = xyz(p[0]);
    // from the tip down with two outputs waiting, as the clump rebuilds (the
    // cut parameter is at most 1)
    F3 pending1 = root, pending2 = root;
    for (uint32_t cv = n - 1u; cv >= 1u; --cv) {
        const F3 q = cut > 0.0f ? curve_eval(p, n, step * static_cast<float>(cv)) : root;
        if (cv + 2u < n) p.store(cv + 2u, pending2);
        pending2 = pending1;
        pending1 = q;
    }
    if (n > 2u) p.store(2u, pending2);
    p.store(1u, pending1);
}

// One strand of the runtime chain (src/xgen_classic_runtime.cpp:1619-1858):
// the cut parameter of a kept strand (1 without a CutFX), -1 for a culled
// one; *err gets the NXG_CLASSIC_* bit where the reference throws.
// NOISE compiles the NoiseFX step in: its frame state raises the kernel's
// register use, so plans without a NoiseFX run a variant without it.
template <bool NOISE, class Curve>
__device__ __forceinline__ float runtime_strand(const CutArgs &a, const Curve &source, uint32_t s, uint32_t *err) {
    const uint32_t n = a.cvs;
    if (a.counts && a.counts[s] != n) {
        *err = NXG_CLASSIC_VARIABLE_COUNTS;
        return -1.0f;
    }
    const ClassicContext c = classic_context(a, s);
    // cLength of the input curve: with defer_length it is measured only if
    // the first ClumpingFX leaves this strand alone (no guide, mask 0), since
    // a clump re-measures it and nothing reads it before; the culling test
    // then uses the chord root-tip, a lower bound of the sampled spline's
    // polyline, with a margin well above the sum's rounding
    const float chord = sqrtf(length_squared(xyz(source[n - 1u]) - xyz(source[0])));
    const bool deferred = a.defer_length && chord > 1.001e-4f;
    float c_length = deferred ? 0.0f : spline_length(a.length_samples, source, n);
    float length_scale = 1.0f;
    if (has_program(a, NXG_CLASSIC_EXPR_LENGTH)) {
        if (!classic_value(a, NXG_CLASSIC_EXPR_LENGTH, c, c_length, 0.0f, 0u, length_scale)) {
            *err = NXG_CLASSIC_EXPRESSION_ERROR;
            return -1.0f;
        }
        if (!isfinite(length_scale) || length_scale < 0.0f) {
            *err = NXG_CLASSIC_NEGATIVE_LENGTH;
            return -1.0f;
        }
        c_length *= length_scale;
    }
    const bool certain = deferred && chord * length_scale > 1.001e-4f;  // cLength >= chord * scale > 1e-4
    if (deferred && !certain) c_length = spline_length(a.length_samples, source, n) * length_scale;
    const bool pending = deferred && certain;  // cLength not measured yet
    const bool scaled = has_program(a, NXG_CLASSIC_EXPR_LENGTH) && !scale_in_place(source, n, length_scale);
    const ScaledCurve<Curve> p{source, xyz(source[0]), length_scale, scaled};
    float width;
    if (has_program(a, NXG_CLASSIC_EXPR_WIDTH)) {
        if (!classic_value(a, NXG_CLASSIC_EXPR_WIDTH, c, c_length, 0.0f, 0u, width)) {
            *err = NXG_CLASSIC_EXPRESSION_ERROR;
            return -1.0f;
        }
    } else {
        width = a.radius_scale > 0.0f ? 2.0f * a.points[uint64_t{s} * n].w / a.radius_scale : 0.0f;
    }
    if (!isfinite(width)) {
        *err = NXG_CLASSIC_NONFINITE_WIDTH;
        return -1.0f;
    }
    if ((!pending && c_length < 1.0e-4f) || width < 1.0e-4f) return -1.0f;
    const float c_width = width;
    bool clumped = false;  // the context's t is 1: the last effect that set it ran its CV loop
    constexpr bool kStaged = std::is_same<Curve, StagedRows>::value;
    // the effects in list order (:1679-1826); effects that change the CVs
    // run on staged strands (at most 64 CVs), a trailing CutFX also on
    // strands read from memory (the write resamples them)
    float cut = 1.0f;
    bool keep = true;
#pragma unroll 1
    for (uint32_t e = 0; e < a.effect_count; ++e) {
        const uint32_t type = a.effects[e] >> 8u, module = a.effects[e] & 0xffu;
        if (type == NXG_CLASSIC_EFFECT_CUT) {
            float amount;
            if (!classic_value(a, NXG_CLASSIC_EXPR_CUT_AMOUNT_OF(module), c, c_length, c_width, 0u, amount,
                               clumped)) {
                *err = NXG_CLASSIC_EXPRESSION_ERROR;
                return -1.0f;
            }
            if (!isfinite(amount)) {
                *err = NXG_CLASSIC_NONFINITE_AMOUNT;
                return -1.0f;
            }
            cut = cut_parameter(a.walk_samples, a.walk_count, p, n, amount < 0.0f ? 0.0f : amount);
            if (cut < 1.0e-4f) keep = false;  // culled, but the later effects still run
            if (e + 1u < a.effect_count) {
                if constexpr (kStaged) {  // later effects read the cut strand: rebuild it now
                    rebuild_staged(a, source, cut);
                    cut = 1.0f;  // nothing left for the write to resample
                    c_length = spline_length(a.length_samples, source, n);
                }
            } else if (a.final_length) {  // the post-effect programs read the rebuilt curve's cLength
                c_length = spline_length(a.length_samples, RebuiltCurve<ScaledCurve<Curve>>{p, n, cut}, n);
            }
        } else if (type == NXG_CLASSIC_EFFECT_NOISE) {
            if constexpr (kStaged && NOISE) {
                if (!runtime_noise(a, module, c, source, s, c_length, c_width, clumped, err)) return -1.0f;
            }
        } else {
            if constexpr (kStaged) {
                if (!runtime_clump(a, module, c, source, s, c_length, c_width, clumped, err))
                    return -1.0f;
            }
        }
        if (e == 0u && pending && !clumped) {
            // the clump left the strand as it was: cLength of the input (the
            // tile may hold it scaled, the memory copy does not), then scaled
            c_length = spline_length(a.length_samples, MemoryCurve{a.points + uint64_t{s} * n}, n);
            if (has_program(a, NXG_CLASSIC_EXPR_LENGTH)) c_length *= length_scale;
        }
    }
    float taper = 0.0f, taper_start = 0.0f;
    if (has_program(a, NXG_CLASSIC_EXPR_TAPER) &&
        !classic_value(a, NXG_CLASSIC_EXPR_TAPER, c, c_length, c_width, 0u, taper, clumped)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return -1.0f;
    }
    if (has_program(a, NXG_CLASSIC_EXPR_TAPER_START) &&
        !classic_value(a, NXG_CLASSIC_EXPR_TAPER_START, c, c_length, c_width, 0u, taper_start, clumped)) {
        *err = NXG_CLASSIC_EXPRESSION_ERROR;
        return -1.0f;
    }
    if (!isfinite(taper) || !isfinite(taper_start)) {
        *err = NXG_CLASSIC_NONFINITE_TAPER;
        return -1.0f;
    }
    if (a.profile) {  // the width profile's errors (its values are the write's)
        for (uint32_t cv = 0u; cv < n; ++cv) {
            float scale = taper_scale(cv_parameter(cv, n), taper, taper_start);
            if (has_program(a, NXG_CLASSIC_EXPR_WIDTH_RAMP)) {
                float ramp;
                if (!classic_value(a, NXG_CLASSIC_EXPR_WIDTH_RAMP, c, c_length, c_width, cv, ramp)) {
                    *err = NXG_CLASSIC_EXPRESSION_ERROR;
                    return -1.0f;
                }
                scale *= ramp;
            }
            if (!isfinite(c_width * scale)) {
                *err = NXG_CLASSIC_NONFINITE_PROFILE;
                return -1.0f;
            }
        }
    }
    RuntimeState st{};
    st.length_scale = length_scale;
    st.c_length = c_length;
    st.c_width = c_width;
    st.taper = taper;
    st.taper_start = taper_start;
    a.state[s] = st;
    return keep ? cut : -1.0f;
}

// With a NoiseFX the kernel is held to 168 VGPRs (3 wavefronts per SIMD,
// which the 12 KB tiles allow) at the cost of a small register spill:
// 2.82 ms against 3.27 ms at 195 VGPRs and 2 wavefronts on the classic_noise
// workload.
template <bool STAGED, bool NOISE>
__global__ void __launch_bounds__(kWave, NOISE ? 3 : 1) runtime_search_kernel(CutArgs a) {
    extern __shared__ float4 tile[];
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    float *rows = reinterpret_cast<float *>(tile);
    if constexpr (STAGED) stage_rows(a, rows, first, count);
    const uint32_t s = first + lane;
    float cut = -1.0f;
    uint32_t err = 0u;
    if (lane < count) {
        if constexpr (STAGED)
            cut = runtime_strand<NOISE>(a, StagedRows{rows, lane}, s, &err);
        else
            cut = runtime_strand<false>(a, MemoryCurve{a.points + uint64_t{s} * n}, s, &err);
        a.cut[s] = cut;
    }
    if constexpr (STAGED) {
        if (a.rewrites) {  // every row as the effects left it, for the write: one contiguous run per tile
            __syncthreads();
            float4 *dst = a.clumped + uint64_t{first} * n;
            const float4 *src = a.points + uint64_t{first} * n;  // the radii
            const uint32_t total = count * n;
            for (uint32_t base = lane; base < total; base += kStageBatch * kWave) {
                float radius[kStageBatch];
#pragma unroll
                for (uint32_t k = 0u; k < kStageBatch; ++k) radius[k] = src[min(base + k * kWave, total - 1u)].w;
#pragma unroll
                for (uint32_t k = 0u; k < kStageBatch; ++k) {
                    const uint32_t e = min(base + k * kWave, total - 1u);  // as in stage_tile
                    const uint32_t r = __umulhi(e, a.div_magic);
                    const uint32_t i = row_index(e - r * n, r);
                    dst[e] = make_float4(rows[i], rows[i + 1u], rows[i + 2u], radius[k]);
                }
            }
        }
    }
    const uint64_t kept = __ballot(cut >= 0.0f);
    uint32_t bits = 0u;  // the wave's NXG_CLASSIC_* bits (all below 1 << 16)
#pragma unroll
    for (uint32_t b = 0u; b < 16u; ++b)
        if (__ballot((err >> b) & 1u)) bits |= 1u << b;
    if (lane == 0u) {
        a.tiles[t] = static_cast<uint32_t>(__popcll(kept)) | (bits ? 1u << kErrShift : 0u);
        if (bits) atomicOr(a.err_bits, bits);
    }
}

// The kept strands at their compacted slots: the scaled (and, with a CutFX,
// rebuilt) positions, and radii from the width profile when the plan has
// one (src/xgen_classic_runtime.cpp:1836-1858), else the input radii.
template <bool STAGED>
__global__ void __launch_bounds__(kWave) runtime_write_kernel(CutArgs a) {
    extern __shared__ float4 tile[];
    if (*a.status_word != 0u) return;  // the reference throws: nothing is written
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    const uint32_t s = first + lane;
    const float cut = lane < count ? a.cut[s] : -1.0f;
    const uint64_t kept = __ballot(cut >= 0.0f);
    if (kept == 0u) return;  // the block is this one wavefront
    if constexpr (STAGED) stage_tile<false>(a, tile, first, count);
    if (!(cut >= 0.0f)) return;
    const uint32_t rank = static_cast<uint32_t>(__popcll(kept & ((uint64_t{1} << lane) - 1u)));
    const uint32_t slot = a.groups[t / kScanBlock] + a.tiles[t] + rank;
    const RuntimeState st = a.state[s];
    const ClassicContext c = classic_context(a, s);
    float4 *dst = a.out_points + uint64_t{slot} * n;
    auto write = [&](const auto &source) {
        using Curve = typename std::decay<decltype(source)>::type;
        const ScaledCurve<Curve> p{source, xyz(source[0]), st.length_scale, has_program(a, NXG_CLASSIC_EXPR_LENGTH)};
        const float step = cut / static_cast<float>(n - 1u);
        for (uint32_t i = 0u; i < n; ++i) {
            float4 q = p[i];
            if (a.has_cut && i > 0u) {
                const F3 r = curve_eval(p, n, step * static_cast<float>(i));
                q = make_float4(r.x, r.y, r.z, q.w);
            }
            if (a.profile) {
                float scale = taper_scale(cv_parameter(i, n), st.taper, st.taper_start);
                if (has_program(a, NXG_CLASSIC_EXPR_WIDTH_RAMP)) {
                    float ramp = 0.0f;
                    classic_value(a, NXG_CLASSIC_EXPR_WIDTH_RAMP, c, st.c_length, st.c_width, i, ramp);
                    scale *= ramp;
                }
                q.w = 0.5f * max_std(st.c_width * scale, 0.0f) * a.radius_scale;
            }
            dst[i] = q;
        }
    };
    if constexpr (STAGED)
        write(StagedCurve<false>{tile, lane});
    else
        write(MemoryCurve{a.points + uint64_t{s} * n});
    if (a.out_counts) a.out_counts[slot] = n;
    if (a.out_roots) a.out_roots[slot] = a.roots[s];
    if (a.out_uvs) a.out_uvs[slot] = a.uvs[s];
}

// The write with each strand held in registers (n <= MAXN), laid out in
// output order through the tile's LDS and stored as one contiguous run per
// wave, as cut_write_held_kernel does.
// CVs per output pass of the held write: a pass compacts every kept
// strand's next kHeldPass CVs in LDS (rows padded by one float4) and stores
// them as runs of kHeldPass * 16 bytes.
constexpr uint32_t kHeldPass = 8u;

template <uint32_t MAXN>
__global__ void __launch_bounds__(kWave) runtime_write_held_kernel(CutArgs a) {
    extern __shared__ float4 tile[];
    if (*a.status_word != 0u) return;  // the reference throws: nothing is written
    const uint32_t n = a.cvs;
    const uint32_t t = blockIdx.x, lane = threadIdx.x;
    const uint32_t first = t * kWave;
    const uint32_t count = a.strands - first < kWave ? a.strands - first : kWave;
    const uint32_t s = first + lane;
    const float cut = lane < count ? a.cut[s] : -1.0f;
    const bool keep = cut >= 0.0f;
    const uint64_t kept = __ballot(keep);
    if (kept == 0u) return;  // the block is this one wavefront
    // positions-only rows (the tile's 12 KB at 16 CVs let 13 tiles share a
    // CU, where whole CVs with a whole compaction area would allow 9)
    float *rows = reinterpret_cast<float *>(tile);
    stage_rows(a, rows, first, count);
    const uint32_t rank = static_cast<uint32_t>(__popcll(kept & ((uint64_t{1} << lane) - 1u)));
    const uint32_t slot0 = a.groups[t / kScanBlock] + a.tiles[t];
    F3 held[MAXN];
    RuntimeState st{};
    if (keep) {  // positions
        st = a.state[s];
        const StagedRows p{rows, lane};
        if (has_program(a, NXG_CLASSIC_EXPR_LENGTH)) scale_in_place(p, n, st.length_scale);
        const float step = cut / static_cast<float>(n - 1u);
#pragma unroll
        for (uint32_t i = 0u; i < MAXN; ++i) {
            if (i >= n) continue;
            held[i] = a.has_cut && i > 0u ? curve_eval(p, n, step * static_cast<float>(i)) : xyz(p[i]);
        }
    }
    // the rows are free after this barrier: the kept lanes' ids, then the
    // compaction area of one pass
    __syncthreads();
    uint32_t *lane_of = reinterpret_cast<uint32_t *>(rows);
    float4 *pass = tile + kWave / 4u;
    constexpr uint32_t kPitch = kHeldPass + 1u;
    if (keep) lane_of[rank] = lane;
    const ClassicContext c = keep && a.profile ? classic_context(a, s) : ClassicContext{};
    const uint32_t kept_count = static_cast<uint32_t>(__popcll(kept));
    float4 *dst = a.out_points + uint64_t{slot0} * n;
    const float4 *src = a.points + uint64_t{first} * n;  // the input radii (no width profile)
    for (uint32_t c0 = 0u; c0 < n; c0 += kHeldPass) {
        const uint32_t h = n - c0 < kHeldPass ? n - c0 : kHeldPass;
        if (c0) __syncthreads();  // the previous pass's reads are done
        if (keep) {
#pragma unroll
            for (uint32_t i = 0u; i < MAXN; ++i)
                if (i >= c0 && i < c0 + h) pass[rank * kPitch + (i - c0)] = make_float4(held[i].x, held[i].y,
                                                                                        held[i].z, 0.0f);
            if (a.profile) {  // the width profile's radii (src/xgen_classic_runtime.cpp:1836-1858)
                for (uint32_t i = c0; i < c0 + h; ++i) {
                    float scale = taper_scale(cv_parameter(i, n), st.taper, st.taper_start);
                    if (has_program(a, NXG_CLASSIC_EXPR_WIDTH_RAMP)) {
                        float ramp = 0.0f;
                        classic_value(a, NXG_CLASSIC_EXPR_WIDTH_RAMP, c, st.c_length, st.c_width, i, ramp);
                        scale *= ramp;
                    }
                    pass[rank * kPitch + (i - c0)].w = 0.5f * max_std(st.c_width * scale, 0.0f) * a.radius_scale;
                }
            }
        }
        __syncthreads();
        // element e of the pass: strand e / h (f / h for f < 64 * 8 as
        // floor((f + 0.5) / h) in float, exact as in flush_chunk), CV c0 + e % h;
        // branch-free with the input radii's loads ahead of the stores
        const uint32_t total = kept_count * h;
        const float inv_h = 1.0f / static_cast<float>(h);
        for (uint32_t base = lane; base < total; base += kStageBatch * kWave) {
            float radius[kStageBatch];
            if (!a.profile) {
#pragma unroll
                for (uint32_t k = 0u; k < kStageBatch; ++k) {
                    const uint32_t e = min(base + k * kWave, total - 1u);
                    const uint32_t r = static_cast<uint32_t>((static_cast<float>(e) + 0.5f) * inv_h);
                    radius[k] = src[lane_of[r] * n + c0 + (e - r * h)].w;
                }
            }
#pragma unroll
            for (uint32_t k = 0u; k < kStageBatch; ++k) {
                const uint32_t e = min(base + k * kWave, total - 1u);  // past the end: the last element again
                const uint32_t r = static_cast<uint32_t>((static_cast<float>(e) + 0.5f) * inv_h);
                const uint32_t j = e - r * h;
                float4 v = pass[r * kPitch + j];
                if (!a.profile) v.w = radius[k];
                dst[r * n + c0 + j] = v;
            }
        }
    }
    if (!keep) return;
    const uint32_t slot = slot0 + rank;
    if (a.out_counts) a.out_counts[slot] = n;
    if (a.out_roots) a.out_roots[slot] = a.roots[s];
    if (a.out_uvs) a.out_uvs[slot] = a.uvs[s];
}

}  // namespace
}  // namespace nxg

namespace nxg {
namespace {

// The cut walk's parameter sequence (1 - step, repeatedly, clamped at 0) up
// to its first value <= 1e-10, counted in the same float steps.
uint32_t walk_count(uint32_t cvs) {
    const float walk_step = 1.0f / (2.0f * static_cast<float>(cvs) + 4.0f);
    float previous = 1.0f;
    uint32_t count = 0u;
    while (count < kClassicTableEntries) {
        const float lowered = previous - walk_step;
        previous = lowered < 0.0f ? 0.0f : lowered;
        ++count;
        if (previous <= 1.0e-10f) break;
    }
    return count;
}

// The arguments every chain shares: inputs, outputs and workspace regions.
CutArgs chain_args(const nxg_classic_curves *curves, const nxg_classic_output *output, uint8_t *ws,
                   const ClassicLayout &layout) {
    const uint32_t n = curves->strand_count, cvs = curves->cvs_per_strand;
    CutArgs a{};
    a.points = reinterpret_cast<const float4 *>(curves->points);
    a.counts = curves->point_counts;
    a.roots = curves->roots;
    a.uvs = curves->root_uvs;
    a.strands = n;
    a.cvs = cvs;
    a.div_magic = static_cast<uint32_t>(0xffffffffu / cvs + 1u);
    a.cut = reinterpret_cast<float *>(ws + layout.cut_offset);
    a.tiles = reinterpret_cast<uint32_t *>(ws + layout.tiles_offset);
    a.groups = reinterpret_cast<uint32_t *>(ws + layout.groups_offset);
    a.status_word = reinterpret_cast<uint32_t *>(ws + layout.status_offset);
    a.tile_count = classic_tiles(n);
    a.group_count = classic_groups(n);
    a.length_samples = reinterpret_cast<Sample *>(ws + layout.table_offset);
    a.walk_samples = a.length_samples + kClassicTableEntries;
    a.walk_count = walk_count(cvs);
    a.out_points = reinterpret_cast<float4 *>(output->points);
    a.out_counts = output->point_counts;
    a.out_roots = output->roots;
    a.out_uvs = output->root_uvs;
    a.kept = output->kept_count;
    a.status = output->status;
    return a;
}

}  // namespace
}  // namespace nxg

extern "C" uint32_t nxg_classic_cut_apply(const nxg_classic_curves *curves, const nxg_classic_cut_params *cut,
                                          const nxg_classic_output *output, void *workspace,
                                          uint64_t workspace_bytes, void *stream_ptr) {
    using namespace nxg;
    if (uint32_t s = nxg_validate_classic_cut(curves, cut, output); s != NXG_OK) return s;
    const uint32_t n = curves->strand_count, cvs = curves->cvs_per_strand;
    const ClassicLayout layout = classic_layout(n);
    if (!workspace || workspace_bytes < layout.total || reinterpret_cast<uintptr_t>(workspace) % 256u != 0u)
        return NXG_ERROR_WORKSPACE_TOO_SMALL;
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    CutArgs a = chain_args(curves, output, ws, layout);
    a.scale = cut->amount_scale;
    a.offset = cut->amount_offset;
    a.radius_scale = cut->radius_scale;
    a.multiplier = cut->multiplier;
    a.multiplier_stride = cut->multiplier_stride;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    const bool staged = cvs <= kStagedCvs;
    const size_t tile_lds = staged ? size_t{kWave} * cvs * sizeof(float4) : 0u;
    if (a.tile_count) {
        hipLaunchKernelGGL(cut_table_kernel, dim3(1), dim3(256), 0, stream, a);
        if (staged)
            hipLaunchKernelGGL(cut_search_kernel<true>, dim3(a.tile_count), dim3(kWave), tile_lds, stream, a);
        else
            hipLaunchKernelGGL(cut_search_kernel<false>, dim3(a.tile_count), dim3(kWave), 0, stream, a);
        hipLaunchKernelGGL(cut_scan_tiles_kernel, dim3(a.group_count), dim3(kScanThreads), 0, stream, a);
    }
    hipLaunchKernelGGL(cut_scan_groups_kernel, dim3(1), dim3(kScanThreads), 0, stream, a);
    if (a.tile_count) {
        const size_t held_lds = size_t{kWave} * (cvs + 1u) * sizeof(float4);
        if (cvs <= 8u)
            hipLaunchKernelGGL(cut_write_held_kernel<8>, dim3(a.tile_count), dim3(kWave), held_lds, stream, a);
        else if (cvs <= 16u)
            hipLaunchKernelGGL(cut_write_held_kernel<16>, dim3(a.tile_count), dim3(kWave), held_lds, stream, a);
        else if (cvs <= 32u)
            hipLaunchKernelGGL(cut_write_held_kernel<32>, dim3(a.tile_count), dim3(kWave), held_lds, stream, a);
        else if (staged)
            hipLaunchKernelGGL(cut_write_kernel<true>, dim3(a.tile_count), dim3(kWave), tile_lds, stream, a);
        else
            hipLaunchKernelGGL(cut_write_kernel<false>, dim3(a.tile_count), dim3(kWave), 0, stream, a);
    }
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

extern "C" uint32_t nxg_classic_runtime_apply(const nxg_classic_curves *curves,
                                              const nxg_classic_runtime_inputs *inputs,
                                              const nxg_classic_output *output, void *workspace,
                                              uint64_t workspace_bytes, void *stream_ptr) {
    using namespace nxg;
    if (uint32_t s = nxg_validate_classic_runtime(curves, inputs, output); s != NXG_OK) return s;
    const uint32_t n = curves->strand_count, cvs = curves->cvs_per_strand;
    const nxg_classic_plan_info &plan = *inputs->plan;
    uint32_t guides = 0u;
    for (uint32_t m = 0; m < plan.clump_count; ++m) guides += inputs->clumps[m].guides->guide_count;
    const ClassicLayout layout = classic_layout(n);
    const ClassicRuntimeLayout rt = classic_runtime_layout(n, cvs, guides, plan.rewrites != 0u);
    if (!workspace || workspace_bytes < rt.total || reinterpret_cast<uintptr_t>(workspace) % 256u != 0u)
        return NXG_ERROR_WORKSPACE_TOO_SMALL;
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    CutArgs a = chain_args(curves, output, ws, layout);
    a.radius_scale = inputs->radius_scale;
    a.image = static_cast<const uint8_t *>(inputs->plan_image);
    for (uint32_t k = 0; k < NXG_CLASSIC_EXPR_COUNT; ++k) a.program_offset[k] = plan.program_offset[k];
    a.present = plan.present;
    a.has_cut = plan.trailing_cut;
    a.final_length = plan.final_length;
    constexpr uint64_t one = 1u;
    a.profile = (plan.present & ((one << NXG_CLASSIC_EXPR_WIDTH) | (one << NXG_CLASSIC_EXPR_TAPER) |
                                 (one << NXG_CLASSIC_EXPR_TAPER_START) | (one << NXG_CLASSIC_EXPR_WIDTH_RAMP))) != 0u;
    a.ids = inputs->primitive_ids;
    a.prefixes = inputs->random_prefixes;
    a.columns = inputs->columns;
    a.column_count = inputs->column_count;
    a.description_id = static_cast<float>(plan.description_id);
    a.name_seed = static_cast<float>(plan.description_name_seed);
    a.uniform = plan.uniform & static_cast<uint32_t>(plan.present);
    a.effect_count = plan.effect_count;
    for (uint32_t e = 0; e < plan.effect_count; ++e) a.effects[e] = plan.effects[e].type << 8u | plan.effects[e].module;
    a.rewrites = plan.rewrites;
    a.root_relative = inputs->root_relative ? 1u : 0u;
    a.surface_tangents = inputs->surface_tangents;
    a.noise_domain = inputs->noise_domain_positions;
    a.uniform_values = reinterpret_cast<float2 *>(ws + rt.uniform_offset);
    a.state = reinterpret_cast<RuntimeState *>(ws + rt.state_offset);
    a.err_bits = a.status_word + 1;
    if (plan.rewrites) a.clumped = reinterpret_cast<float4 *>(ws + rt.clumped_offset);
    if (plan.clump_count) {
        a.clump_count = plan.clump_count;
        a.clump_length = plan.clump_length;
        // cLength's first measure can wait for a leading ClumpingFX when no
        // program evaluated before its re-measure reads it: the length and
        // width, and that module's mask, noise, noiseScale, clump and
        // clumpScale (its frequency and correlation run per guide)
        if (plan.effect_count && plan.effects[0].type == NXG_CLASSIC_EFFECT_CLUMP) {
            const uint32_t m0 = plan.effects[0].module;
            uint64_t early = (one << NXG_CLASSIC_EXPR_LENGTH) | (one << NXG_CLASSIC_EXPR_WIDTH);
            for (uint32_t r : {NXG_CLASSIC_EXPR_CLUMP_MASK, NXG_CLASSIC_EXPR_CLUMP_NOISE,
                               NXG_CLASSIC_EXPR_CLUMP_NOISE_SCALE, NXG_CLASSIC_EXPR_CLUMP, NXG_CLASSIC_EXPR_CLUMP_SCALE})
                early |= one << NXG_CLASSIC_EXPR_CLUMP_OF(m0, r);
            a.defer_length = (plan.length_readers & plan.present & early) == 0u ? 1u : 0u;
        }
        uint32_t base = 0u;
        for (uint32_t m = 0; m < plan.clump_count; ++m) {
            const nxg_classic_clump_binding &b = inputs->clumps[m];
            const nxg_clump_guides &g = *b.guides;
            ClumpModule &cm = a.clumps[m];
            cm.guide_count = g.guide_count;
            cm.guide_base = base;
            base += g.guide_count;
            cm.strand_guides = b.strand_guides;
            cm.axes = g.axes;
            cm.normals = g.normals;
            cm.tangents = g.tangents;
            cm.reference = g.reference_positions;
            cm.uvs = b.guide_uvs;
            cm.prefixes = b.guide_random_prefixes;
            cm.local_axes = b.local_axes;
            cm.local_frame = inputs->root_relative && b.local_axes ? 1u : 0u;
        }
        a.clumped = reinterpret_cast<float4 *>(ws + rt.clumped_offset);
        a.guide_render = reinterpret_cast<float4 *>(ws + rt.render_offset);
        a.guide_noise = reinterpret_cast<float4 *>(ws + rt.noise_offset);
        a.guide_info = reinterpret_cast<float2 *>(ws + rt.guide_offset);
    }
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    const bool staged = cvs <= kStagedCvs;
    const size_t tile_lds = staged ? size_t{kWave} * cvs * sizeof(float4) : 0u;
    if (a.tile_count) {
        hipLaunchKernelGGL(cut_table_kernel, dim3(1), dim3(256), 0, stream, a);
        hipLaunchKernelGGL(clear_word_kernel, dim3(1), dim3(1), 0, stream, a.err_bits);
        if (a.uniform) hipLaunchKernelGGL(classic_uniform_kernel, dim3(1), dim3(256), 0, stream, a);
        for (uint32_t m = 0; m < a.clump_count; ++m)
            hipLaunchKernelGGL(runtime_guide_kernel, dim3((a.clumps[m].guide_count + 63u) / 64u), dim3(64), 0,
                               stream, a, m);
        if (staged)
        {
            const size_t rows_lds = size_t{cvs} * kRowStride * sizeof(float);  // runtime_search_kernel's rows
            if (plan.noise_count)
                hipLaunchKernelGGL((runtime_search_kernel<true, true>), dim3(a.tile_count), dim3(kWave), rows_lds,
                                   stream, a);
            else
                hipLaunchKernelGGL((runtime_search_kernel<true, false>), dim3(a.tile_count), dim3(kWave), rows_lds,
                                   stream, a);
        }
        else
            hipLaunchKernelGGL((runtime_search_kernel<false, false>), dim3(a.tile_count), dim3(kWave), 0, stream, a);
        hipLaunchKernelGGL(cut_scan_tiles_kernel, dim3(a.group_count), dim3(kScanThreads), 0, stream, a);
    }
    hipLaunchKernelGGL(cut_scan_groups_kernel, dim3(1), dim3(kScanThreads), 0, stream, a);
    if (a.tile_count) {
        // after an effect that changes the CVs the write starts from the
        // search's copy, already scaled (and rebuilt by every non-trailing cut)
        CutArgs w = a;
        if (a.rewrites) {
            w.points = a.clumped;
            w.present &= ~(uint64_t{1} << NXG_CLASSIC_EXPR_LENGTH);
        }
        // runtime_write_held_kernel: the staged rows, then the kept lanes' ids and one pass's compaction area
        const size_t held_lds = std::max(size_t{cvs} * kRowStride * sizeof(float),
                                         kWave * sizeof(uint32_t) + size_t{kWave} * (kHeldPass + 1u) * sizeof(float4));
        if (cvs <= 8u)
            hipLaunchKernelGGL(runtime_write_held_kernel<8>, dim3(a.tile_count), dim3(kWave), held_lds, stream, w);
        else if (cvs <= 16u)
            hipLaunchKernelGGL(runtime_write_held_kernel<16>, dim3(a.tile_count), dim3(kWave), held_lds, stream, w);
        else if (cvs <= 32u)
            hipLaunchKernelGGL(runtime_write_held_kernel<32>, dim3(a.tile_count), dim3(kWave), held_lds, stream, w);
        else if (staged)
            hipLaunchKernelGGL(runtime_write_kernel<true>, dim3(a.tile_count), dim3(kWave), tile_lds, stream, w);
        else
            hipLaunchKernelGGL(runtime_write_kernel<false>, dim3(a.tile_count), dim3(kWave), 0, stream, w);
    }
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

extern "C" uint32_t nxg_classic_add_endpoints(const nxg_packed_point *points, uint32_t strand_count,
                                              uint32_t cvs_per_strand, const uint32_t *device_strand_count,
                                              nxg_packed_point *out_points, uint64_t out_point_capacity,
                                              uint32_t *out_point_counts, uint64_t out_point_count_capacity,
                                              void *stream_ptr) {
    using namespace nxg;
    const uint64_t n = strand_count, cvs = cvs_per_strand;
    if (cvs < 2u || cvs > NXG_CLASSIC_MAX_CVS || n > NXG_CLASSIC_MAX_STRANDS) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (n == 0u) return NXG_OK;
    if (!points || !out_points) return NXG_ERROR_NULL_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(points) % 16u != 0u || reinterpret_cast<uintptr_t>(out_points) % 16u != 0u)
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (out_point_capacity < n * (cvs + 2u)) return NXG_ERROR_POINT_CAPACITY_TOO_SMALL;
    if (out_point_counts && out_point_count_capacity < n) return NXG_ERROR_POINT_COUNT_CAPACITY_TOO_SMALL;
    const uintptr_t in0 = reinterpret_cast<uintptr_t>(points), in1 = in0 + n * cvs * 16u;
    const uintptr_t o0 = reinterpret_cast<uintptr_t>(out_points), o1 = o0 + n * (cvs + 2u) * 16u;
    if (o0 < in1 && in0 < o1) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    EndpointArgs a{};
    a.points = reinterpret_cast<const float4 *>(points);
    a.out = reinterpret_cast<float4 *>(out_points);
    a.counts = out_point_counts;
    a.device_count = device_strand_count;
    a.strands = strand_count;
    a.cvs = cvs_per_strand;
    a.magic = static_cast<uint32_t>(0xffffffffu / (cvs_per_strand + 2u) + 1u);
    hipLaunchKernelGGL(endpoints_kernel, dim3(classic_tiles(strand_count)), dim3(256), 0,
                       static_cast<hipStream_t>(stream_ptr), a);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

namespace nxg {
namespace {
struct ClumpLayout {
    uint64_t render_offset, noise_offset, lengths_offset, table_offset, total;
};
ClumpLayout clump_layout(uint32_t guides, uint32_t cvs) {
    auto al = [](uint64_t v) { return (v + 255u) & ~uint64_t{255u}; };
    ClumpLayout l{};
    const uint64_t axis = uint64_t{guides} * cvs * sizeof(float4);
    l.render_offset = 0u;
    l.noise_offset = al(axis);
    l.lengths_offset = l.noise_offset + al(axis);
    l.table_offset = l.lengths_offset + al(uint64_t{guides} * 4u);
    l.total = l.table_offset + al(uint64_t{kClassicTableEntries} * sizeof(Sample));
    return l;
}
}  // namespace
}  // namespace nxg

extern "C" uint64_t nxg_classic_clump_workspace_size(uint32_t guide_count, uint32_t cvs_per_strand) {
    return nxg::clump_layout(guide_count, cvs_per_strand).total;
}

extern "C" uint32_t nxg_classic_clump_apply(const nxg_packed_point *points, uint32_t strand_count,
                                            uint32_t cvs_per_strand, const uint32_t *strand_guides,
                                            const nxg_clump_guides *guides, const nxg_clump_params *params,
                                            nxg_packed_point *out_points, uint64_t out_point_capacity,
                                            uint32_t *status, void *workspace, uint64_t workspace_bytes,
                                            void *stream_ptr) {
    using namespace nxg;
    if (!guides || !params) return NXG_ERROR_NULL_ARGUMENT;
    const uint64_t n = strand_count, cvs = cvs_per_strand;
    if (cvs < 3u || cvs > NXG_CLUMP_MAX_CVS || n > NXG_CLASSIC_MAX_STRANDS) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    const float k[7] = {params->mask, params->amount, params->clump_scale, params->noise, params->noise_scale,
                        params->noise_frequency, params->noise_correlation};
    for (float v : k)
        if (!std::isfinite(v)) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (params->amount_multiplier && params->amount_stride == 0u) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (n == 0u) return NXG_OK;
    if (!points || !out_points || !strand_guides || !guides->axes || !guides->normals || !guides->tangents)
        return NXG_ERROR_NULL_ARGUMENT;
    if (guides->guide_count == 0u) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (reinterpret_cast<uintptr_t>(points) % 16u != 0u || reinterpret_cast<uintptr_t>(out_points) % 16u != 0u)
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (out_point_capacity < n * cvs) return NXG_ERROR_POINT_CAPACITY_TOO_SMALL;
    const uintptr_t in0 = reinterpret_cast<uintptr_t>(points), in1 = in0 + n * cvs * 16u;
    const uintptr_t o0 = reinterpret_cast<uintptr_t>(out_points), o1 = o0 + n * cvs * 16u;
    if (o0 != in0 && o0 < in1 && in0 < o1) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    const ClumpLayout layout = clump_layout(guides->guide_count, cvs_per_strand);
    if (!workspace || workspace_bytes < layout.total || reinterpret_cast<uintptr_t>(workspace) % 256u != 0u)
        return NXG_ERROR_WORKSPACE_TOO_SMALL;
    // the module's constants, evaluated as the reference evaluates them per strand
    const float mask = params->mask < 0.0f ? 0.0f : (1.0f < params->mask ? 1.0f : params->mask);
    const float noise = params->noise < 0.0f ? 0.0f : params->noise;
    const float noise_scale = params->noise_scale < 0.0f ? 0.0f : params->noise_scale;
    const bool noisy = noise > 1.0e-5f && mask > 1.0e-4f;
    const float magnitude = noisy ? mask * noise * noise_scale : 0.0f;
    if (!std::isfinite(magnitude)) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    ClumpArgs a{};
    a.points = reinterpret_cast<const float4 *>(points);
    a.out = reinterpret_cast<float4 *>(out_points);
    a.strand_guides = strand_guides;
    a.axes = guides->axes;
    a.normals = guides->normals;
    a.tangents = guides->tangents;
    a.reference = guides->reference_positions;
    a.strands = strand_count;
    a.cvs = cvs_per_strand;
    a.guides = guides->guide_count;
    a.div_magic = static_cast<uint32_t>(0xffffffffu / cvs_per_strand + 1u);
    a.mask = mask;
    a.amount = params->amount;
    a.clump_scale = params->clump_scale;
    a.magnitude = magnitude;
    a.frequency = params->noise_frequency < 0.0f ? 0.0f : params->noise_frequency;
    const float corr = params->noise_correlation * 0.01f;
    a.correlation = corr < 0.0f ? 0.0f : (1.0f < corr ? 1.0f : corr);
    a.amount_multiplier = params->amount_multiplier;
    a.amount_stride = params->amount_stride;
    a.render = reinterpret_cast<float4 *>(ws + layout.render_offset);
    a.noise = reinterpret_cast<float4 *>(ws + layout.noise_offset);
    a.lengths = reinterpret_cast<float *>(ws + layout.lengths_offset);
    a.length_samples = reinterpret_cast<const Sample *>(ws + layout.table_offset);
    a.status = status;
    CutArgs table{};
    table.cvs = cvs_per_strand;
    table.length_samples = reinterpret_cast<Sample *>(ws + layout.table_offset);
    table.walk_count = 0u;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    if (status) hipLaunchKernelGGL(clear_word_kernel, dim3(1), dim3(1), 0, stream, status);
    hipLaunchKernelGGL(cut_table_kernel, dim3(1), dim3(256), 0, stream, table);
    hipLaunchKernelGGL(clump_guide_kernel, dim3((a.guides + 63u) / 64u), dim3(64), 0, stream, a);
    const uint32_t tiles = classic_tiles(strand_count);
    const size_t tile_lds = size_t{kWave} * cvs * sizeof(float4);
    const size_t held_lds = size_t{kWave} * (cvs + 1u) * sizeof(float4);
    if (cvs <= 8u)
        hipLaunchKernelGGL(clump_strand_kernel<8>, dim3(tiles), dim3(kWave), held_lds, stream, a);
    else if (cvs <= 16u)
        hipLaunchKernelGGL(clump_strand_kernel<16>, dim3(tiles), dim3(kWave), held_lds, stream, a);
    else if (cvs <= 32u)
        hipLaunchKernelGGL(clump_strand_kernel<32>, dim3(tiles), dim3(kWave), held_lds, stream, a);
    else
        hipLaunchKernelGGL(clump_strand_kernel<0>, dim3(tiles), dim3(kWave), tile_lds, stream, a);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}
