// The renderer curve payload (build_curve_batches, src/curve_payload.cpp:
// 150-301) on the GPU: curves cut into batches of at most
// max_strands_per_batch strands, points transformed, widths turned into radii,
// curves optionally resampled, motion samples alike.
//
//   offsets_block_kernel    per curve set (blockIdx.y: base, then each motion
//   offsets_groups_kernel   sample), the exclusive scan of its point counts
//   offsets_add_kernel      (or of the fixed count) into CV offsets, checking
//                           counts >= 2 and their sum against point_count.
//   batch_points_kernel     each batch's first output point.
//   payload_points_kernel   one wavefront per 64-strand tile, its lanes over
//                           the tile's contiguous run of output points (a
//                           search among the tile's 65 offsets in LDS finds
//                           each point's strand): transform, radius,
//                           resampling; point counts and root uvs per strand.
//   payload_motion_kernel   the same run per motion sample (blockIdx.y),
//                           written batch-major and sample-major in a batch.
// Everything is HBM traffic plus a few flops per point: the bound is HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/nanoxgen_amd.h"
#include "block_scan.h"
#include "internal.h"

namespace nxg {
namespace {

constexpr uint32_t kTile = 64u;
constexpr uint32_t kBlockValues = 1024u;  // counts scanned per block (4 per thread)
constexpr uint32_t kMaxSets = 1u + NXG_PAYLOAD_MAX_MOTION_SAMPLES;

struct ScanJob {
    const uint32_t *counts;  // NULL: every curve has `fixed` CVs
    uint32_t *offsets;       // strands + 1
    uint32_t *blocks;        // one per kBlockValues strands
    uint64_t point_count;
    uint32_t fixed;
};

struct ScanArgs {
    ScanJob job[kMaxSets];
    uint32_t strands, block_count;
    uint32_t *status;  // workspace word
};

__global__ void __launch_bounds__(kScanThreads) offsets_block_kernel(ScanArgs a) {
    const ScanJob &j = a.job[blockIdx.y];
    const uint32_t base = blockIdx.x * kBlockValues + threadIdx.x * 4u;
    uint32_t v[4], sum = 0u, bad = 0u;
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        v[k] = base + k < a.strands ? (j.counts ? j.counts[base + k] : j.fixed) : 0u;
        if (base + k < a.strands && v[k] < 2u) bad = 1u;
        sum += v[k];
    }
    uint32_t total, flags;
    uint32_t run = block_scan(sum, bad, total, flags);
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k) {
        if (base + k < a.strands) j.offsets[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0u) {
        j.blocks[blockIdx.x] = total;
        if (flags) atomicOr(a.status, NXG_PAYLOAD_INVALID_COUNTS);
    }
}

__global__ void __launch_bounds__(kScanThreads) offsets_groups_kernel(ScanArgs a) {
    const ScanJob &j = a.job[blockIdx.x];
    uint64_t carry = 0u;
    for (uint32_t first = 0u; first < a.block_count; first += kScanThreads) {
        const uint32_t i = first + threadIdx.x;
        const uint32_t v = i < a.block_count ? j.blocks[i] : 0u;
        uint32_t total, flags;
        const uint32_t ex = block_scan(v, 0u, total, flags);
        if (i < a.block_count) j.blocks[i] = static_cast<uint32_t>(carry) + ex;
        carry += total;
    }
    if (threadIdx.x == 0u) {
        j.offsets[a.strands] = static_cast<uint32_t>(carry);
        if (carry != j.point_count) atomicOr(a.status, NXG_PAYLOAD_INVALID_COUNTS);
    }
}

__global__ void __launch_bounds__(kScanThreads) offsets_add_kernel(ScanArgs a) {
    const ScanJob &j = a.job[blockIdx.y];
    const uint32_t base = blockIdx.x * kBlockValues + threadIdx.x * 4u;
    const uint32_t add = j.blocks[blockIdx.x];
#pragma unroll
    for (uint32_t k = 0; k < 4u; ++k)
        if (base + k < a.strands) j.offsets[base + k] += add;
}

__global__ void clear_status_kernel(uint32_t *word, uint32_t *caller) {
    *word = 0u;
    if (caller) *caller = 0u;
}
__global__ void copy_status_kernel(const uint32_t *word, uint32_t *caller) { *caller = *word; }

struct MotionSet {
    const nxg_vec3 *points;
    const uint32_t *offsets;  // NULL: `fixed` CVs per curve
    uint32_t fixed;
};

struct PayloadArgs {
    const nxg_vec3 *points;
    const float *widths;
    const nxg_root_sample *roots;
    const nxg_vec2 *root_uvs;
    const uint32_t *offsets;  // base CV offsets (strands + 1); NULL: `fixed` CVs per curve
    uint32_t fixed;
    MotionSet motion[NXG_PAYLOAD_MAX_MOTION_SAMPLES];
    uint32_t motion_count;
    uint32_t strands, per_batch, batch_count, resample;
    float constant_width, radius_scale, radius_factor;
    float m[12];
    uint32_t identity;
    const uint32_t *status;  // the scan's word: nothing is written when set
    nxg_packed_point *out_points;
    uint32_t *out_counts;
    nxg_vec2 *out_uvs;
    nxg_vec3 *out_motion;
    uint64_t *batch_points;
};

__device__ __forceinline__ uint32_t in_offset(const PayloadArgs &a, uint32_t s) {
    return a.offsets ? a.offsets[s] : s * a.fixed;
}
__device__ __forceinline__ uint32_t motion_offset(const MotionSet &m, uint32_t s) {
    return m.offsets ? m.offsets[s] : s * m.fixed;
}

// Output offset of strand s: s * resample, or its input offset.
__device__ __forceinline__ uint64_t out_offset(const PayloadArgs &a, uint32_t s) {
    return a.resample ? uint64_t{s} * a.resample : in_offset(a, s);
}

__device__ __forceinline__ nxg_vec3 transform(const PayloadArgs &a, nxg_vec3 p) {
    if (a.identity) return p;  // AffineTransform::apply_point (src/curve_payload.cpp:128-134)
    return {a.m[0] * p.x + a.m[1] * p.y + a.m[2] * p.z + a.m[3], a.m[4] * p.x + a.m[5] * p.y + a.m[6] * p.z + a.m[7],
            a.m[8] * p.x + a.m[9] * p.y + a.m[10] * p.z + a.m[11]};
}

// sample_point (src/curve_payload.cpp:73-82): linear in index space.
__device__ __forceinline__ nxg_vec3 sample_point(const nxg_vec3 *pts, uint32_t first_cv, uint32_t count, float t,
                                                 uint32_t *first, float *frac) {
    const float tc = t < 0.0f ? 0.0f : (1.0f < t ? 1.0f : t);
    const float x = tc * static_cast<float>(count - 1u);
    const uint32_t i = static_cast<uint32_t>(x);
    *first = i;
    if (i >= count - 1u) return pts[first_cv + count - 1u];
    const float f = x - static_cast<float>(i);
    *frac = f;
    const nxg_vec3 a = pts[first_cv + i], b = pts[first_cv + i + 1u];
    return {a.x * (1.0f - f) + b.x * f, a.y * (1.0f - f) + b.y * f, a.z * (1.0f - f) + b.z * f};
}

__global__ void __launch_bounds__(kScanThreads) batch_points_kernel(PayloadArgs a) {
    if (*a.status) return;
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b <= a.batch_count; b += gridDim.x * blockDim.x) {
        const uint64_t s = uint64_t{b} * a.per_batch;
        a.batch_points[b] = out_offset(a, s < a.strands ? static_cast<uint32_t>(s) : a.strands);
    }
}

// The tile's 65 output offsets into LDS; the strand owning output point e.
struct TileRun {
    uint64_t first, end;
    uint32_t count;
};
__device__ __forceinline__ TileRun tile_run(const PayloadArgs &a, uint64_t *offs, uint32_t first) {
    const uint32_t count = a.strands - first < kTile ? a.strands - first : kTile;
    for (uint32_t i = threadIdx.x; i <= count; i += blockDim.x) offs[i] = out_offset(a, first + i);
    __syncthreads();
    return {offs[0], offs[count], count};
}
__device__ __forceinline__ uint32_t owner(const uint64_t *offs, uint32_t count, uint64_t e) {
    uint32_t lo = 0u, hi = count;  // offs[lo] <= e < offs[hi]
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) / 2u;
        if (offs[mid] <= e) lo = mid; else hi = mid;
    }
    return lo;
}

// Output points in flight per thread: a loop that stores each point before
// loading the next waits one memory round trip per point.  The batched
// loops are branch-free: a thread past the tile's end loads and stores its
// last point again (the same value to the same place).
constexpr uint32_t kBatch = 8u;

__global__ void __launch_bounds__(kTile) payload_points_kernel(PayloadArgs a) {
    __shared__ uint64_t offs[kTile + 1u];
    if (*a.status) return;
    const uint32_t first = blockIdx.x * kTile;
    const TileRun run = tile_run(a, offs, first);
    if (!a.resample) {  // src/curve_payload.cpp:215-246: point e from input point e
        for (uint64_t base = run.first + threadIdx.x; base < run.end; base += kBatch * kTile) {
            nxg_vec3 p[kBatch];
            float w[kBatch];
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) p[j] = a.points[min(base + j * kTile, run.end - 1u)];
            if (a.widths) {
#pragma unroll
                for (uint32_t j = 0; j < kBatch; ++j) w[j] = a.widths[min(base + j * kTile, run.end - 1u)];
            } else {
#pragma unroll
                for (uint32_t j = 0; j < kBatch; ++j) w[j] = a.constant_width;
            }
#pragma unroll
            for (uint32_t j = 0; j < kBatch; ++j) {
                const nxg_vec3 q = transform(a, p[j]);
                a.out_points[min(base + j * kTile, run.end - 1u)] = {q.x, q.y, q.z, w[j] * a.radius_factor};
            }
        }
    }
    for (uint64_t e = run.first + threadIdx.x; a.resample && e < run.end; e += kTile) {
        const uint32_t r = owner(offs, run.count, e);
        const uint32_t s = first + r;
        const uint32_t cv = static_cast<uint32_t>(e - offs[r]);
        // resampled (src/curve_payload.cpp:247-259)
        const uint32_t c0 = in_offset(a, s), count = in_offset(a, s + 1u) - c0;
        const float t = static_cast<float>(cv) / static_cast<float>(a.resample - 1u);
        uint32_t i;
        float f = 0.0f;
        const nxg_vec3 p = transform(a, sample_point(a.points, c0, count, t, &i, &f));
        float width = a.constant_width;  // sample_width (src/curve_payload.cpp:84-93)
        if (a.widths)
            width = i >= count - 1u ? a.widths[c0 + count - 1u]
                                    : a.widths[c0 + i] * (1.0f - f) + a.widths[c0 + i + 1u] * f;
        a.out_points[e] = {p.x, p.y, p.z, 0.5f * width * a.radius_scale};
    }
    if (threadIdx.x < run.count) {
        const uint32_t s = first + threadIdx.x;
        a.out_counts[s] = static_cast<uint32_t>(offs[threadIdx.x + 1u] - offs[threadIdx.x]);
        if (a.out_uvs) a.out_uvs[s] = a.root_uvs ? a.root_uvs[s] : a.roots[s].uv;
    }
}

__global__ void __launch_bounds__(kTile) payload_motion_kernel(PayloadArgs a) {
    __shared__ uint64_t offs[kTile + 1u];
    if (*a.status) return;
    const uint32_t k = blockIdx.y;
    const MotionSet m = a.motion[k];
    const uint32_t first = blockIdx.x * kTile;
    const TileRun run = tile_run(a, offs, first);
    // kBatch output points per thread with their loads in flight together:
    // first every point's strand, CV and motion offsets (whose loads, when
    // the set has offsets, would otherwise wait on the point loads before
    // them), then its two source indices without branches (the copied point
    // of src/curve_payload.cpp:270-272 loads the same CV twice; sample_point's
    // clamp at the last CV likewise) and every load, then the stores
    for (uint64_t base = run.first + threadIdx.x; base < run.end; base += kBatch * kTile) {
        uint32_t strand[kBatch], cv[kBatch], out_count[kBatch], m0[kBatch], m_count[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint64_t e = min(base + j * kTile, run.end - 1u);  // as in payload_points_kernel
            const uint32_t r = owner(offs, run.count, e);
            strand[j] = first + r;
            cv[j] = static_cast<uint32_t>(e - offs[r]);
            out_count[j] = static_cast<uint32_t>(offs[r + 1u] - offs[r]);
            m0[j] = motion_offset(m, strand[j]);
            m_count[j] = motion_offset(m, strand[j] + 1u) - m0[j];
        }
        nxg_vec3 p0[kBatch], p1[kBatch];
        float frac[kBatch];
        bool blend[kBatch];
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const bool copy = !a.resample && m_count[j] == out_count[j];
            // sample_point (src/curve_payload.cpp:73-82) at t = cv / (out_count - 1)
            const float t = static_cast<float>(cv[j]) / static_cast<float>(out_count[j] - 1u);
            const float tc = t < 0.0f ? 0.0f : (1.0f < t ? 1.0f : t);
            const float x = tc * static_cast<float>(m_count[j] - 1u);
            const uint32_t i = static_cast<uint32_t>(x);
            const bool last = i >= m_count[j] - 1u;
            blend[j] = !copy && !last;
            frac[j] = x - static_cast<float>(i);
            const uint32_t i0 = copy ? cv[j] : (last ? m_count[j] - 1u : i);
            const uint32_t i1 = blend[j] ? i + 1u : i0;
            p0[j] = m.points[m0[j] + i0];
            p1[j] = m.points[m0[j] + i1];
        }
#pragma unroll
        for (uint32_t j = 0; j < kBatch; ++j) {
            const uint64_t e = min(base + j * kTile, run.end - 1u);
            const uint32_t b = strand[j] / a.per_batch;
            const uint64_t b0 = a.batch_points[b], b1 = a.batch_points[b + 1u];
            const float f = frac[j];
            const nxg_vec3 q = blend[j] ? nxg_vec3{p0[j].x * (1.0f - f) + p1[j].x * f, p0[j].y * (1.0f - f) + p1[j].y * f,
                                                   p0[j].z * (1.0f - f) + p1[j].z * f}
                                        : p0[j];
            a.out_motion[uint64_t{a.motion_count} * b0 + uint64_t{k} * (b1 - b0) + (e - b0)] = transform(a, q);
        }
    }
}

struct PayloadLayout {
    uint64_t offsets[kMaxSets], blocks[kMaxSets], status, total;
};
PayloadLayout payload_layout(uint32_t strands, uint32_t motion_count) {
    auto al = [](uint64_t v) { return (v + 255u) & ~uint64_t{255u}; };
    PayloadLayout l{};
    uint64_t at = 0u;
    const uint64_t blocks = (uint64_t{strands} + kBlockValues - 1u) / kBlockValues;
    for (uint32_t k = 0; k <= motion_count && k < kMaxSets; ++k) {
        l.offsets[k] = at;
        at += al((uint64_t{strands} + 1u) * 4u);
        l.blocks[k] = at;
        at += al(blocks * 4u);
    }
    l.status = at;
    l.total = at + 256u;
    return l;
}

void put(char *msg, uint64_t cap, const char *text) {
    if (msg && cap) std::snprintf(msg, cap, "%s", text);
}

bool is_identity(const float *m) {
    static const float id[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    return std::memcmp(m, id, sizeof(id)) == 0;
}

}  // namespace
}  // namespace nxg

extern "C" {

uint32_t nxg_curve_payload_plan(const nxg_curve_set *base, const nxg_curve_set *motion, uint32_t motion_count,
                                const nxg_curve_payload_options *options, uint32_t *batch_count,
                                uint64_t *output_points, char *err, uint64_t err_capacity) {
    using nxg::put;
    auto fail = [&](const char *text) {
        put(err, err_capacity, text);
        return static_cast<uint32_t>(NXG_ERROR_INVALID_BUILD_INPUT);
    };
    if (!base || !options || (motion_count && !motion)) return NXG_ERROR_NULL_ARGUMENT;
    // src/curve_payload.cpp:155-196, in its order
    if (options->max_strands_per_batch == 0u) return fail("curve batches need a positive strand limit");
    if (options->resample_cv_count == 1u) return fail("curve resampling needs at least two CVs");
    if (!std::isfinite(options->radius_scale) || options->radius_scale < 0.0f)
        return fail("curve radius scale must be finite and non-negative");
    auto check_view = [&](const nxg_curve_set &v, bool is_base) -> uint32_t {
        if (v.strand_count == 0u) return fail(is_base ? "base curves has no strands" : "motion curves has no strands");
        if (!v.point_counts && v.fixed_cvs_per_strand < 2u)
            return fail(is_base ? "base curves needs fixed or variable CV counts"
                                : "motion curves needs fixed or variable CV counts");
        if (!v.point_counts && v.point_count != uint64_t{v.strand_count} * v.fixed_cvs_per_strand)
            return fail(is_base ? "base curves point array size mismatch" : "motion curves point array size mismatch");
        if (!v.points || v.point_count >= (uint64_t{1} << 32)) return NXG_ERROR_INVALID_BUILD_INPUT;
        if (!v.widths && (!std::isfinite(v.constant_width) || v.constant_width < 0.0f))
            return fail(is_base ? "base curves has an invalid constant width"
                                : "motion curves has an invalid constant width");
        return NXG_OK;
    };
    if (uint32_t s = check_view(*base, true); s != NXG_OK) return s;
    if (motion_count > NXG_PAYLOAD_MAX_MOTION_SAMPLES) return NXG_ERROR_INVALID_BUILD_INPUT;
    for (uint32_t k = 0; k < motion_count; ++k) {
        if (!std::isfinite(motion[k].time)) return fail("motion sample time must be finite");
        if (motion[k].strand_count != base->strand_count) return fail("motion sample strand-count mismatch");
        if (uint32_t s = check_view(motion[k], false); s != NXG_OK) return s;
    }
    for (float v : options->transform)
        if (!std::isfinite(v)) return fail("curve transform must contain finite values");
    const uint64_t n = base->strand_count, per = options->max_strands_per_batch;
    if (batch_count) *batch_count = static_cast<uint32_t>((n + per - 1u) / per);
    if (output_points)
        *output_points = options->resample_cv_count ? n * options->resample_cv_count : base->point_count;
    return NXG_OK;
}

uint64_t nxg_curve_payload_workspace_size(uint32_t strand_count, uint32_t motion_count) {
    return nxg::payload_layout(strand_count, motion_count).total;
}

uint32_t nxg_curve_payload_build(const nxg_curve_set *base, const nxg_curve_set *motion, uint32_t motion_count,
                                 const nxg_curve_payload_options *options, const nxg_curve_payload_output *out,
                                 void *workspace, uint64_t workspace_bytes, void *stream_ptr) {
    using namespace nxg;
    uint32_t batches = 0u;
    uint64_t total = 0u;
    if (uint32_t s = nxg_curve_payload_plan(base, motion, motion_count, options, &batches, &total, nullptr, 0u);
        s != NXG_OK)
        return s;
    if (!out || !out->points || !out->point_counts || !out->batch_points) return NXG_ERROR_NULL_ARGUMENT;
    const uint64_t n = base->strand_count;
    const bool has_uvs = base->root_uvs || base->roots;
    if (out->point_capacity < total) return NXG_ERROR_POINT_CAPACITY_TOO_SMALL;
    if (out->point_count_capacity < n) return NXG_ERROR_POINT_COUNT_CAPACITY_TOO_SMALL;
    if (has_uvs && (!out->root_uvs || out->root_uv_capacity < n)) return NXG_ERROR_ROOT_UV_CAPACITY_TOO_SMALL;
    if (motion_count && (!out->motion_points || out->motion_point_capacity < total * motion_count))
        return NXG_ERROR_MOTION_POINT_CAPACITY_TOO_SMALL;
    if (out->batch_capacity < uint64_t{batches} + 1u) return NXG_ERROR_BUFFER_TOO_SMALL;
    const PayloadLayout layout = payload_layout(base->strand_count, motion_count);
    if (!workspace || workspace_bytes < layout.total || reinterpret_cast<uintptr_t>(workspace) % 256u != 0u)
        return NXG_ERROR_WORKSPACE_TOO_SMALL;
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    ScanArgs sa{};
    sa.strands = base->strand_count;
    sa.block_count = static_cast<uint32_t>((n + kBlockValues - 1u) / kBlockValues);
    sa.status = reinterpret_cast<uint32_t *>(ws + layout.status);
    for (uint32_t k = 0; k <= motion_count; ++k) {
        const nxg_curve_set &v = k ? motion[k - 1u] : *base;
        sa.job[k] = {v.point_counts, reinterpret_cast<uint32_t *>(ws + layout.offsets[k]),
                     reinterpret_cast<uint32_t *>(ws + layout.blocks[k]), v.point_count, v.fixed_cvs_per_strand};
    }
    PayloadArgs a{};
    a.points = base->points;
    a.widths = base->widths;
    a.roots = base->roots;
    a.root_uvs = base->root_uvs;
    // fixed-count sets need no scan: their offsets are s * fixed
    bool scan = false;
    for (uint32_t k = 0; k <= motion_count; ++k) scan |= (k ? motion[k - 1u] : *base).point_counts != nullptr;
    a.offsets = base->point_counts ? sa.job[0].offsets : nullptr;
    a.fixed = base->fixed_cvs_per_strand;
    for (uint32_t k = 0; k < motion_count; ++k)
        a.motion[k] = {motion[k].points, motion[k].point_counts ? sa.job[k + 1u].offsets : nullptr,
                       motion[k].fixed_cvs_per_strand};
    a.motion_count = motion_count;
    a.strands = base->strand_count;
    a.per_batch = options->max_strands_per_batch;
    a.batch_count = batches;
    a.resample = options->resample_cv_count;
    a.constant_width = base->constant_width;
    a.radius_scale = options->radius_scale;
    a.radius_factor = 0.5f * options->radius_scale;
    std::memcpy(a.m, options->transform, sizeof(a.m));
    a.identity = is_identity(options->transform) ? 1u : 0u;
    a.status = sa.status;
    a.out_points = out->points;
    a.out_counts = out->point_counts;
    a.out_uvs = has_uvs ? out->root_uvs : nullptr;
    a.out_motion = out->motion_points;
    a.batch_points = out->batch_points;
    hipLaunchKernelGGL(clear_status_kernel, dim3(1), dim3(1), 0, stream, sa.status, out->status);
    if (scan) {
        hipLaunchKernelGGL(offsets_block_kernel, dim3(sa.block_count, motion_count + 1u), dim3(kScanThreads), 0,
                           stream, sa);
        hipLaunchKernelGGL(offsets_groups_kernel, dim3(motion_count + 1u), dim3(kScanThreads), 0, stream, sa);
        hipLaunchKernelGGL(offsets_add_kernel, dim3(sa.block_count, motion_count + 1u), dim3(kScanThreads), 0,
                           stream, sa);
    }
    hipLaunchKernelGGL(batch_points_kernel, dim3((batches + kScanThreads) / kScanThreads), dim3(kScanThreads), 0,
                       stream, a);
    const uint32_t tiles = static_cast<uint32_t>((n + kTile - 1u) / kTile);
    hipLaunchKernelGGL(payload_points_kernel, dim3(tiles), dim3(kTile), 0, stream, a);
    if (motion_count)
        hipLaunchKernelGGL(payload_motion_kernel, dim3(tiles, motion_count), dim3(kTile), 0, stream, a);
    if (out->status) hipLaunchKernelGGL(copy_status_kernel, dim3(1), dim3(1), 0, stream, sa.status, out->status);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

}  // extern "C"
