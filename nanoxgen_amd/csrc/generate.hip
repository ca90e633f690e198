// gfx950 strand generation: the device half of libnanoxgen_amd.so.
//
// Two kernels per request, both on the caller's stream:
//   guide_prep_kernel   one lane per (guide, output CV): the guide's current
//                       root/normal/support radius (generate.h:137-172) and
//                       its offset sample_guide(t_cv) - guide_root
//                       (generate.h:189-205, 257-258) into the workspace.
//                       The result depends only on the guide and t, so every
//                       strand that blends the guide reads it instead of
//                       re-interpolating.
//   strand_kernel       one lane per strand, a 64-lane wavefront per tile
//                       of 64 strands: alias-table root sample, stencil
//                       blend, base polyline, noise cursor with the
//                       transported frame, taper, and float4 output written
//                       through LDS as whole 128-byte lines.
// The arithmetic follows the reference's generate.h operation by operation,
// compiled without FP contraction and with correctly rounded float division
// and square root, so everything that involves only +,-,*,/,sqrt (root
// identity, stencil membership, gradient noise values) is bit-identical to
// the reference's CPU path; acosf/sinf/cosf/powf come from the device math
// library and feed positions only.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>

#include "../../include/nanoxgen_amd.h"
#include "expr_device.h"
#include "expr_format.h"
#include "internal.h"

namespace nxg {
namespace {

constexpr int kWave = 64;
constexpr int kChunk = 8;        // CVs staged in LDS between stores
constexpr int kStageStride = 66; // floats per staged row (bank-skewed)

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return {x, y, z}; }
__device__ __forceinline__ V3 ld(const nxg_vec3 &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return {a.x * s, a.y * s, a.z * s}; }
__device__ __forceinline__ V3 operator/(V3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ float len2(V3 v) { return dot(v, v); }
__device__ __forceinline__ float fmax_ref(float a, float b) { return a > b ? a : b; }
__device__ __forceinline__ float fmin_ref(float a, float b) { return a < b ? a : b; }
__device__ __forceinline__ float clamp_ref(float x, float lo, float hi) {
    return fmin_ref(fmax_ref(x, lo), hi);
}
__device__ __forceinline__ V3 normalize(V3 v) {  // generate.h:67-70
    const float l2 = len2(v);
    return l2 > 1.0e-20f ? v / sqrtf(l2) : v3(0.0f, 1.0f, 0.0f);
}

// Positions-only arithmetic.  Everything below that is marked "fast" feeds
// only point positions, which are checked against the reference to a
// tolerance; it uses fused multiply-add and the hardware's 1-ulp reciprocal,
// square root and reciprocal square root.  Root samples, stencil membership,
// blend weights and radii never go through these helpers.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float fast_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float fma3(V3 a, V3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
__device__ __forceinline__ V3 fma_v(V3 a, float s, V3 c) {  // a * s + c
    return {fmaf(a.x, s, c.x), fmaf(a.y, s, c.y), fmaf(a.z, s, c.z)};
}
__device__ __forceinline__ V3 cross_fast(V3 a, V3 b) {
    return {fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)), fmaf(a.x, b.y, -(a.y * b.x))};
}
__device__ __forceinline__ V3 normalize_fast(V3 v) {
    const float l2 = fma3(v, v);
    return l2 > 1.0e-20f ? v * fast_rsqrt(l2) : v3(0.0f, 1.0f, 0.0f);
}

__device__ __forceinline__ uint32_t hash32(uint32_t x) {  // generate.h:72-79
    x ^= x >> 16u;
    x *= 0x7feb352du;
    x ^= x >> 15u;
    x *= 0x846ca68bu;
    x ^= x >> 16u;
    return x;
}
__device__ __forceinline__ float random01(uint32_t seed, uint32_t strand, uint32_t dim) {
    const uint32_t h = hash32(seed ^ hash32(strand + 0x9e3779b9u * (dim + 1u)));
    return static_cast<float>(h >> 8u) * (1.0f / 16777216.0f);
}

// The LCG of noise_hash (generate.h:353-364) for lattice point (x, y, z)
// is x*a^2 + y*a + z + c*(a^2 + a + 1) modulo 2^32 with a = 1664525 and
// c = 1013904223, so along one moving axis it is one multiply plus a
// per-strand constant.  Tempering is applied to that state as in the
// reference; only its low 8 bits select the gradient.
constexpr uint32_t kLcgA = 1664525u;
constexpr uint32_t kLcgA2 = kLcgA * kLcgA;
constexpr uint32_t kLcgC = 1013904223u;
constexpr uint32_t kLcgK = kLcgC * kLcgA2 + kLcgC * kLcgA + kLcgC;

__device__ __forceinline__ uint32_t temper_index(uint32_t seed) {
    seed ^= seed >> 11u;
    seed = xor_and(seed, seed << 7u, 0x9d2c5680u);
    seed = xor_and(seed, seed << 15u, 0xefc60000u);
    seed ^= seed >> 18u;
    return (((seed & 0x00ff0000u) >> 4u) + (seed & 0xffu)) & 0xffu;
}

// noise_s_curve (generate.h:366-368), fast form
__device__ __forceinline__ float s_curve(float t) {
    return (t * t * t) * fmaf(t, fmaf(t, 6.0f, -15.0f), 10.0f);
}

// (1 - alpha) * a + alpha * b (generate.h:397-398), fast form
__device__ __forceinline__ float lerp_fast(float a, float b, float alpha) {
    return fmaf(alpha, b - a, a);
}

// Per-strand constants of the three noise samples.  A strand's noise point
// for output axis A is the domain point moved along lattice axis A by
// distance = arc * frequency (generate.h:486-495), so the two other lattice
// coordinates never change along the strand.
struct NoiseDomain {
    float d[3];           // domain point
    float w[3], alpha[3]; // fractional lattice offset and its s-curve per axis
    uint32_t lcg[3];      // lattice coordinate of each axis times its LCG weight
};

__device__ __forceinline__ NoiseDomain make_domain(V3 d) {
    NoiseDomain n;
    n.d[0] = d.x;
    n.d[1] = d.y;
    n.d[2] = d.z;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const float f = floorf(n.d[ax]);
        n.w[ax] = n.d[ax] - f;
        n.alpha[ax] = s_curve(n.w[ax]);
        const uint32_t mult = ax == 0 ? kLcgA2 : (ax == 1 ? kLcgA : 1u);
        n.lcg[ax] = static_cast<uint32_t>(static_cast<int32_t>(f)) * mult;
    }
    return n;
}

// The lattice cube of gradient_noise (generate.h:370-402) has two faces
// across the moving axis A, at lattice coordinates j and j + 1.  Each
// corner's contribution g . (w - corner) is linear in the moving offset, and
// the trilinear reduction is linear in the corner values, so a face reduces
// over the two fixed axes to  value(w_A) = w_A * p + q.  A strand crosses a
// new lattice plane only every few CVs, so faces are cached and the noise at
// a CV costs one face (four corners) at most instead of eight corners.
struct FaceCache {
    int32_t cell;          // lattice coordinate of face 0 (face 1 is cell + 1)
    float p0, q0, p1, q1;
};

template <int A>
__device__ __forceinline__ void noise_face(const NoiseDomain &n, int32_t j, const float4 *grad, float &p,
                                           float &q) {
    constexpr int F1 = A == 0 ? 1 : 0;  // fixed axes in x, y, z order
    constexpr int F2 = A == 2 ? 1 : 2;
    constexpr uint32_t kMult[3] = {kLcgA2, kLcgA, 1u};
    // the LCG seed of corner (x, y, z) is x*a^2 + y*a + z + K (generate.h:356-358)
    const uint32_t h = static_cast<uint32_t>(j) * kMult[A] + n.lcg[F1] + n.lcg[F2] + kLcgK;
    float gp[4], gq[4];
#pragma unroll
    for (int f = 0; f < 4; ++f) {
        const float o1 = static_cast<float>(f & 1), o2 = static_cast<float>(f >> 1);
        const uint32_t corner = (f & 1 ? kMult[F1] : 0u) + (f & 2 ? kMult[F2] : 0u);
        const float4 g = grad[temper_index(h + corner)];
        const float ga = A == 0 ? g.x : (A == 1 ? g.y : g.z);
        const float g1 = F1 == 0 ? g.x : g.y;
        const float g2 = F2 == 1 ? g.y : g.z;
        gp[f] = ga;
        gq[f] = fmaf(g2, n.w[F2] - o2, g1 * (n.w[F1] - o1));
    }
    p = lerp_fast(lerp_fast(gp[0], gp[1], n.alpha[F1]), lerp_fast(gp[2], gp[3], n.alpha[F1]), n.alpha[F2]);
    q = lerp_fast(lerp_fast(gq[0], gq[1], n.alpha[F1]), lerp_fast(gq[2], gq[3], n.alpha[F1]), n.alpha[F2]);
}

__device__ __forceinline__ FaceCache empty_face_cache() {
    FaceCache c;
    c.cell = -2147483647 - 1;
    c.p0 = c.q0 = c.p1 = c.q1 = 0.0f;
    return c;
}

// 2 * (gradient_noise - 0.5) (generate.h:370-402) at the domain point moved
// by `distance` along axis A, with the gradient table in LDS.
template <int A>
__device__ __forceinline__ float gradient_noise_axis(const NoiseDomain &n, float distance, FaceCache &fc,
                                                     const float4 *grad) {
    const float x = n.d[A] + distance;
    const float fl = floorf(x);
    const int32_t i = static_cast<int32_t>(fl);
    const float w = x - fl;
    if (i != fc.cell) {
        if (i != fc.cell + 1) {
            noise_face<A>(n, i, grad, fc.p0, fc.q0);
        } else {
            fc.p0 = fc.p1;
            fc.q0 = fc.q1;
        }
        noise_face<A>(n, i + 1, grad, fc.p1, fc.q1);
        fc.cell = i;
    }
    // gradient_noise returns 0.5 * v + 0.5 and the caller subtracts 0.5;
    // the caller's scale carries the 0.5
    return lerp_fast(fmaf(w, fc.p0, fc.q0), fmaf(w - 1.0f, fc.p1, fc.q1), s_curve(w));
}

// rotate_by(value, axis, acos(c)) (generate.h:404-418, 477-480) with
// cos(acos c) = c and sin(acos c) = sqrt((1 - c)(1 + c)); the axis is the
// unnormalised cross(tangent, next_tangent) exactly as in the reference.
__device__ __forceinline__ V3 rotate_by_cos(V3 v, V3 a, float c) {
    const float omc = 1.0f - c;
    const float s = fast_sqrt(omc * (1.0f + c));
    const V3 ao = a * omc;
    return {fma3(v, v3(fmaf(a.x, ao.x, c), fmaf(a.x, ao.y, -(a.z * s)), fmaf(a.x, ao.z, a.y * s))),
            fma3(v, v3(fmaf(a.x, ao.y, a.z * s), fmaf(a.y, ao.y, c), fmaf(a.y, ao.z, -(a.x * s)))),
            fma3(v, v3(fmaf(a.x, ao.z, -(a.y * s)), fmaf(a.y, ao.z, a.x * s), fmaf(a.z, ao.z, c)))};
}

struct Args {
    // frame geometry (deformed array or the asset's rest array)
    const nxg_vec3 *positions;
    const nxg_vec3 *normals;
    const nxg_vec3 *rest_positions;
    const nxg_vec2 *texcoords;  // null when the asset has none
    const nxg_uint3 *triangles;
    const nxg_alias_entry *alias;
    const uint32_t *stencil;
    const nxg_vec3 *gradients;
    const float4 *guide_state;    // 2 float4 per guide
    const float4 *guide_offsets;  // guide-major, cvs float4 per guide
    const nxg_root_sample *explicit_roots;
    const float4 *attributes;  // optional per-root multipliers (NXG_ATTR_*)
    uint32_t strand_count, cvs, seed, triangle_count, guide_count, has_normals;
    uint32_t noise, octave_count;  // runtime form of the template switches (batched path)
    uint32_t vec3_points;          // positions only, 3 floats per point (motion samples)
    const uint32_t *order;         // processing order of the strands (null: index order)
    float support_scale, weight_power, rejection_cos, length_scale;
    float root_width, tip_width, radius_scale;
    float noise_mask, decorrelation_scale, preserve, inv_span;
    float amp[NXG_MAX_NOISE_OCTAVES];
    float freq[NXG_MAX_NOISE_OCTAVES];
    float min_freq;  // lowest octave frequency
    nxg_packed_point *points;
    nxg_root_sample *roots;
    nxg_vec2 *root_uvs;
    uint32_t *point_counts;
    uint32_t *stencil_indices;
    float *stencil_weights;
};

struct Blend {
    uint32_t idx[8];
    float w[8];
    float sum;
    float inv_sum;  // 1 / sum (fast path of the weighted mean)
    float length_scale;  // length_scale * length attribute
    uint32_t nearest;
};

struct Root {
    V3 position, normal;
    float u, v;
    uint32_t tri;
    float b1, b2;
};

// sample_root (generate.h:101-135)
// The alias-table pick of sample_root (generate.h:107-112)
__device__ __forceinline__ uint32_t sample_triangle(const nxg_alias_entry *alias, uint32_t seed,
                                                    uint32_t triangle_count, uint32_t strand) {
    const float choose = random01(seed, strand, 0u) * static_cast<float>(triangle_count);
    uint32_t bucket = static_cast<uint32_t>(choose);
    if (bucket >= triangle_count) bucket = triangle_count - 1u;
    const nxg_alias_entry e = alias[bucket];
    const uint32_t tri = random01(seed, strand, 1u) < e.probability ? bucket : e.alias;
    // validate_asset does not range-check alias targets; a corrupt entry
    // falls back to its own bucket instead of reading past the triangles
    return tri < triangle_count ? tri : bucket;
}

// sample_root's barycentric draw and surface interpolation on triangle tri
// (generate.h:113-135); the alias and the inverse-CDF samplers share it
__device__ __forceinline__ Root root_on_triangle(const Args &a, uint32_t strand, uint32_t tri) {
    const float r1 = random01(a.seed, strand, 2u);
    const float r2 = random01(a.seed, strand, 3u);
    const float s = sqrtf(r1);
    const float b0 = 1.0f - s;
    const float b1 = s * (1.0f - r2);
    const float b2 = s * r2;
    const nxg_uint3 t = a.triangles[tri];
    const V3 p0 = ld(a.positions[t.x]), p1 = ld(a.positions[t.y]), p2 = ld(a.positions[t.z]);
    Root r;
    r.position = p0 * b0 + p1 * b1 + p2 * b2;
    if (a.has_normals) {
        r.normal = normalize(ld(a.normals[t.x]) * b0 + ld(a.normals[t.y]) * b1 + ld(a.normals[t.z]) * b2);
    } else {
        r.normal = normalize(cross(p1 - p0, p2 - p0));
    }
    r.u = 0.0f;
    r.v = 0.0f;
    if (a.texcoords) {
        const nxg_vec2 u0 = a.texcoords[t.x], u1 = a.texcoords[t.y], u2 = a.texcoords[t.z];
        r.u = u0.x * b0 + u1.x * b1 + u2.x * b2;
        r.v = u0.y * b0 + u1.y * b1 + u2.y * b2;
    }
    r.tri = tri;
    r.b1 = b1;
    r.b2 = b2;
    return r;
}

__device__ __forceinline__ Root sample_root(const Args &a, uint32_t strand) {
    return root_on_triangle(a, strand, sample_triangle(a.alias, a.seed, a.triangle_count, strand));
}

// The opt-in inverse-CDF sampler's target for a strand: floor(r T / 2^32)
// with r = h0 2^32 + h1, h0 and h1 the hashes behind random01's dimensions
// 0 and 1; it lies in [0, T 2^32), and target / 2^32 is the strand's place
// along the triangles in units of one triangle's mean mass.
__device__ __forceinline__ uint64_t cdf_target(uint32_t seed, uint32_t strand, uint32_t triangle_count) {
    const uint64_t h0 = hash32(seed ^ hash32(strand + 0x9e3779b9u * 1u));
    const uint64_t h1 = hash32(seed ^ hash32(strand + 0x9e3779b9u * 2u));
    return h0 * triangle_count + ((h1 * triangle_count) >> 32);
}

// pow(x, max(power, 0)) of make_guide_blend_state (generate.h:236-238) for
// x in (0, 1].  Powers 0, 1 and 2 (2 is the default guide_weight_power,
// types.h:114) are the correctly rounded results 1, x and x * x; any other
// power calls the device powf.  A libm powf is within 1 ulp of these (glibc's
// powf(x, 2) differs from x * x on 0.07% of the floats in [2^-10, 1], by one
// ulp), and the branch is uniform over the launch.
__device__ __forceinline__ float blend_weight(float x, float power) {
    const float e = fmax_ref(power, 0.0f);
    if (e == 2.0f) return x * x;
    if (e == 1.0f) return x;
    if (e == 0.0f) return 1.0f;
    return powf(x, e);
}

// make_guide_blend_state (generate.h:214-244) over the per-guide state
__device__ __forceinline__ Blend make_blend(const Args &a, const Root &r) {
    Blend b;
    b.sum = 0.0f;
    b.nearest = NXG_INVALID_INDEX;
    float nearest_d2 = 3.402823466e+38f;
    const uint4 *row = reinterpret_cast<const uint4 *>(a.stencil + uint64_t{r.tri} * 8u);
    const uint4 s0 = row[0], s1 = row[1];
    const uint32_t slots[8] = {s0.x, s0.y, s0.z, s0.w, s1.x, s1.y, s1.z, s1.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        b.idx[j] = NXG_INVALID_INDEX;
        b.w[j] = 0.0f;
        const uint32_t gi = slots[j];
        // out-of-range stencil entries (not range-checked by validate_asset)
        // are treated like empty slots
        if (gi >= a.guide_count) continue;
        const float4 gs0 = a.guide_state[2u * gi];
        const float4 gs1 = a.guide_state[2u * gi + 1u];
        const V3 groot = v3(gs0.x, gs0.y, gs0.z);
        const float d2 = len2(r.position - groot);
        if (dot(r.normal, v3(gs1.x, gs1.y, gs1.z)) < a.rejection_cos) continue;
        if (d2 < nearest_d2) {
            nearest_d2 = d2;
            b.nearest = gi;
        }
        const float radius = gs0.w * a.support_scale;
        if (radius <= 0.0f || d2 >= radius * radius) continue;
        const float w = blend_weight(1.0f - sqrtf(d2) / radius, a.weight_power);
        b.idx[j] = gi;
        b.w[j] = w;
        b.sum += w;
    }
    b.inv_sum = 1.0f / b.sum;
    return b;
}

// The blend as the CV loop reads it: per stencil slot the byte offset of
// the guide's row in the offset table and the slot's weight, an empty slot
// reading guide 0's row with weight 0 (it adds +-0), so that the eight
// gathers of a CV are unconditional and independent; the nearest-guide
// fallback (no weight in reach) is one slot of weight 1 and scale 1, and a
// strand with no guide at all follows its normal.
struct BlendRows {
    uint32_t off[8];
    float w[8];
    float scale;         // 1 / sum, or 1 for the nearest guide
    float length_scale;  // length_scale * length attribute
    bool along_normal;   // no guide in reach and no nearest guide
};

__device__ __forceinline__ BlendRows blend_rows(const Args &a, const Blend &b) {
    BlendRows br;
    const uint32_t row = a.cvs * 16u;
    const bool weighted = b.sum > 1.0e-12f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const bool use = weighted ? b.idx[j] != NXG_INVALID_INDEX : (j == 0 && b.nearest != NXG_INVALID_INDEX);
        const uint32_t g = weighted ? b.idx[j] : b.nearest;
        br.off[j] = use ? g * row : 0u;
        br.w[j] = use ? (weighted ? b.w[j] : 1.0f) : 0.0f;
    }
    br.scale = weighted ? b.inv_sum : 1.0f;
    br.length_scale = b.length_scale;
    br.along_normal = !weighted && b.nearest == NXG_INVALID_INDEX;
    return br;
}

// evaluate_base_strand_cv (generate.h:246-267, 300-310) from the offset
// table, fast form (positions only): the weighted mean multiplies by 1/sum.
// At cv 0 every offset is exactly zero, so the first CV is the root.  The
// row base of the CV is wave-uniform, the slot offsets per lane.
__device__ __forceinline__ V3 base_cv(const Args &a, const Root &r, const BlendRows &b, uint32_t cv) {
    V3 off;
    if (!b.along_normal) {
        const uint8_t *rows = reinterpret_cast<const uint8_t *>(a.guide_offsets) + cv * 16u;
        float4 o[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = *reinterpret_cast<const float4 *>(rows + b.off[j]);
        off = v3(0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int j = 0; j < 8; ++j) off = fma_v(v3(o[j].x, o[j].y, o[j].z), b.w[j], off);
        off = off * b.scale;
    } else {
        const float t = static_cast<float>(cv) / static_cast<float>(a.cvs - 1u);
        off = r.normal * t;
    }
    return fma_v(off, b.length_scale, r.position);
}

// root_surface_u (generate.h:312-341)
__device__ __forceinline__ V3 root_surface_u(const Args &a, const Root &r) {
    const nxg_uint3 t = a.triangles[r.tri];
    V3 tangent = v3(0.0f, 0.0f, 0.0f);
    if (a.texcoords) {
        const V3 p0 = ld(a.positions[t.x]);
        const V3 e1 = ld(a.positions[t.y]) - p0;
        const V3 e2 = ld(a.positions[t.z]) - p0;
        const nxg_vec2 u0 = a.texcoords[t.x], u1 = a.texcoords[t.y], u2 = a.texcoords[t.z];
        const float d1x = u1.x - u0.x, d1y = u1.y - u0.y;
        const float d2x = u2.x - u0.x, d2y = u2.y - u0.y;
        const float det = d1x * d2y - d1y * d2x;
        if ((det < 0.0f ? -det : det) > 1.0e-12f) {
            tangent = (e1 * d2y - e2 * d1y) / det;
            tangent = tangent - r.normal * dot(tangent, r.normal);
        }
    }
    if (len2(tangent) > 1.0e-20f) return normalize(tangent);
    const V3 axis = r.normal.z > -0.999f && r.normal.z < 0.999f ? v3(0.0f, 0.0f, 1.0f)
                                                                : v3(0.0f, 1.0f, 0.0f);
    return normalize(cross(r.normal, axis));
}

struct Cursor {
    uint32_t cv;
    V3 cur, next, tangent, tn;
    float arc;
    float step;  // |base(cv) - base(cv - 1)|, added to arc when cv is reached
};

template <int OCT>
struct NoiseState {
    NoiseDomain dom;
    float freq[OCT];
    float mag[OCT];  // noise_mask * amplitude
    FaceCache face[OCT][3];
};

// evaluate_and_advance_noise_cursor (generate.h:459-515); OCT octaves share
// the transported frame and the arc length and their displacements add.
template <int OCT, bool RT = false>
__device__ __forceinline__ V3 advance(const Args &a, const Root &r, const BlendRows &b,
                                      NoiseState<OCT> &ns, Cursor &c, const float4 *grad) {
    if (c.cv > 0u) c.arc += c.step;
    V3 nt = c.tangent;
    if (c.cv + 1u < a.cvs) {
        const V3 seg = c.next - c.cur;
        const float seg2 = fma3(seg, seg);
        c.step = fast_sqrt(seg2);
        if (seg2 > 0.0f) nt = normalize_fast(seg);
    }
    const V3 axis = cross_fast(c.tangent, nt);
    if (fma3(axis, axis) > 0.0f) {
        const float cosine = clamp_ref(fma3(c.tangent, nt), -1.0f, 1.0f);
        c.tn = normalize_fast(rotate_by_cos(c.tn, axis, cosine));
    }
    const V3 normal = c.tn;
    const V3 binormal = cross_fast(normal, nt);
    const V3 tangent = cross_fast(binormal, normal);
    V3 point = c.cur;
    if (c.cv > 0u) {
        const float half_t = 0.5f * static_cast<float>(c.cv) * a.inv_span;
        float lx = 0.0f, ly = 0.0f, lz = 0.0f;
#pragma unroll
        for (int k = 0; k < OCT; ++k) {
            if (RT && k >= static_cast<int>(a.octave_count)) break;
            const float distance = c.arc * ns.freq[k];
            const float m = ns.mag[k] * half_t;  // noise_mask * amplitude * t / 2
            lx = fmaf(gradient_noise_axis<0>(ns.dom, distance, ns.face[k][0], grad), m, lx);
            ly = fmaf(gradient_noise_axis<1>(ns.dom, distance, ns.face[k][1], grad), m, ly);
            lz = fmaf(gradient_noise_axis<2>(ns.dom, distance, ns.face[k][2], grad), m, lz);
        }
        point = fma_v(tangent, lz, fma_v(binormal, ly, fma_v(normal, lx, point)));
    }
    c.tangent = nt;
    if (c.cv + 1u < a.cvs) {
        c.cur = c.next;
        if (c.cv + 2u < a.cvs) c.next = base_cv(a, r, b, c.cv + 2u);
    }
    ++c.cv;
    return point;
}

template <int OCT>
__device__ __forceinline__ Cursor make_cursor(const Args &a, const Root &r, const BlendRows &b, V3 surface_u) {
    Cursor c;
    c.cv = 0u;
    c.cur = base_cv(a, r, b, 0u);
    c.next = base_cv(a, r, b, 1u);
    c.step = 0.0f;
    c.tangent = r.normal;
    c.tn = surface_u;
    c.arc = 0.0f;
    return c;
}

struct Prep {
    const nxg_guide_record *guides;
    const nxg_vec3 *guide_cvs;
    const nxg_uint3 *triangles;
    const nxg_vec3 *d_positions, *d_normals, *d_guide_cvs, *frame_positions;
    uint32_t guide_count, triangle_count, cvs;
    float4 *state, *offsets;
};

// guide state and offset table: one lane per (guide, cv), i < guide_count * cvs
__device__ __forceinline__ void guide_prep_lane(const Prep &p, uint64_t i) {
    const nxg_guide_record *guides = p.guides;
    const nxg_vec3 *guide_cvs = p.guide_cvs, *d_positions = p.d_positions, *d_normals = p.d_normals;
    const nxg_vec3 *d_guide_cvs = p.d_guide_cvs, *frame_positions = p.frame_positions;
    const nxg_uint3 *triangles = p.triangles;
    const uint32_t triangle_count = p.triangle_count, cvs = p.cvs;
    float4 *state = p.state, *offsets = p.offsets;
    const uint32_t gi = static_cast<uint32_t>(i / cvs);
    const uint32_t cv = static_cast<uint32_t>(i % cvs);
    const nxg_guide_record g = guides[gi];
    const bool bound = g.triangle_index < triangle_count;
    // guide_root (generate.h:137-151)
    V3 root = ld(g.root_position);
    if (d_guide_cvs) {
        root = ld(d_guide_cvs[g.first_cv]);
    } else if (d_positions && bound) {
        const nxg_uint3 t = triangles[g.triangle_index];
        const float b1 = g.barycentric.x, b2 = g.barycentric.y, b0 = 1.0f - b1 - b2;
        root = ld(d_positions[t.x]) * b0 + ld(d_positions[t.y]) * b1 + ld(d_positions[t.z]) * b2;
    }
    // follow_deformed_guide_root (generate.h:174-187)
    const bool follow = !d_guide_cvs && d_positions && bound;
    const nxg_vec3 *cv_src = d_guide_cvs ? d_guide_cvs : guide_cvs;
    // sample_guide (generate.h:189-205) at t = cv / (cvs - 1)
    const float t = static_cast<float>(cv) / static_cast<float>(cvs - 1u);
    V3 sample;
    if (g.cv_count <= 1u) {
        sample = ld(cv_src[g.first_cv]);
    } else {
        const float x = clamp_ref(t, 0.0f, 1.0f) * static_cast<float>(g.cv_count - 1u);
        uint32_t k = static_cast<uint32_t>(x);
        if (k >= uint32_t{g.cv_count} - 1u) k = uint32_t{g.cv_count} - 2u;
        const float f = x - static_cast<float>(k);
        sample = ld(cv_src[g.first_cv + k]) * (1.0f - f) + ld(cv_src[g.first_cv + k + 1u]) * f;
    }
    if (follow) sample = sample + root - ld(g.root_position);
    const V3 off = sample - root;
    offsets[i] = make_float4(off.x, off.y, off.z, 0.0f);
    if (cv == 0u) {
        // guide_normal (generate.h:153-172)
        V3 n = ld(g.root_normal);
        if (bound && (d_positions || d_normals)) {
            const nxg_uint3 tr = triangles[g.triangle_index];
            const float b1 = g.barycentric.x, b2 = g.barycentric.y, b0 = 1.0f - b1 - b2;
            if (d_normals) {
                n = normalize(ld(d_normals[tr.x]) * b0 + ld(d_normals[tr.y]) * b1 + ld(d_normals[tr.z]) * b2);
            } else {
                const V3 p0 = ld(frame_positions[tr.x]);
                n = normalize(cross(ld(frame_positions[tr.y]) - p0, ld(frame_positions[tr.z]) - p0));
            }
        }
        state[2u * gi] = make_float4(root.x, root.y, root.z, g.support_radius);
        state[2u * gi + 1u] = make_float4(n.x, n.y, n.z, 0.0f);
    }
}

__global__ void __launch_bounds__(256) guide_prep_kernel(Prep p) {
    const uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x;
    if (i < uint64_t{p.guide_count} * p.cvs) guide_prep_lane(p, i);
}

// Writes the wave's staged chunk [cv0, cv0 + n) for its 64 strands as
// contiguous runs of n float4 per strand.
// Stores a staged chunk of n CVs of the wavefront's 64 strands: lane L of
// pass `it` writes flat element it*64 + L of the strand-major chunk, so each
// store instruction covers consecutive float4s.  N is n when it is known at
// compile time (full chunks), else 0.
template <uint32_t N>
__device__ __forceinline__ void flush_chunk(const Args &a, const float *stage, const uint32_t *ids,
                                            uint64_t wave_first, uint32_t cv0, uint32_t n, uint32_t lane) {
    // f / n for f < 64 * n <= 1024 as floor((f + 0.5) / n) in float: the
    // half keeps exact quotients away from rounding below an integer
    const float inv_n = 1.0f / static_cast<float>(N ? N : n);
#pragma unroll 1
    for (uint32_t it = 0; it < (N ? N : n); ++it) {
        const uint32_t f = it * kWave + lane;
        const uint32_t s = static_cast<uint32_t>((static_cast<float>(f) + 0.5f) * inv_n);
        const uint32_t k = f - s * (N ? N : n);
        if (wave_first + s < a.strand_count) {
            const uint64_t strand = ids[s];
            float4 p;
            p.x = stage[(k * 4u + 0u) * kStageStride + s];
            p.y = stage[(k * 4u + 1u) * kStageStride + s];
            p.z = stage[(k * 4u + 2u) * kStageStride + s];
            p.w = stage[(k * 4u + 3u) * kStageStride + s];
            reinterpret_cast<float4 *>(a.points)[strand * a.cvs + cv0 + k] = p;
        }
    }
}

// The same chunk as positions only (motion-sample output): 3 floats per
// point, stored as consecutive dwords of the strand-major chunk.
__device__ __forceinline__ void flush_chunk_vec3(const Args &a, const float *stage, const uint32_t *ids,
                                                 uint64_t wave_first, uint32_t cv0, uint32_t n, uint32_t lane) {
    float *out = reinterpret_cast<float *>(a.points);
    const uint32_t per = 3u * n;
    // f / per for f < 64 * per, per <= 48, as in flush_chunk
    const float inv_per = 1.0f / static_cast<float>(per);
#pragma unroll 1
    for (uint32_t it = 0; it < per; ++it) {
        const uint32_t f = it * kWave + lane;
        const uint32_t s = static_cast<uint32_t>((static_cast<float>(f) + 0.5f) * inv_per);
        const uint32_t rem = f - s * per;
        const uint32_t k = rem / 3u, c = rem - 3u * k;
        if (wave_first + s < a.strand_count)
            out[(uint64_t{ids[s]} * a.cvs + cv0 + k) * 3u + c] = stage[(k * 4u + c) * kStageStride + s];
    }
}

// Orders this wavefront's LDS traffic: every lane's earlier LDS writes are
// visible to every lane's later reads (the staging area is per wavefront,
// so no workgroup barrier is needed and wavefronts may work on different
// descriptions).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wavefront generates strands [wave_first, wave_first + 64) of the
// description `a`: stage is this wavefront's CHUNK-CV staging area, grad the
// description's gradient table in LDS.  NOISE/OCT fix the noise path at
// compile time; with RT the octave count comes from a.octave_count (<= OCT).
template <bool NOISE, int OCT, bool RT, int CHUNK>
__device__ __forceinline__ void strand_tile(const Args &a, uint64_t wave_first, uint32_t lane, float *stage,
                                            uint32_t *ids, const float4 *grad) {
    // position wave_first + lane of the processing order; lanes past the end
    // compute the last position's strand and store nothing, so every lane
    // runs the same CV loop
    const uint64_t slot64 = wave_first + lane;
    const bool live = slot64 < a.strand_count;
    const uint32_t slot = live ? static_cast<uint32_t>(slot64) : a.strand_count - 1u;
    const uint32_t strand = a.order ? a.order[slot] : slot;
    ids[lane] = strand;  // read by the chunk flushes, after a wavefront sync

    Root r;
    if (a.explicit_roots) {
        const nxg_root_sample e = a.explicit_roots[strand];
        r.position = ld(e.position);
        r.normal = ld(e.normal);
        r.u = e.uv.x;
        r.v = e.uv.y;
        // explicit roots are the caller's contract (asset.cpp:468-483 checks
        // them on the host); an out-of-range triangle is clamped so that the
        // strand reads stay inside the asset
        r.tri = e.triangle_index < a.triangle_count ? e.triangle_index : a.triangle_count - 1u;
        r.b1 = e.barycentric.x;
        r.b2 = e.barycentric.y;
    } else {
        r = sample_root(a, strand);
    }
    Blend b = make_blend(a, r);
    // per-root attribute multipliers; 1.0 (no attributes) is exact
    const float4 attr = a.attributes ? a.attributes[strand] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    b.length_scale = a.length_scale * attr.x;
    if (live) {
        if (a.roots) {
            nxg_root_sample o;
            o.position = {r.position.x, r.position.y, r.position.z};
            o.normal = {r.normal.x, r.normal.y, r.normal.z};
            o.uv = {r.u, r.v};
            o.triangle_index = r.tri;
            o.barycentric = {r.b1, r.b2};
            o.surface_face_id = r.tri;
            a.roots[strand] = o;
        }
        if (a.root_uvs) a.root_uvs[strand] = {r.u, r.v};
        // every count is cvs, so writing them by processing position fills the
        // same array with coalesced stores whatever the order
        if (a.point_counts) a.point_counts[slot] = a.cvs;
        if (a.stencil_indices) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a.stencil_indices[uint64_t{strand} * 8u + j] = b.idx[j];
                a.stencil_weights[uint64_t{strand} * 8u + j] = b.w[j];
            }
        }
    }

    const BlendRows br = blend_rows(a, b);
    NoiseState<OCT> ns;
    V3 surface_u = v3(0.0f, 0.0f, 0.0f);
    float base_length = 0.0f;
    if constexpr (NOISE) {
        surface_u = root_surface_u(a, r);
        // base_strand_length (generate.h:420-433) sets the effective
        // frequency max(0.5 / length, f) and the preserve-length target.
        // Without preserve-length the length is needed only when 0.5 / length
        // could reach the lowest octave frequency: the chord from the root to
        // the last base CV bounds the polyline's length from below, so a
        // chord past 0.5 / f_min (with a margin for the float sums) fixes
        // every frequency at f and the pass over the base CVs is skipped.
        const V3 tip = base_cv(a, r, br, a.cvs - 1u);
        const float chord = sqrtf(len2(tip - r.position));
        if (a.preserve > 0.001f || !(a.min_freq > 0.0f) || !(chord * a.min_freq > 0.5f * 1.001f)) {
            V3 prev = base_cv(a, r, br, 0u);
            for (uint32_t cv = 1u; cv < a.cvs; ++cv) {
                const V3 c = base_cv(a, r, br, cv);
                base_length += sqrtf(len2(c - prev));
                prev = c;
            }
        }
#pragma unroll
        for (int k = 0; k < OCT; ++k) {
            ns.freq[k] = base_length > 0.0f ? fmax_ref(0.5f / base_length, a.freq[k]) : a.freq[k];
            ns.mag[k] = a.noise_mask * a.amp[k] * attr.z;
        }
        // rest_root_position (generate.h:343-351), noise domain (generate.h:548-552)
        const nxg_uint3 t = a.triangles[r.tri];
        const float b0 = 1.0f - r.b1 - r.b2;
        const V3 rest = ld(a.rest_positions[t.x]) * b0 + ld(a.rest_positions[t.y]) * r.b1 +
                        ld(a.rest_positions[t.z]) * r.b2;
        ns.dom = make_domain((rest + v3(0.419276f, 0.184247f, 0.805721f)) * a.decorrelation_scale);
    }

    // With preserve-length on, a first pass runs the same cursor to measure
    // the noisy length (noisy_strand_length, generate.h:517-532) and the
    // second pass writes the scaled points; both passes share one copy of
    // the cursor code.  The pass count is uniform across the launch.
    const uint32_t passes = (NOISE && a.preserve > 0.001f) ? 2u : 1u;
    float preserve_scale = 1.0f;
    for (uint32_t pass = 0; pass < passes; ++pass) {
        const bool emit = pass + 1u == passes;
        Cursor cur;
        if constexpr (NOISE) {
            cur = make_cursor<OCT>(a, r, br, surface_u);
#pragma unroll
            for (int k = 0; k < OCT; ++k)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) ns.face[k][ax] = empty_face_cache();
        }
        V3 last = v3(0.0f, 0.0f, 0.0f);
        float noisy = 0.0f;
        for (uint32_t cv0 = 0; cv0 < a.cvs; cv0 += CHUNK) {
            const uint32_t n = a.cvs - cv0 < CHUNK ? a.cvs - cv0 : CHUNK;
            // evaluate_strand_width's taper (generate.h:577-581) of the
            // chunk's CVs, one per lane, read back per CV with readlane
            float taper = 0.0f;
            if (emit) {
                const float t = static_cast<float>(cv0 + lane) / static_cast<float>(a.cvs - 1u);
                taper = a.root_width * (1.0f - t) + a.tip_width * t;
            }
#pragma unroll 1
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t cv = cv0 + k;
                V3 p;
                if constexpr (NOISE) {
                    p = advance<OCT, RT>(a, r, br, ns, cur, grad);
                } else {
                    p = base_cv(a, r, br, cv);
                }
                if (!emit) {
                    if (cv > 0u) noisy += sqrtf(len2(p - last));
                    last = p;
                    continue;
                }
                if (preserve_scale != 1.0f) p = r.position + (p - r.position) * preserve_scale;
                // evaluate_strand_width (generate.h:577-581) and the packed radius
                const float width =
                    __int_as_float(__builtin_amdgcn_readlane(__float_as_int(taper), static_cast<int>(k))) * attr.y;
                stage[(k * 4u + 0u) * kStageStride + lane] = p.x;
                stage[(k * 4u + 1u) * kStageStride + lane] = p.y;
                stage[(k * 4u + 2u) * kStageStride + lane] = p.z;
                stage[(k * 4u + 3u) * kStageStride + lane] = 0.5f * width * a.radius_scale;
            }
            if (emit) {
                wave_sync();
                if (a.vec3_points) flush_chunk_vec3(a, stage, ids, wave_first, cv0, n, lane);
                else flush_chunk<0>(a, stage, ids, wave_first, cv0, n, lane);
                wave_sync();
            }
        }
        if (!emit && noisy > 0.0f) {  // generate.h:553-563
            const float target = base_length * a.preserve + noisy * (1.0f - a.preserve);
            const float diff = noisy - target;
            if ((diff < 0.0f ? -diff : diff) >= 0.0001f) preserve_scale = target / noisy;
        }
    }
}


// Up to two octaves fit 168 VGPRs, so that three wavefronts share each SIMD.
template <bool NOISE, int OCT>
__global__ void __launch_bounds__(256, (OCT <= 2 ? 3 : 2)) strand_kernel(Args a) {
    __shared__ float4 grad[NXG_NOISE_GRADIENT_COUNT];
    __shared__ float stage_all[256 / kWave][kChunk * 4 * kStageStride];
    __shared__ uint32_t ids_all[256 / kWave][kWave];
    for (uint32_t i = threadIdx.x; i < NXG_NOISE_GRADIENT_COUNT; i += blockDim.x) {
        const nxg_vec3 g = a.gradients[i];
        grad[i] = make_float4(g.x, g.y, g.z, 0.0f);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t wave = threadIdx.x / kWave;
    const uint64_t wave_first = uint64_t{blockIdx.x} * blockDim.x + wave * kWave;
    strand_tile<NOISE, OCT, false, kChunk>(a, wave_first, lane, stage_all[wave], ids_all[wave], grad);
}

// ---- batched descriptions ------------------------------------------------

// Guide preparation of every item in one launch: item k owns the 256-lane
// blocks [block_end[k-1], block_end[k]).  Block 0 also zeroes the tile
// counter of the strand launch that follows on the same stream.
struct BatchPrepArgs {
    Prep items[NXG_BATCH_MAX_ITEMS];
    uint32_t block_end[NXG_BATCH_MAX_ITEMS];
    uint32_t count;
    uint32_t *counter;
};

__global__ void __launch_bounds__(256) batch_prep_kernel(BatchPrepArgs b) {
    if (blockIdx.x == 0u && threadIdx.x == 0u) *b.counter = 0u;
    uint32_t item = 0u;
    while (item + 1u < b.count && blockIdx.x >= b.block_end[item]) ++item;
    const uint32_t first = item ? b.block_end[item - 1u] : 0u;
    const Prep &p = b.items[item];
    const uint64_t i = uint64_t{blockIdx.x - first} * 256u + threadIdx.x;
    if (i < uint64_t{p.guide_count} * p.cvs) guide_prep_lane(p, i);
}

struct BatchArgs {
    Args items[NXG_BATCH_MAX_ITEMS];
    uint32_t tile_end[NXG_BATCH_MAX_ITEMS];  // inclusive scan of 64-strand tiles, processing order
    uint32_t order[NXG_BATCH_MAX_ITEMS];     // item processed at each position
    uint32_t count;
    uint32_t total_tiles;
    uint32_t tiles_per_claim;
    uint32_t *counter;                       // zeroed by batch_prep_kernel
};

// Persistent: every wavefront claims runs of tiles_per_claim 64-strand
// tiles from one atomic counter until the work list is exhausted (every
// wavefront leaves once its claim starts past the last tile); each tile's
// description comes from a binary search of the tile scan held in LDS.
template <int WAVES, int CHUNK, int OCTMAX>
__global__ void __launch_bounds__(WAVES * 64, (OCTMAX <= 2 && CHUNK < 16 ? 3 : 2)) batch_kernel(BatchArgs b) {
    __shared__ float4 grad_all[WAVES][NXG_NOISE_GRADIENT_COUNT];
    __shared__ float stage_all[WAVES][CHUNK * 4 * kStageStride];
    __shared__ uint32_t ids_all[WAVES][kWave];
    __shared__ uint32_t scan[NXG_BATCH_MAX_ITEMS];
    if (threadIdx.x < b.count) scan[threadIdx.x] = b.tile_end[threadIdx.x];
    __syncthreads();
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    float4 *grad = grad_all[wave];
    uint32_t loaded = 0xffffffffu;
    for (;;) {
        uint32_t claim = 0u;
        if (lane == 0u) claim = atomicAdd(b.counter, 1u);
        const uint32_t first_tile = __builtin_amdgcn_readfirstlane(__shfl(claim, 0, kWave)) * b.tiles_per_claim;
        if (first_tile >= b.total_tiles) break;
        const uint32_t last_tile = min(first_tile + b.tiles_per_claim, b.total_tiles);
        for (uint32_t tile = first_tile; tile < last_tile; ++tile) {
            uint32_t lo = 0u, hi = b.count - 1u;  // first position whose scan exceeds tile
            while (lo < hi) {
                const uint32_t mid = (lo + hi) / 2u;
                if (scan[mid] > tile) hi = mid; else lo = mid + 1u;
            }
            lo = __builtin_amdgcn_readfirstlane(lo);
            const uint32_t item = b.order[lo];
            const uint32_t item_first = lo ? scan[lo - 1u] : 0u;
            const Args &a = b.items[item];
            if (item != loaded) {
#pragma unroll  // four loads in flight, not one round trip each
                for (uint32_t i = lane; i < NXG_NOISE_GRADIENT_COUNT; i += kWave) {
                    const nxg_vec3 g = a.gradients[i];
                    grad[i] = make_float4(g.x, g.y, g.z, 0.0f);
                }
                wave_sync();
                loaded = item;
            }
            const uint64_t wave_first = uint64_t{tile - item_first} * kWave;
            if (a.noise) strand_tile<true, OCTMAX, true, CHUNK>(a, wave_first, lane, stage_all[wave], ids_all[wave], grad);
            else strand_tile<false, 1, false, CHUNK>(a, wave_first, lane, stage_all[wave], ids_all[wave], grad);
            wave_sync();
        }
    }
}


// ---- strand order ----------------------------------------------------------
// Strands are independent, so they may run in any order; grouping them by
// root triangle makes the 64 strands of a wavefront share stencil guides,
// so their guide-offset reads hit the same cache lines (and their root and
// noise-lattice gathers cluster too).  A counting sort into kOrderBins bins
// of 2^shift consecutive triangles, with no device-wide atomics: each tile
// of kOrderTile strands histograms its bins in LDS (stored tile-major), the
// counts are scanned in bin-major order, and each tile places its strands
// from its LDS copy of the starts.  The order within a bin depends on LDS
// atomic timing, which changes no output.

struct OrderArgs {
    const nxg_alias_entry *alias;
    const nxg_root_sample *roots;  // roots already sampled (their triangles), or null: the alias draw
    uint32_t seed, triangle_count, strand_count, cvs;
    uint32_t shift, tiles;
    uint32_t *counts, *partials, *order;
    uint32_t cdf_key;  // sort by the inverse-CDF target / 2^32 instead (monotone in the triangle drawn)
    uint32_t *keys;    // per strand: bin << kOrderRankBits | rank within its tile's bin
};

__device__ __forceinline__ uint32_t order_triangle(const OrderArgs &o, uint32_t s) {
    if (o.cdf_key) return static_cast<uint32_t>(cdf_target(o.seed, s, o.triangle_count) >> 32);
    if (o.roots) {
        const uint32_t t = o.roots[s].triangle_index;
        return t < o.triangle_count ? t : o.triangle_count - 1u;
    }
    return sample_triangle(o.alias, o.seed, o.triangle_count, s);
}

// Every counting-sort tile kernel runs kOrderThreads threads, each taking
// kOrderTile / kOrderThreads strands; a thread's triangle draws (alias-table
// gathers) are all issued before its first histogram update.  A thread past
// the last strand draws the last strand again and adds nothing.
constexpr uint32_t kOrderThreads = 1024u;
constexpr uint32_t kOrderPerThread = kOrderTile / kOrderThreads;
static_assert(kOrderTile % kOrderThreads == 0u && kOrderBins % kOrderThreads == 0u, "order tile shape");

__device__ __forceinline__ void order_bins(const OrderArgs &o, uint32_t first, uint32_t (&bin)[kOrderPerThread]) {
#pragma unroll
    for (uint32_t j = 0; j < kOrderPerThread; ++j) {
        const uint32_t s = min(first + threadIdx.x + j * kOrderThreads, o.strand_count - 1u);
        bin[j] = order_triangle(o, s) >> o.shift;
    }
}

// A strand's key packs its bin and its rank among its tile's strands of
// that bin (the histogram count its update returned), so the scatter needs
// neither the triangle draw again nor atomics.
constexpr uint32_t kOrderRankBits = 13u;
static_assert(kOrderTile <= (1u << kOrderRankBits) && kOrderBins <= (1u << (32u - kOrderRankBits)), "order keys");

__device__ __forceinline__ void order_count_tile(const OrderArgs &o, uint32_t tile) {
    __shared__ uint32_t hist[kOrderBins];
#pragma unroll
    for (uint32_t j = 0; j < kOrderBins / kOrderThreads; ++j) hist[threadIdx.x + j * kOrderThreads] = 0u;
    __syncthreads();
    const uint32_t first = tile * kOrderTile;
    uint32_t bin[kOrderPerThread];
    order_bins(o, first, bin);
#pragma unroll
    for (uint32_t j = 0; j < kOrderPerThread; ++j) {
        const uint32_t s = first + threadIdx.x + j * kOrderThreads;
        const bool live = s < o.strand_count;
        const uint32_t rank = atomicAdd(&hist[bin[j]], live ? 1u : 0u);
        if (live) o.keys[s] = bin[j] << kOrderRankBits | rank;
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kOrderBins / kOrderThreads; ++j) {
        const uint32_t b = threadIdx.x + j * kOrderThreads;
        o.counts[tile * kOrderBins + b] = hist[b];
    }
}

__device__ __forceinline__ void order_scatter_tile(const OrderArgs &o, uint32_t tile) {
    __shared__ uint32_t start[kOrderBins];
    const uint32_t first = tile * kOrderTile;
    uint32_t key[kOrderPerThread];  // in flight with the starts' loads below
#pragma unroll
    for (uint32_t j = 0; j < kOrderPerThread; ++j)
        key[j] = o.keys[min(first + threadIdx.x + j * kOrderThreads, o.strand_count - 1u)];
    // the scan left each count's exclusive prefix within its kScanBlock
    // block of the bin-major order, and each block's exclusive prefix in partials
#pragma unroll
    for (uint32_t j = 0; j < kOrderBins / kOrderThreads; ++j) {
        const uint32_t b = threadIdx.x + j * kOrderThreads;
        start[b] = o.counts[tile * kOrderBins + b] + o.partials[(b * o.tiles + tile) / kScanBlock];
    }
    __syncthreads();
#pragma unroll
    for (uint32_t j = 0; j < kOrderPerThread; ++j) {
        const uint32_t s = first + threadIdx.x + j * kOrderThreads;
        if (s < o.strand_count)
            o.order[start[key[j] >> kOrderRankBits] + (key[j] & ((1u << kOrderRankBits) - 1u))] = s;
    }
}

// Exclusive scan of one value per thread of a 256-thread block.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t *wave_totals, uint32_t &total) {
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    uint32_t inc = v;
#pragma unroll
    for (uint32_t d = 1; d < static_cast<uint32_t>(kWave); d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, kWave);
        if (lane >= d) inc += y;
    }
    if (lane == kWave - 1) wave_totals[wave] = inc;
    __syncthreads();
    uint32_t before = 0u;
    total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < 256u / kWave; ++w) {
        const uint32_t t = wave_totals[w];
        if (w < wave) before += t;
        total += t;
    }
    __syncthreads();
    return before + inc - v;
}

// Exclusive scan of the block totals in place, 256 at a time (one block).
__device__ __forceinline__ void scan_partials(uint32_t *partials, uint32_t count) {
    __shared__ uint32_t wave_totals[256 / kWave];
    uint32_t carry = 0u;
    for (uint32_t first = 0; first < count; first += 256u) {
        const uint32_t i = first + threadIdx.x;
        const uint32_t v = i < count ? partials[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, wave_totals, total);
        if (i < count) partials[i] = carry + ex;
        carry += total;
    }
}

__device__ __forceinline__ uint32_t order_count_values(const OrderArgs &o) { return kOrderBins * o.tiles; }
__device__ __forceinline__ uint32_t order_scan_blocks(const OrderArgs &o) {
    return (order_count_values(o) + kScanBlock - 1u) / kScanBlock;
}

__global__ void __launch_bounds__(kOrderThreads) order_count_kernel(OrderArgs o) { order_count_tile(o, blockIdx.x); }
__global__ void __launch_bounds__(kOrderThreads) order_scatter_kernel(OrderArgs o) {
    order_scatter_tile(o, blockIdx.x);
}
// The counts are stored tile-major (bin b of tile t at t * kOrderBins + b, so
// each tile writes its histogram and reads its starts as one run) and
// scanned in bin-major order (element e = b * tiles + t): block `block` of
// kScanBlock elements, exclusive within the block, and the block's total.
__device__ __forceinline__ void scan_order_block(const OrderArgs &o, uint32_t block) {
    __shared__ uint32_t wave_totals[256 / kWave];
    const uint32_t n = order_count_values(o);
    const uint32_t base = block * kScanBlock + threadIdx.x * 4u;
    uint32_t b = base / o.tiles, t = base - b * o.tiles;  // element base, then along the bin-major order
    uint32_t at[4], v[4], sum = 0u;
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        at[j] = t * kOrderBins + b;
        v[j] = base + j < n ? o.counts[at[j]] : 0u;
        sum += v[j];
        if (++t == o.tiles) {
            t = 0u;
            ++b;
        }
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, wave_totals, total);
#pragma unroll
    for (uint32_t j = 0; j < 4u; ++j) {
        if (base + j < n) o.counts[at[j]] = run;
        run += v[j];
    }
    if (threadIdx.x == 0u) o.partials[block] = total;
}

__global__ void __launch_bounds__(256) order_scan_blocks_kernel(OrderArgs o) { scan_order_block(o, blockIdx.x); }
__global__ void __launch_bounds__(256) order_scan_partials_kernel(OrderArgs o) {
    scan_partials(o.partials, order_scan_blocks(o));
}

// The same four steps over every item of a batch: block ranges per item
// come from inclusive scans of tiles and of scan blocks.
struct BatchOrderArgs {
    OrderArgs items[NXG_BATCH_MAX_ITEMS];
    uint32_t tile_end[NXG_BATCH_MAX_ITEMS];
    uint32_t scan_end[NXG_BATCH_MAX_ITEMS];
    uint32_t count;
};

__device__ __forceinline__ uint32_t batch_item_of(const uint32_t *end, uint32_t count, uint32_t block,
                                                  uint32_t &first) {
    uint32_t item = 0u;
    while (item + 1u < count && block >= end[item]) ++item;
    first = item ? end[item - 1u] : 0u;
    return item;
}

__global__ void __launch_bounds__(kOrderThreads) batch_order_count_kernel(BatchOrderArgs b) {
    uint32_t first;
    const uint32_t k = batch_item_of(b.tile_end, b.count, blockIdx.x, first);
    order_count_tile(b.items[k], blockIdx.x - first);
}
__global__ void __launch_bounds__(kOrderThreads) batch_order_scatter_kernel(BatchOrderArgs b) {
    uint32_t first;
    const uint32_t k = batch_item_of(b.tile_end, b.count, blockIdx.x, first);
    order_scatter_tile(b.items[k], blockIdx.x - first);
}
__global__ void __launch_bounds__(256) batch_order_scan_blocks_kernel(BatchOrderArgs b) {
    uint32_t first;
    const uint32_t k = batch_item_of(b.scan_end, b.count, blockIdx.x, first);
    scan_order_block(b.items[k], blockIdx.x - first);
}
__global__ void __launch_bounds__(256) batch_order_scan_partials_kernel(BatchOrderArgs b) {
    scan_partials(b.items[blockIdx.x].partials, order_scan_blocks(b.items[blockIdx.x]));
}

// ---- float expression runtime ------------------------------------------

// gradient_noise (generate.h:370-402) operation for operation, without
// contraction: the expression runtime's noise3 is bit-identical to it.
__device__ float gradient_noise_exact(float sx, float sy, float sz, const float4 *grad) {
    const float fx = floorf(sx), fy = floorf(sy), fz = floorf(sz);
    const int32_t ix = static_cast<int32_t>(fx), iy = static_cast<int32_t>(fy), iz = static_cast<int32_t>(fz);
    const float wx = sx - fx, wy = sy - fy, wz = sz - fz;
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        const uint32_t seed = static_cast<uint32_t>(ix + ox) * kLcgA2 + static_cast<uint32_t>(iy + oy) * kLcgA +
                              static_cast<uint32_t>(iz + oz) + kLcgK;
        const float4 g = grad[temper_index(seed)];
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

// A texel map's point sample (the generator's texel op): face
// clamp(floor(f)), column clamp(floor(u R)), row clamp(floor(v R)).
__device__ __forceinline__ float texel_sample(const nxg_texel_map &m, float f, float u, float v) {
    const float R = static_cast<float>(m.resolution);
    const float ff = clamp_std(floorf(f), 0.0f, static_cast<float>(m.face_count - 1u));
    const float fu = clamp_std(floorf(u * R), 0.0f, R - 1.0f);
    const float fv = clamp_std(floorf(v * R), 0.0f, R - 1.0f);
    return m.texels[(static_cast<uint64_t>(ff) * m.resolution + static_cast<uint32_t>(fv)) * m.resolution +
                    static_cast<uint32_t>(fu)];
}

struct AttrArgs {
    const nxg_vec3 *positions;
    const nxg_vec3 *normals;
    const nxg_vec2 *texcoords;
    const nxg_uint3 *triangles;
    const nxg_alias_entry *alias;
    const nxg_vec3 *gradients;
    const uint8_t *programs;  // packed set image
    nxg_texel_map maps[NXG_EXPR_MAX_TEXEL_MAPS];
    uint32_t map_count, strand_count, seed, triangle_count, has_normals, program_count;
    uint32_t program_offset[NXG_EXPR_MAX_PROGRAMS];
    float4 *attributes;
    uint32_t *error_word;
    // verification modes (nxg_verify_attributes)
    uint32_t mode;          // kAttrAll, kAttrCheck or kAttrFallback
    uint32_t stride;        // kAttrCheck: roots s with s % stride == stride / 2
    float threshold;        // kAttrCheck: relative error bound
    uint32_t *tile_flags;   // one word per 64-root tile
    uint32_t *stats;        // [0] tiles re-evaluated, [1] samples checked
    uint32_t split;         // one program per grid row (blockIdx.y), only its attribute written
};

enum AttrMode : uint32_t { kAttrAll = 0u, kAttrCheck = 1u, kAttrFallback = 2u };

// One lane per root.  The program is the same for every lane, so opcode
// dispatch is wave-uniform (scalar branches, instructions fetched with scalar
// loads).  With REG the value slots are a kRegSlots-entry register array,
// indexed with the wave-uniform slot numbers (M0-relative register moves);
// otherwise they live in LDS at values[slot * blockDim + lane].
constexpr uint32_t kRegSlots = 32u;

template <bool REG>
__global__ void __launch_bounds__(256) attribute_kernel(AttrArgs a) {
    extern __shared__ float lds[];
    float4 *grad = reinterpret_cast<float4 *>(lds);
    float *values = lds + 4u * NXG_NOISE_GRADIENT_COUNT;
    for (uint32_t i = threadIdx.x; i < NXG_NOISE_GRADIENT_COUNT; i += blockDim.x) {
        const nxg_vec3 g = a.gradients[i];
        grad[i] = make_float4(g.x, g.y, g.z, 0.0f);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x, B = blockDim.x;
    uint64_t strand64 = uint64_t{blockIdx.x} * B + lane;
    if (a.mode == kAttrCheck) strand64 = strand64 * a.stride + a.stride / 2u;
    const bool live = strand64 < a.strand_count;
    const uint32_t strand = live ? static_cast<uint32_t>(strand64) : a.strand_count - 1u;
    // fallback: a wavefront covers exactly one 64-root tile (B is a multiple
    // of 64), so this exit is uniform per wavefront
    if (a.mode == kAttrFallback && a.tile_flags[strand / 64u] == 0u) return;

    Args ra{};
    ra.positions = a.positions;
    ra.normals = a.normals;
    ra.texcoords = a.texcoords;
    ra.triangles = a.triangles;
    ra.alias = a.alias;
    ra.seed = a.seed;
    ra.triangle_count = a.triangle_count;
    ra.has_normals = a.has_normals;
    const Root r = sample_root(ra, strand);
    const float vars[NXG_ROOT_VAR_COUNT] = {r.position.x, r.position.y, r.position.z, r.normal.x,
                                            r.normal.y, r.normal.z, r.u, r.v, static_cast<float>(r.tri),
                                            r.b1, r.b2, static_cast<float>(strand)};
    const float face_seed = hash_to_unit(hash_step(hash_step(hash_step(0u, 0.0f), 0.0f), static_cast<float>(r.tri)));

    // the root variables live in their reserved slots for the whole set
    // value slots: a plain local array (kept in registers: a struct around it
    // would send it to scratch) or the LDS rows
    float regs[kRegSlots];
#define SLOT_GET(x) (REG ? regs[__builtin_amdgcn_readfirstlane(x) & (kRegSlots - 1u)] : values[(x) * B + lane])
#define SLOT_SET(x, v)                                                        \
    do {                                                                      \
        if constexpr (REG) regs[__builtin_amdgcn_readfirstlane(x) & (kRegSlots - 1u)] = (v); \
        else values[(x) * B + lane] = (v);                                    \
    } while (0)
#pragma unroll
    for (uint32_t q = 0; q < NXG_ROOT_VAR_COUNT; ++q) SLOT_SET(q, vars[q]);

    const ExprContext ctx{r.u, r.v, face_seed, 0.0f, 0u, false};
    float result[4] = {1.0f, 1.0f, 1.0f, 1.0f};
    // the reference rejects a non-finite context (xgen_expression.cpp:991-994)
    bool bad = !isfinite(r.u) || !isfinite(r.v);
    // the programs of this launch row: all of them, or with `split` the one
    // of blockIdx.y (the verification passes run few roots; one program per
    // wavefront shortens each wavefront's interpreter pass)
    const uint32_t k_first = a.split ? blockIdx.y : 0u, k_end = a.split ? blockIdx.y + 1u : a.program_count;
    for (uint32_t k = k_first; k < k_end; ++k) {
        // the image is read-only for the whole launch: reading it through
        // the constant address space lets every (wave-uniform) fetch be a
        // scalar load
        const Image image = (Image)a.programs;
        ImageOf(PackedProgram) prog = (ImageOf(PackedProgram))(image + a.program_offset[k]);
        ImageOf(PackedInstruction) ins = (ImageOf(PackedInstruction))(image + prog->instructions_offset);
        ImageOf(uint32_t) operand_slots = (ImageOf(uint32_t))(image + prog->operands_offset);
        ImageOf(nxg_expr_ramp) ramps = (ImageOf(nxg_expr_ramp))(image + prog->ramps_offset);
        ImageOf(nxg_expr_ramp_point) points = (ImageOf(nxg_expr_ramp_point))(image + prog->ramp_points_offset);
        const uint32_t n = prog->instruction_count;
        // inputs are not dispatched: the root variables the program reads are
        // checked once here (an input is an error when its value is not finite)
        const uint32_t var_mask = prog->var_mask;
#pragma unroll
        for (uint32_t q = 0; q < NXG_ROOT_VAR_COUNT; ++q)
            if (((var_mask >> q) & 1u) && !isfinite(vars[q])) bad = true;
        // the next instruction is fetched while the current one runs (the
        // array ends with a zeroed entry)
        PackedInstruction next = load_instruction(ins);
        for (uint32_t i = 0; i < n; ++i) {
            const PackedInstruction in = next;
            next = load_instruction(ins + i + 1u);
#define OPERAND(e)                                                                          \
    (((e) & kImmediate) ? (((e) & 1u) ? in.immediate1 : in.immediate) : SLOT_GET(e))
#define EXPR_OPV(j) OPERAND(uint32_t{in.slot[j]})
#define EXPR_OPA(j) OPERAND(operand_slots[in.operand_offset + (j)])
#define EXPR_NOISE3(x, y, z) gradient_noise_exact(x, y, z, grad)
#define EXPR_TEXEL(map, f, u, v) texel_sample(a.maps[(map) < a.map_count ? (map) : 0u], f, u, v)
            float value = 0.0f;
#include "expr_ops.inc"
#undef EXPR_OPV
#undef EXPR_OPA
#undef EXPR_NOISE3
#undef EXPR_TEXEL
#undef OPERAND
            if (!isfinite(value)) bad = true;
            SLOT_SET(uint32_t{in.dest}, value);
        }
        // constant indices keep result[] in registers
        const float rv = SLOT_GET(prog->result);
#pragma unroll
        for (uint32_t q = 0; q < NXG_EXPR_MAX_PROGRAMS; ++q) result[q] = q == k ? rv : result[q];
    }
#undef SLOT_GET
#undef SLOT_SET
    if (!live) return;
    if (bad) {
        const float nan = __builtin_nanf("");
        result[0] = result[1] = result[2] = result[3] = nan;
        if (a.error_word) atomicOr(a.error_word, 1u);
    }
    if (a.mode == kAttrCheck) {
        const float4 g = a.attributes[strand];
        const float guess[4] = {g.x, g.y, g.z, g.w};
        bool fail = bad;
#pragma unroll
        for (uint32_t k = 0; k < NXG_EXPR_MAX_PROGRAMS; ++k) {
            if (k < k_first || k >= k_end) continue;
            const float e = result[k], d = guess[k] - e;
            const float mag = (e < 0.0f ? -e : e) > 1.0e-6f ? (e < 0.0f ? -e : e) : 1.0e-6f;
            if (!((d < 0.0f ? -d : d) <= a.threshold * mag)) fail = true;
        }
        if (a.stats && k_first == 0u) atomicAdd(&a.stats[1], 1u);
        if (fail && atomicOr(&a.tile_flags[strand / 64u], 1u) == 0u && a.stats) atomicAdd(&a.stats[0], 1u);
    }
    if (a.split) {
        float *row = reinterpret_cast<float *>(a.attributes + strand);
#pragma unroll
        for (uint32_t k = 0; k < NXG_EXPR_MAX_PROGRAMS; ++k)
            if (k == k_first) row[k] = result[k];
    } else {
        a.attributes[strand] = make_float4(result[0], result[1], result[2], result[3]);
    }
}

__global__ void clear_words_kernel(uint32_t *words, uint64_t count, uint32_t *stats) {
    const uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x;
    if (i < count) words[i] = 0u;
    if (i < 2u && stats) stats[i] = 0u;
}

bool noise_enabled(const nxg_generation_params &p, const nxg_noise_octaves *o) {
    if (p.noise_mask == 0.0f) return false;  // generate.h:291-293
    if (!o || o->count <= 1u) return p.noise_amplitude != 0.0f;
    for (uint32_t k = 0; k < o->count; ++k)
        if (o->amplitude[k] != 0.0f) return true;
    return false;
}

template <typename T>
const T *section(const nxg_device_asset &asset, uint64_t offset) {
    return offset ? reinterpret_cast<const T *>(static_cast<const uint8_t *>(asset.data) + offset)
                  : nullptr;
}

enum Stage : uint32_t { kPrepare = 1u, kStrands = 2u };

// Everything one description's launch needs, validated.
struct Plan {
    Args a;
    Prep p;
    uint64_t prep_lanes;
    bool noise;
    bool ordered;    // strands run grouped by root triangle
    OrderArgs order;
};

uint32_t make_plan(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                   const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                   const nxg_root_sample *explicit_roots, const float *attributes, uint64_t attribute_capacity,
                   const nxg_packed_output *output, const nxg_stencil_output *stencil,
                   const nxg_launch_config *config, void *workspace, uint64_t workspace_bytes, uint32_t stages,
                   Plan &plan) {
    if (uint32_t s = validate_common(asset, deformed, params, config); s != NXG_OK) return s;
    if (stages & kStrands) {
        if (uint32_t s = validate_packed_output(*params, output); s != NXG_OK) return s;
    }
    if (!valid_octaves(octaves)) return NXG_ERROR_INVALID_OCTAVES;
    if (attributes && attribute_capacity < uint64_t{params->strand_count} * NXG_ATTR_COUNT)
        return NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL;
    if (stencil && (stencil->indices || stencil->weights)) {
        if (!stencil->indices || !stencil->weights ||
            stencil->capacity < uint64_t{params->strand_count} * 8u)
            return NXG_ERROR_STENCIL_CAPACITY_TOO_SMALL;
    }
    const nxg_asset_header &h = asset->header;
    const WorkspaceLayout L =
        workspace_layout(h.guide_count, params->cvs_per_strand, params->strand_count, h.triangle_count);
    if (!workspace || workspace_bytes < L.total) return NXG_ERROR_WORKSPACE_TOO_SMALL;
    // the strand kernel addresses the guide-offset table with 32-bit byte offsets
    if (uint64_t{h.guide_count} * params->cvs_per_strand * 16u > 0xFFFFFFFFull)
        return NXG_ERROR_INVALID_GENERATION_PARAMETERS;

    const nxg_vec3 *d_pos = deformed ? deformed->positions : nullptr;
    const nxg_vec3 *d_nrm = deformed ? deformed->normals : nullptr;
    const nxg_vec3 *d_gcv = deformed ? deformed->guide_cvs : nullptr;
    const nxg_vec3 *rest_pos = section<nxg_vec3>(*asset, h.positions_offset);
    const nxg_vec3 *frame_pos = d_pos ? d_pos : rest_pos;
    float4 *state = reinterpret_cast<float4 *>(static_cast<uint8_t *>(workspace) + L.guide_state_offset);
    float4 *offsets = reinterpret_cast<float4 *>(static_cast<uint8_t *>(workspace) + L.guide_offsets_offset);

    plan.prep_lanes = uint64_t{h.guide_count} * params->cvs_per_strand;
    plan.p = Prep{section<nxg_guide_record>(*asset, h.guides_offset), section<nxg_vec3>(*asset, h.guide_cvs_offset),
                  section<nxg_uint3>(*asset, h.triangles_offset), d_pos, d_nrm, d_gcv, frame_pos,
                  h.guide_count, h.triangle_count, params->cvs_per_strand, state, offsets};
    plan.noise = noise_enabled(*params, octaves);

    Args &a = plan.a;
    a = Args{};
    a.positions = frame_pos;
    a.normals = d_nrm ? d_nrm : section<nxg_vec3>(*asset, h.normals_offset);
    a.rest_positions = rest_pos;
    a.texcoords = (h.flags & NXG_ASSET_HAS_TEXCOORDS) ? section<nxg_vec2>(*asset, h.texcoords_offset) : nullptr;
    a.triangles = section<nxg_uint3>(*asset, h.triangles_offset);
    a.alias = section<nxg_alias_entry>(*asset, h.alias_table_offset);
    a.stencil = section<uint32_t>(*asset, h.triangle_guides_offset);
    a.gradients = section<nxg_vec3>(*asset, h.noise_gradients_offset);
    a.guide_state = state;
    a.guide_offsets = offsets;
    a.explicit_roots = explicit_roots;
    a.attributes = reinterpret_cast<const float4 *>(attributes);
    a.strand_count = params->strand_count;
    a.cvs = params->cvs_per_strand;
    a.seed = params->seed;
    a.triangle_count = h.triangle_count;
    a.guide_count = h.guide_count;
    a.has_normals = (h.flags & NXG_ASSET_HAS_NORMALS) != 0u;
    a.support_scale = params->guide_support_scale;
    a.weight_power = params->guide_weight_power;
    a.rejection_cos = params->normal_rejection_cos;
    a.length_scale = params->length_scale;
    a.root_width = params->root_width;
    a.tip_width = params->tip_width;
    a.radius_scale = output ? output->radius_scale : 1.0f;
    a.noise_mask = params->noise_mask;
    const float decorrelation = 1.0f - params->noise_correlation;
    a.decorrelation_scale = 100.0f * decorrelation * decorrelation;
    a.preserve = params->noise_preserve_length;
    a.inv_span = 1.0f / static_cast<float>(params->cvs_per_strand - 1u);
    a.octave_count = 1u;
    if (octaves && octaves->count > 1u) {
        a.octave_count = octaves->count;
        for (uint32_t k = 0; k < a.octave_count; ++k) {
            a.amp[k] = octaves->amplitude[k];
            a.freq[k] = octaves->frequency[k];
        }
    } else {
        a.amp[0] = params->noise_amplitude;
        a.freq[0] = params->noise_frequency;
    }
    a.min_freq = a.freq[0];
    for (uint32_t k = 1; k < a.octave_count; ++k) a.min_freq = a.freq[k] < a.min_freq ? a.freq[k] : a.min_freq;
    a.noise = plan.noise ? 1u : 0u;
    if (output) {
        a.points = output->points;
        a.roots = output->roots;
        a.root_uvs = output->root_uvs;
        a.point_counts = output->point_counts;
    }
    a.stencil_indices = stencil ? stencil->indices : nullptr;
    a.stencil_weights = stencil ? stencil->weights : nullptr;
    // explicit roots keep the caller's order; sampled roots run grouped by triangle
    plan.ordered = explicit_roots == nullptr;
    uint8_t *ws = static_cast<uint8_t *>(workspace);
    plan.order = OrderArgs{a.alias, nullptr, a.seed, a.triangle_count, a.strand_count, a.cvs,
                           order_shift(a.triangle_count), order_tiles(a.strand_count),
                           reinterpret_cast<uint32_t *>(ws + L.counts_offset),
                           reinterpret_cast<uint32_t *>(ws + L.partials_offset),
                           reinterpret_cast<uint32_t *>(ws + L.order_offset)};
    plan.order.keys = reinterpret_cast<uint32_t *>(ws + L.keys_offset);
    a.order = plan.ordered ? plan.order.order : nullptr;
    return NXG_OK;
}

// The counting sort of the strands by root triangle (four small kernels).
uint32_t launch_order(const OrderArgs &o, hipStream_t stream) {
    const uint32_t scan_blocks = (kOrderBins * o.tiles + kScanBlock - 1u) / kScanBlock;
    hipLaunchKernelGGL(order_count_kernel, dim3(o.tiles), dim3(kOrderThreads), 0, stream, o);
    hipLaunchKernelGGL(order_scan_blocks_kernel, dim3(scan_blocks), dim3(256), 0, stream, o);
    hipLaunchKernelGGL(order_scan_partials_kernel, dim3(1), dim3(256), 0, stream, o);
    hipLaunchKernelGGL(order_scatter_kernel, dim3(o.tiles), dim3(kOrderThreads), 0, stream, o);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t launch_plan(const Plan &plan, const nxg_generation_params &params, hipStream_t stream, uint32_t stages) {
    if (stages & kPrepare) {
        hipLaunchKernelGGL(guide_prep_kernel, dim3(static_cast<uint32_t>((plan.prep_lanes + 255u) / 256u)),
                           dim3(256), 0, stream, plan.p);
        if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
        // the strand order depends on the asset, seed and strand count only
        if (plan.ordered)
            if (uint32_t s = launch_order(plan.order, stream); s != NXG_OK) return s;
    }
    if (!(stages & kStrands)) return NXG_OK;
    const Args &a = plan.a;
    const dim3 grid((params.strand_count + 255u) / 256u), block(256);
    if (!plan.noise) {
        hipLaunchKernelGGL((strand_kernel<false, 1>), grid, block, 0, stream, a);
    } else {
        switch (a.octave_count) {
            case 1: hipLaunchKernelGGL((strand_kernel<true, 1>), grid, block, 0, stream, a); break;
            case 2: hipLaunchKernelGGL((strand_kernel<true, 2>), grid, block, 0, stream, a); break;
            case 3: hipLaunchKernelGGL((strand_kernel<true, 3>), grid, block, 0, stream, a); break;
            default: hipLaunchKernelGGL((strand_kernel<true, 4>), grid, block, 0, stream, a); break;
        }
    }
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t launch(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                const nxg_root_sample *explicit_roots, const float *attributes,
                uint64_t attribute_capacity, const nxg_packed_output *output,
                const nxg_stencil_output *stencil, const nxg_launch_config *config, void *workspace,
                uint64_t workspace_bytes, void *stream_ptr, uint32_t stages) {
    Plan plan;
    if (uint32_t s = make_plan(asset, deformed, params, octaves, explicit_roots, attributes, attribute_capacity,
                               output, stencil, config, workspace, workspace_bytes, stages, plan);
        s != NXG_OK)
        return s;
    return launch_plan(plan, *params, static_cast<hipStream_t>(stream_ptr), stages);
}

// Motion samples: every sample is planned (and validated) before the first
// launch; each writes its positions at sample * strand_count * cvs.
uint32_t launch_motion(const nxg_device_asset *asset, const nxg_motion_sample *samples, uint32_t count,
                       const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                       const nxg_motion_output *output, const nxg_launch_config *config, void *workspace,
                       uint64_t workspace_bytes, hipStream_t stream) {
    if (uint32_t s = nxg_validate_motion_generation(asset, samples, count, params, output, config); s != NXG_OK)
        return s;
    if (!valid_octaves(octaves)) return NXG_ERROR_INVALID_OCTAVES;
    const uint64_t one =
        workspace_layout(asset->header.guide_count, params->cvs_per_strand, params->strand_count,
                         asset->header.triangle_count)
            .total;
    if (!workspace || workspace_bytes < one * count) return NXG_ERROR_WORKSPACE_TOO_SMALL;
    const uint64_t per_sample = uint64_t{params->strand_count} * params->cvs_per_strand;
    for (uint32_t pass = 0; pass < 2u; ++pass) {  // pass 0 validates every sample, pass 1 launches
        for (uint32_t k = 0; k < count; ++k) {
            nxg_packed_output out{};
            out.points = reinterpret_cast<nxg_packed_point *>(output->points + per_sample * k);
            out.point_capacity = per_sample;
            out.radius_scale = 1.0f;
            Plan plan;
            if (uint32_t s = make_plan(asset, &samples[k].deformation, params, octaves, nullptr, nullptr, 0u, &out,
                                       nullptr, config, static_cast<uint8_t *>(workspace) + one * k, one,
                                       kPrepare | kStrands, plan);
                s != NXG_OK)
                return s;
            plan.a.vec3_points = 1u;
            if (pass == 1u)
                if (uint32_t s = launch_plan(plan, *params, stream, kPrepare | kStrands); s != NXG_OK) return s;
        }
    }
    return NXG_OK;
}

struct AttrLaunch {
    AttrArgs a;
    uint32_t block, lds;
    bool registers;  // value slots in registers (attribute_kernel<true>)
    uint64_t lanes;
};

uint32_t make_attr_launch(const nxg_device_asset *asset, const nxg_generation_params *params,
                          const nxg_expr_set *set, const nxg_texel_map *maps, uint32_t map_count, float *attributes,
                          uint64_t attribute_capacity, uint32_t *error_word, uint32_t mode, uint32_t stride,
                          float threshold, uint32_t *tile_flags, uint32_t *stats, AttrLaunch &out) {
    if (uint32_t s = validate_common(asset, nullptr, params, nullptr); s != NXG_OK) return s;
    if (!set || !set->image || reinterpret_cast<uintptr_t>(set->image) % 16u != 0u ||
        set->byte_size < sizeof(PackedSet) || set->program_count == 0u ||
        set->program_count > NXG_EXPR_MAX_PROGRAMS || set->max_values == 0u ||
        set->max_values > NXG_EXPR_MAX_VALUES + kVarSlots)
        return NXG_ERROR_INVALID_PROGRAM;
    for (uint32_t k = 0; k < set->program_count; ++k)
        if (set->program_offset[k] < sizeof(PackedSet) || set->program_offset[k] >= set->byte_size)
            return NXG_ERROR_INVALID_PROGRAM;
    if (!attributes || attribute_capacity < uint64_t{params->strand_count} * NXG_ATTR_COUNT)
        return NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL;
    if (map_count > NXG_EXPR_MAX_TEXEL_MAPS || (map_count && !maps) || map_count < set->texel_maps_required)
        return NXG_ERROR_INVALID_PROGRAM;
    for (uint32_t m = 0; m < map_count; ++m)
        if (!maps[m].texels || maps[m].face_count == 0u || maps[m].resolution == 0u) return NXG_ERROR_INVALID_PROGRAM;
    const nxg_asset_header &h = asset->header;
    AttrArgs &a = out.a;
    a = AttrArgs{};
    a.positions = section<nxg_vec3>(*asset, h.positions_offset);
    a.normals = section<nxg_vec3>(*asset, h.normals_offset);
    a.texcoords = (h.flags & NXG_ASSET_HAS_TEXCOORDS) ? section<nxg_vec2>(*asset, h.texcoords_offset) : nullptr;
    a.triangles = section<nxg_uint3>(*asset, h.triangles_offset);
    a.alias = section<nxg_alias_entry>(*asset, h.alias_table_offset);
    a.gradients = section<nxg_vec3>(*asset, h.noise_gradients_offset);
    a.programs = static_cast<const uint8_t *>(set->image);
    for (uint32_t m = 0; m < map_count; ++m) a.maps[m] = maps[m];
    a.map_count = map_count;
    a.strand_count = params->strand_count;
    a.seed = params->seed;
    a.triangle_count = h.triangle_count;
    a.has_normals = (h.flags & NXG_ASSET_HAS_NORMALS) != 0u;
    a.program_count = set->program_count;
    for (uint32_t k = 0; k < set->program_count; ++k) a.program_offset[k] = set->program_offset[k];
    a.attributes = reinterpret_cast<float4 *>(attributes);
    a.error_word = error_word;
    a.mode = mode;
    a.stride = stride;
    a.threshold = threshold;
    a.tile_flags = tile_flags;
    a.stats = stats;
    // values[max_values][block] floats plus the gradient table, within 64 KB
    // of LDS where a block of at least 64 lanes allows it
    uint32_t block = 256u;
    while (block > 64u && uint64_t{set->max_values} * block * 4u + 4096u > 65536u) block -= 64u;
    out.block = block;
    out.registers = set->max_values <= kRegSlots;
    out.lds = (out.registers ? 0u : set->max_values * block * 4u) + 4u * 4u * NXG_NOISE_GRADIENT_COUNT;
    out.lanes = mode == kAttrCheck ? (uint64_t{params->strand_count} + stride - 1u) / stride : params->strand_count;
    a.split = mode != kAttrAll ? 1u : 0u;
    return NXG_OK;
}

uint32_t launch_attributes(const AttrLaunch &l, hipStream_t stream) {
    if (l.lanes == 0u) return NXG_OK;
    if (l.lds > 65536u &&
        hipFuncSetAttribute(reinterpret_cast<const void *>(&attribute_kernel<false>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(l.lds)) != hipSuccess)
        return NXG_ERROR_HIP;
    const dim3 grid(static_cast<uint32_t>((l.lanes + l.block - 1u) / l.block), l.a.split ? l.a.program_count : 1u),
        block(l.block);
    if (l.registers) hipLaunchKernelGGL(attribute_kernel<true>, grid, block, l.lds, stream, l.a);
    else hipLaunchKernelGGL(attribute_kernel<false>, grid, block, l.lds, stream, l.a);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t evaluate_attributes(const nxg_device_asset *asset, const nxg_generation_params *params,
                             const nxg_expr_set *set, const nxg_texel_map *maps, uint32_t map_count,
                             float *attributes, uint64_t attribute_capacity, uint32_t *error_word,
                             void *stream_ptr, uint32_t mode = kAttrAll, uint32_t stride = 1u,
                             float threshold = 0.0f, uint32_t *tile_flags = nullptr,
                             uint32_t *stats = nullptr) {
    AttrLaunch l;
    if (uint32_t s = make_attr_launch(asset, params, set, maps, map_count, attributes, attribute_capacity, error_word,
                                      mode, stride, threshold, tile_flags, stats, l);
        s != NXG_OK)
        return s;
    return launch_attributes(l, static_cast<hipStream_t>(stream_ptr));
}

// ---- batch host side --------------------------------------------------

constexpr uint32_t kBatchWaveChunk[][2] = {{1, 4}, {1, 8}, {1, 16}, {2, 4}, {2, 8}, {2, 16}, {4, 4}, {4, 8}};

template <int OCTMAX>
const void *batch_kernel_of(uint32_t waves, uint32_t chunk) {
    if (waves == 1u && chunk == 4u) return reinterpret_cast<const void *>(&batch_kernel<1, 4, OCTMAX>);
    if (waves == 1u && chunk == 8u) return reinterpret_cast<const void *>(&batch_kernel<1, 8, OCTMAX>);
    if (waves == 1u && chunk == 16u) return reinterpret_cast<const void *>(&batch_kernel<1, 16, OCTMAX>);
    if (waves == 2u && chunk == 4u) return reinterpret_cast<const void *>(&batch_kernel<2, 4, OCTMAX>);
    if (waves == 2u && chunk == 8u) return reinterpret_cast<const void *>(&batch_kernel<2, 8, OCTMAX>);
    if (waves == 2u && chunk == 16u) return reinterpret_cast<const void *>(&batch_kernel<2, 16, OCTMAX>);
    if (waves == 4u && chunk == 4u) return reinterpret_cast<const void *>(&batch_kernel<4, 4, OCTMAX>);
    if (waves == 4u && chunk == 8u) return reinterpret_cast<const void *>(&batch_kernel<4, 8, OCTMAX>);
    return nullptr;
}

// The batch kernel carries noise state for two octaves when no item uses
// more (fewer registers, three wavefronts per SIMD), else for four.
const void *batch_kernel_for(uint32_t waves, uint32_t chunk, uint32_t max_octaves = 2u) {
    return max_octaves <= 2u ? batch_kernel_of<2>(waves, chunk) : batch_kernel_of<4>(waves, chunk);
}

nxg_batch_tuning default_tuning() { return nxg_batch_tuning{4u, 8u, 1u, 0u, NXG_BATCH_ORDER_HEAVY_FIRST}; }

bool valid_tuning(const nxg_batch_tuning &t) {
    return batch_kernel_for(t.waves_per_block, t.chunk_cvs) && t.tiles_per_claim >= 1u && t.tiles_per_claim <= 16u &&
           t.blocks_per_cu <= 64u && t.order <= NXG_BATCH_ORDER_HEAVY_FIRST;
}

// Relative cost of one strand: CV count times noise passes and octaves.
double strand_cost(const Args &a) {
    const double passes = (a.noise && a.preserve > 0.001f) ? 2.0 : 1.0;
    return static_cast<double>(a.cvs) * passes * (a.noise ? 1.0 + 3.0 * a.octave_count : 1.0);
}

struct BatchLaunch {
    AttrLaunch attrs[NXG_BATCH_MAX_ITEMS];
    bool has_attrs[NXG_BATCH_MAX_ITEMS];
    const nxg_batch_item *items;  // the caller's items (surrogate launches read them)
    BatchOrderArgs order;
    uint32_t order_tiles, order_scan_blocks;
    BatchPrepArgs prep;
    BatchArgs strands;
    uint32_t prep_blocks;
    uint32_t max_octaves;
    nxg_batch_tuning tuning;
};

uint32_t make_batch(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning, uint32_t *counter,
                    BatchLaunch &out) {
    if (!items || !counter) return NXG_ERROR_NULL_ARGUMENT;
    if (count == 0u || count > NXG_BATCH_MAX_ITEMS) return NXG_ERROR_INVALID_BATCH;
    out.tuning = tuning ? *tuning : default_tuning();
    if (!valid_tuning(out.tuning)) return NXG_ERROR_INVALID_BATCH;
    std::memset(&out.prep, 0, sizeof(out.prep));
    std::memset(&out.strands, 0, sizeof(out.strands));
    out.max_octaves = 1u;
    uint64_t blocks = 0u;
    double cost[NXG_BATCH_MAX_ITEMS];
    for (uint32_t k = 0; k < count; ++k) {
        const nxg_batch_item &it = items[k];
        if (!it.params || !it.output) return NXG_ERROR_NULL_ARGUMENT;
        out.has_attrs[k] = it.exprs != nullptr;
        if (it.surrogate) {  // the checks nxg_surrogate_attributes and nxg_verify_attributes make
            const nxg_texel_map *m = it.surrogate_map;
            if (!it.exprs || !m || !m->texels || m->face_count == 0u || m->resolution == 0u || !it.surrogate_flags)
                return NXG_ERROR_NULL_ARGUMENT;
            if (it.surrogate_flag_capacity < (uint64_t{it.params->strand_count} + 63u) / 64u)
                return NXG_ERROR_BUFFER_TOO_SMALL;
            if (it.surrogate_stride == 0u || !(it.surrogate_threshold >= 0.0f)) return NXG_ERROR_INVALID_MLP_REQUEST;
        }
        if (it.exprs) {
            if (uint32_t s = make_attr_launch(it.asset, it.params, it.exprs, it.maps, it.map_count, it.attributes,
                                              it.attribute_capacity, nullptr, kAttrAll, 1u, 0.0f, nullptr, nullptr,
                                              out.attrs[k]);
                s != NXG_OK)
                return s;
        }
        Plan plan;
        if (uint32_t s = make_plan(it.asset, it.deformed, it.params, it.octaves, nullptr, it.attributes,
                                   it.attribute_capacity, it.output, it.stencil, nullptr, it.workspace,
                                   it.workspace_bytes, kPrepare | kStrands, plan);
            s != NXG_OK)
            return s;
        out.prep.items[k] = plan.p;
        blocks += (plan.prep_lanes + 255u) / 256u;
        out.prep.block_end[k] = static_cast<uint32_t>(blocks);
        out.strands.items[k] = plan.a;
        out.order.items[k] = plan.order;  // batch items always sample their roots
        if (plan.noise && plan.a.octave_count > out.max_octaves) out.max_octaves = plan.a.octave_count;
        cost[k] = strand_cost(plan.a);
        out.strands.order[k] = k;
    }
    if (out.tuning.order == NXG_BATCH_ORDER_HEAVY_FIRST) {
        // stable insertion sort, heaviest strands first
        for (uint32_t i = 1; i < count; ++i)
            for (uint32_t j = i; j > 0 && cost[out.strands.order[j]] > cost[out.strands.order[j - 1u]]; --j) {
                const uint32_t t = out.strands.order[j];
                out.strands.order[j] = out.strands.order[j - 1u];
                out.strands.order[j - 1u] = t;
            }
    }
    uint64_t tiles = 0u;
    for (uint32_t k = 0; k < count; ++k) {
        tiles += (uint64_t{out.strands.items[out.strands.order[k]].strand_count} + 63u) / 64u;
        out.strands.tile_end[k] = static_cast<uint32_t>(tiles);
    }
    out.order.count = count;
    out.order_tiles = out.order_scan_blocks = 0u;
    for (uint32_t k = 0; k < count; ++k) {
        const OrderArgs &o = out.order.items[k];
        out.order_tiles += o.tiles;
        out.order_scan_blocks += (kOrderBins * o.tiles + kScanBlock - 1u) / kScanBlock;
        out.order.tile_end[k] = out.order_tiles;
        out.order.scan_end[k] = out.order_scan_blocks;
    }
    if (blocks > 0x7fffffffu || tiles > 0x7fffffffu) return NXG_ERROR_OUTPUT_SIZE_OVERFLOW;
    out.prep.count = count;
    out.prep.counter = counter;
    out.prep_blocks = static_cast<uint32_t>(blocks);
    out.strands.count = count;
    out.strands.total_tiles = static_cast<uint32_t>(tiles);
    out.strands.tiles_per_claim = out.tuning.tiles_per_claim;
    out.strands.counter = counter;
    out.items = items;
    return NXG_OK;
}

// Side streams and their events for a captured batch: the items' exact
// attribute programs do not read the strand order, so in a graph they run on
// the side streams beside the order kernels (and beside each other), joining
// before the persistent kernel.  Three sides and the capture stream use the
// four hardware queues a process gets by default (a fourth side was slower).
struct BatchSides {
    static constexpr uint32_t kCount = 3u;
    hipStream_t stream[kCount];
    hipEvent_t fork, join[kCount];
};

uint32_t launch_batch(BatchLaunch &b, hipStream_t stream, const BatchSides *sides = nullptr) {
    uint64_t load[BatchSides::kCount] = {};  // roots given to each side so far
    auto lightest = [&load]() {
        uint32_t side = 0u;
        for (uint32_t i = 1; i < BatchSides::kCount; ++i)
            if (load[i] < load[side]) side = i;
        return side;
    };
    if (sides) {
        if (hipEventRecord(sides->fork, stream) != hipSuccess) return NXG_ERROR_HIP;
        for (uint32_t i = 0; i < BatchSides::kCount; ++i)
            if (hipStreamWaitEvent(sides->stream[i], sides->fork, 0) != hipSuccess) return NXG_ERROR_HIP;
        // largest first, each to the side with the fewest roots so far; a
        // surrogate item's network reads its roots in index order here (the
        // same values as in the strand order), so it need not wait for the order
        bool done[NXG_BATCH_MAX_ITEMS] = {};
        for (;;) {
            uint32_t pick = NXG_BATCH_MAX_ITEMS;
            for (uint32_t k = 0; k < b.strands.count; ++k)
                if (b.has_attrs[k] && !done[k] &&
                    (pick == NXG_BATCH_MAX_ITEMS ||
                     b.items[k].params->strand_count > b.items[pick].params->strand_count))
                    pick = k;
            if (pick == NXG_BATCH_MAX_ITEMS) break;
            done[pick] = true;
            const uint32_t side = lightest();
            load[side] += b.items[pick].params->strand_count;
            const nxg_batch_item &it = b.items[pick];
            if (!it.surrogate) {
                if (uint32_t s = launch_attributes(b.attrs[pick], sides->stream[side]); s != NXG_OK) return s;
                continue;
            }
            if (uint32_t s = nxg_surrogate_attributes(it.asset, it.params, it.surrogate, it.surrogate_map,
                                                      it.attributes, it.attribute_capacity, nullptr, 0u,
                                                      sides->stream[side]);
                s != NXG_OK)
                return s;
            if (uint32_t s = nxg_verify_attributes(it.asset, it.params, it.exprs, it.maps, it.map_count,
                                                   it.attributes, it.attribute_capacity, it.surrogate_threshold,
                                                   it.surrogate_stride, it.surrogate_flags,
                                                   it.surrogate_flag_capacity, it.surrogate_stats,
                                                   sides->stream[side]);
                s != NXG_OK)
                return s;
        }
    }
    if (b.order_tiles) {
        hipLaunchKernelGGL(batch_order_count_kernel, dim3(b.order_tiles), dim3(kOrderThreads), 0, stream, b.order);
        hipLaunchKernelGGL(batch_order_scan_blocks_kernel, dim3(b.order_scan_blocks), dim3(256), 0, stream, b.order);
        hipLaunchKernelGGL(batch_order_scan_partials_kernel, dim3(b.order.count), dim3(256), 0, stream, b.order);
        hipLaunchKernelGGL(batch_order_scatter_kernel, dim3(b.order_tiles), dim3(kOrderThreads), 0, stream, b.order);
        if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    }
    // without sides, attributes after the order: a surrogate reads its roots
    // in the item's strand order
    for (uint32_t k = 0; !sides && k < b.strands.count; ++k) {
        if (!b.has_attrs[k]) continue;
        const nxg_batch_item &it = b.items[k];
        if (!it.surrogate) {
            if (uint32_t s = launch_attributes(b.attrs[k], stream); s != NXG_OK) return s;
            continue;
        }
        if (uint32_t s = nxg_surrogate_attributes(it.asset, it.params, it.surrogate, it.surrogate_map, it.attributes,
                                                  it.attribute_capacity, it.workspace, it.workspace_bytes, stream);
            s != NXG_OK)
            return s;
        if (uint32_t s = nxg_verify_attributes(it.asset, it.params, it.exprs, it.maps, it.map_count, it.attributes,
                                               it.attribute_capacity, it.surrogate_threshold, it.surrogate_stride,
                                               it.surrogate_flags, it.surrogate_flag_capacity, it.surrogate_stats,
                                               stream);
            s != NXG_OK)
            return s;
    }
    if (b.prep_blocks) {
        hipLaunchKernelGGL(batch_prep_kernel, dim3(b.prep_blocks), dim3(256), 0, stream, b.prep);
    } else {
        // no guides anywhere: the counter still has to start at zero
        hipLaunchKernelGGL(batch_prep_kernel, dim3(1), dim3(1), 0, stream, b.prep);
    }
    if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    if (sides) {
        for (uint32_t i = 0; i < BatchSides::kCount; ++i)
            if (hipEventRecord(sides->join[i], sides->stream[i]) != hipSuccess ||
                hipStreamWaitEvent(stream, sides->join[i], 0) != hipSuccess)
                return NXG_ERROR_HIP;
    }
    if (b.strands.total_tiles == 0u) return NXG_OK;
    const nxg_batch_tuning &t = b.tuning;
    const void *kernel = batch_kernel_for(t.waves_per_block, t.chunk_cvs, b.max_octaves);
    const uint32_t threads = t.waves_per_block * kWave;
    int device = 0, cus = 0, resident = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kernel, static_cast<int>(threads), 0) != hipSuccess)
        return NXG_ERROR_HIP;
    if (resident < 1) resident = 1;
    const uint32_t per_cu = t.blocks_per_cu ? min(t.blocks_per_cu, static_cast<uint32_t>(resident))
                                            : static_cast<uint32_t>(resident);
    const uint64_t claims = (uint64_t{b.strands.total_tiles} + t.tiles_per_claim - 1u) / t.tiles_per_claim;
    uint64_t grid = uint64_t{per_cu} * static_cast<uint32_t>(cus);
    grid = std::min<uint64_t>(grid, (claims + t.waves_per_block - 1u) / t.waves_per_block);
    void *args[] = {&b.strands};
    if (hipLaunchKernel(kernel, dim3(static_cast<uint32_t>(grid)), dim3(threads), args, 0, stream) != hipSuccess)
        return NXG_ERROR_HIP;
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

// ---- Opt-in inverse-CDF root sampler (no reference counterpart) ----------
// The reference draws a root triangle from the blob's Walker alias table
// (generate.h:107-112).  This sampler draws from the same distribution
// through a prefix sum and a binary search instead; it sends a strand to a
// different triangle than the alias draw, so it is used only when the caller
// asks for it.  The distribution is read off the alias table in integers,
// in units of 2^-32 of one bucket's probability mass:
//   bucket j keeps keep_j = ceil(prob_j * 2^24) * 2^8, the probability that
//   random01(seed, strand, 1) (one of 2^24 equally likely values k / 2^24)
//   falls below prob_j, and hands 2^32 - keep_j to its alias (to itself when
//   the alias is out of range, as sample_triangle does).
// The masses sum to T * 2^32 exactly and their prefix sums are integers, so
// every summation order gives the same table.  A strand draws
// r = h0 * 2^32 + h1 from the hashes of random01's dimensions 0 and 1,
// target = floor(r * T / 2^32) in [0, T * 2^32), and takes the first triangle
// whose inclusive prefix sum exceeds target; its barycentrics come from
// dimensions 2 and 3, as in the alias draw.  The sampler buffer holds the
// prefix sums (u64 x T), the scan's block sums, and a coarse table of every
// S-th prefix sum (at most kCdfCoarseMax entries) that each block keeps in
// LDS: the search runs over the coarse table in LDS, then over one S-entry
// segment of the prefix sums in memory.

constexpr uint32_t kCdfCoarseMax = 4096u;   // coarse entries in LDS (32 KB)
constexpr uint32_t kCdfBlockItems = 4096u;  // prefix-sum elements per 256-thread scan block

struct CdfLayout {
    uint64_t sums_offset, coarse_offset, total_bytes;
    uint32_t stride_log2, coarse_count, block_count;
};

CdfLayout cdf_layout(uint32_t triangle_count) {
    CdfLayout l{};
    auto align = [](uint64_t x) { return (x + 255u) & ~uint64_t{255u}; };
    l.stride_log2 = 6u;
    while ((uint64_t{triangle_count} + (uint64_t{1} << l.stride_log2) - 1u) >> l.stride_log2 > kCdfCoarseMax)
        ++l.stride_log2;
    l.coarse_count = static_cast<uint32_t>((uint64_t{triangle_count} + (uint64_t{1} << l.stride_log2) - 1u) >>
                                           l.stride_log2);
    l.block_count = (triangle_count + kCdfBlockItems - 1u) / kCdfBlockItems;
    l.sums_offset = align(uint64_t{triangle_count} * 8u);
    l.coarse_offset = l.sums_offset + align(uint64_t{l.block_count} * 8u);
    l.total_bytes = l.coarse_offset + align(uint64_t{l.coarse_count} * 8u);
    return l;
}

struct CdfArgs {
    const nxg_alias_entry *alias;
    uint64_t *prefix, *sums, *coarse;
    uint32_t triangle_count, stride_log2, coarse_count, block_count;
};

__device__ __forceinline__ uint64_t alias_keep(const nxg_alias_entry &e) {
    const float p = e.probability;
    uint32_t k = 0u;  // also for a NaN probability, which never keeps
    if (p >= 1.0f) k = 1u << 24;
    else if (p > 0.0f) k = static_cast<uint32_t>(ceilf(p * 16777216.0f));
    return uint64_t{k} << 8;
}

__global__ void __launch_bounds__(256) cdf_keep_kernel(CdfArgs c) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j < c.triangle_count) c.prefix[j] = alias_keep(c.alias[j]);
}

__global__ void __launch_bounds__(256) cdf_hand_over_kernel(CdfArgs c) {
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= c.triangle_count) return;
    const nxg_alias_entry e = c.alias[j];
    const uint64_t give = (uint64_t{1} << 32) - alias_keep(e);
    if (give) {
        const uint32_t to = e.alias < c.triangle_count ? e.alias : j;
        atomicAdd(reinterpret_cast<unsigned long long *>(&c.prefix[to]), static_cast<unsigned long long>(give));
    }
}

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, uint32_t d) {
    const uint32_t lo = __shfl_up(static_cast<uint32_t>(v), d, kWave);
    const uint32_t hi = __shfl_up(static_cast<uint32_t>(v >> 32), d, kWave);
    return (uint64_t{hi} << 32) | lo;
}

// Exclusive scan of one u64 per thread of a 256-thread block; total is the
// block's sum.
__device__ __forceinline__ uint64_t block_scan_u64(uint64_t v, uint64_t &total) {
    __shared__ uint64_t wave_totals[256u / kWave];
    const uint32_t lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    uint64_t inc = v;
#pragma unroll
    for (uint32_t d = 1; d < static_cast<uint32_t>(kWave); d <<= 1) {
        const uint64_t y = shfl_up_u64(inc, d);
        if (lane >= d) inc += y;
    }
    if (lane == kWave - 1) wave_totals[wave] = inc;
    __syncthreads();
    uint64_t before = 0u;
    total = 0u;
#pragma unroll
    for (uint32_t w = 0; w < 256u / kWave; ++w) {
        if (w < wave) before += wave_totals[w];
        total += wave_totals[w];
    }
    __syncthreads();
    return before + inc - v;
}

// Sum of each kCdfBlockItems-element block of the masses.
__global__ void __launch_bounds__(256) cdf_block_sums_kernel(CdfArgs c) {
    const uint64_t base = uint64_t{blockIdx.x} * kCdfBlockItems;
    uint64_t v = 0u;
    for (uint32_t i = threadIdx.x; i < kCdfBlockItems; i += 256u)
        if (base + i < c.triangle_count) v += c.prefix[base + i];
    uint64_t total;
    block_scan_u64(v, total);
    if (threadIdx.x == 0u) c.sums[blockIdx.x] = total;
}

// Exclusive scan of the block sums in place (one block, 256 at a time).
__global__ void __launch_bounds__(256) cdf_scan_sums_kernel(CdfArgs c) {
    uint64_t carry = 0u;
    for (uint32_t first = 0u; first < c.block_count; first += 256u) {
        const uint32_t i = first + threadIdx.x;
        const uint64_t v = i < c.block_count ? c.sums[i] : 0u;
        uint64_t total;
        const uint64_t ex = block_scan_u64(v, total);
        if (i < c.block_count) c.sums[i] = carry + ex;
        carry += total;
    }
}

// Inclusive prefix sums of one block's masses in place, offset by the
// block's scanned sum, and the coarse entries (every S-th sum, and the last)
// that fall in the block.  Each thread scans 16 consecutive elements staged
// in LDS (row pitch 17 to spread the banks).
__global__ void __launch_bounds__(256) cdf_scan_kernel(CdfArgs c) {
    __shared__ uint64_t tile[256u * 17u];
    const uint64_t base = uint64_t{blockIdx.x} * kCdfBlockItems;
    const uint32_t t = threadIdx.x;
    for (uint32_t i = t; i < kCdfBlockItems; i += 256u)
        tile[(i / 16u) * 17u + i % 16u] = base + i < c.triangle_count ? c.prefix[base + i] : 0u;
    __syncthreads();
    uint64_t run = 0u;
    for (uint32_t k = 0; k < 16u; ++k) run += tile[t * 17u + k];
    uint64_t total;
    uint64_t before = block_scan_u64(run, total) + c.sums[blockIdx.x];
    for (uint32_t k = 0; k < 16u; ++k) {
        before += tile[t * 17u + k];
        tile[t * 17u + k] = before;
    }
    __syncthreads();
    const uint64_t last = uint64_t{c.triangle_count} - 1u;
    const uint64_t stride_mask = (uint64_t{1} << c.stride_log2) - 1u;
    for (uint32_t i = t; i < kCdfBlockItems; i += 256u) {
        const uint64_t g = base + i;
        if (g > last) break;
        const uint64_t value = tile[(i / 16u) * 17u + i % 16u];
        c.prefix[g] = value;
        if ((g & stride_mask) == stride_mask || g == last) c.coarse[g >> c.stride_log2] = value;
    }
}

// One lane per strand (grid-stride over 256-strand chunks of the processing
// order, sorted by target, so that a wavefront's searches and vertex reads
// share cache lines): the triangle by the two-level search, then the alias
// mode's root on it, as a RootSample.
__global__ void __launch_bounds__(256) cdf_roots_kernel(Args a, CdfArgs c, nxg_root_sample *roots) {
    __shared__ uint64_t coarse[kCdfCoarseMax];
    // eight loads in flight per thread: no branches (a thread past the end
    // copies the last entry again), so all issue before the first store
    for (uint32_t base = threadIdx.x; base < c.coarse_count; base += 8u * 256u) {
        uint64_t v[8];
#pragma unroll
        for (uint32_t j = 0u; j < 8u; ++j) v[j] = c.coarse[min(base + j * 256u, c.coarse_count - 1u)];
#pragma unroll
        for (uint32_t j = 0u; j < 8u; ++j) coarse[min(base + j * 256u, c.coarse_count - 1u)] = v[j];
    }
    __syncthreads();
    const uint64_t T = c.triangle_count;
    for (uint64_t slot = uint64_t{blockIdx.x} * 256u + threadIdx.x; slot < a.strand_count;
         slot += uint64_t{gridDim.x} * 256u) {
        const uint32_t strand = a.order ? a.order[slot] : static_cast<uint32_t>(slot);
        const uint64_t target = cdf_target(a.seed, strand, c.triangle_count);
        uint32_t lo = 0u, hi = c.coarse_count - 1u;  // coarse[last] = T 2^32 > target
        while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (coarse[mid] > target) hi = mid;
            else lo = mid + 1u;
        }
        // the first triangle with a prefix sum above target lies in coarse
        // segment lo: [lo S, min((lo + 1) S, T) - 1]
        const uint64_t seg = uint64_t{lo} << c.stride_log2;
        const uint64_t end = seg + (uint64_t{1} << c.stride_log2) < T ? seg + (uint64_t{1} << c.stride_log2) : T;
        uint32_t first = static_cast<uint32_t>(seg), last = static_cast<uint32_t>(end - 1u);
        while (first < last) {
            const uint32_t mid = (first + last) >> 1;
            if (c.prefix[mid] > target) last = mid;
            else first = mid + 1u;
        }
        const Root r = root_on_triangle(a, strand, first);
        nxg_root_sample o;
        o.position = {r.position.x, r.position.y, r.position.z};
        o.normal = {r.normal.x, r.normal.y, r.normal.z};
        o.uv = {r.u, r.v};
        o.triangle_index = r.tri;
        o.barycentric = {r.b1, r.b2};
        o.surface_face_id = r.tri;
        roots[strand] = o;
    }
}

uint32_t cdf_args(const nxg_device_asset *asset, const void *sampler, uint64_t sampler_bytes, CdfArgs &c) {
    if (uint32_t s = validate_asset_view(asset); s != NXG_OK) return s;
    const nxg_asset_header &h = asset->header;
    const CdfLayout l = cdf_layout(h.triangle_count);
    if (!sampler || reinterpret_cast<uintptr_t>(sampler) % 256u != 0u || sampler_bytes < l.total_bytes)
        return NXG_ERROR_INVALID_SAMPLER;
    uint8_t *base = static_cast<uint8_t *>(const_cast<void *>(sampler));
    c.alias = section<nxg_alias_entry>(*asset, h.alias_table_offset);
    c.prefix = reinterpret_cast<uint64_t *>(base);
    c.sums = reinterpret_cast<uint64_t *>(base + l.sums_offset);
    c.coarse = reinterpret_cast<uint64_t *>(base + l.coarse_offset);
    c.triangle_count = h.triangle_count;
    c.stride_log2 = l.stride_log2;
    c.coarse_count = l.coarse_count;
    c.block_count = l.block_count;
    return NXG_OK;
}

uint32_t cdf_build(const nxg_device_asset *asset, void *sampler, uint64_t sampler_bytes, hipStream_t stream) {
    CdfArgs c;
    if (uint32_t s = cdf_args(asset, sampler, sampler_bytes, c); s != NXG_OK) return s;
    const dim3 lanes((c.triangle_count + 255u) / 256u);
    hipLaunchKernelGGL(cdf_keep_kernel, lanes, dim3(256), 0, stream, c);
    hipLaunchKernelGGL(cdf_hand_over_kernel, lanes, dim3(256), 0, stream, c);
    hipLaunchKernelGGL(cdf_block_sums_kernel, dim3(c.block_count), dim3(256), 0, stream, c);
    hipLaunchKernelGGL(cdf_scan_sums_kernel, dim3(1), dim3(256), 0, stream, c);
    hipLaunchKernelGGL(cdf_scan_kernel, dim3(c.block_count), dim3(256), 0, stream, c);
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

// Generation with inverse-CDF roots: guide prep and the processing order
// (by target), the roots sampled into output->roots (required) in that
// order, then the strands from them in the same order.
uint32_t generate_cdf(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                      const nxg_generation_params *params, const nxg_noise_octaves *octaves, const void *sampler,
                      uint64_t sampler_bytes, const nxg_packed_output *output, const nxg_stencil_output *stencil,
                      const nxg_launch_config *config, void *workspace, uint64_t workspace_bytes,
                      hipStream_t stream) {
    if (uint32_t s = validate_common(asset, deformed, params, config); s != NXG_OK) return s;
    if (uint32_t s = validate_packed_output(*params, output); s != NXG_OK) return s;
    if (!output->roots) return NXG_ERROR_ROOT_CAPACITY_TOO_SMALL;
    CdfArgs c;
    if (uint32_t s = cdf_args(asset, sampler, sampler_bytes, c); s != NXG_OK) return s;
    Plan plan;
    if (uint32_t s = make_plan(asset, deformed, params, octaves, output->roots, nullptr, 0u, output, stencil, config,
                               workspace, workspace_bytes, kPrepare | kStrands, plan);
        s != NXG_OK)
        return s;
    // one order for both kernels: by target / 2^32, which is monotone in the
    // triangle drawn, so the strands also run grouped by triangle
    plan.a.roots = nullptr;  // the sampler writes them
    plan.ordered = true;
    plan.order.cdf_key = 1u;
    plan.a.order = plan.order.order;
    if (uint32_t s = launch_plan(plan, *params, stream, kPrepare); s != NXG_OK) return s;
    int device = 0, cus = 0;
    if (hipGetDevice(&device) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return NXG_ERROR_HIP;
    const uint64_t blocks = std::min<uint64_t>((uint64_t{params->strand_count} + 255u) / 256u,
                                               uint64_t{static_cast<uint32_t>(cus)} * 4u);
    if (blocks) {
        hipLaunchKernelGGL(cdf_roots_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, stream, plan.a, c,
                           output->roots);
        if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    }
    return launch_plan(plan, *params, stream, kStrands);
}

}  // namespace
// The static scratch of nxg_generate_packed_noalloc: a file-scope device
// array, one copy per GPU at that GPU's own address.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the
// symbol); it uses the address hipGetSymbolAddress returns for the current
// device, looked up once per device on its first call.
constexpr uint64_t kStaticScratchBytes = NXG_STATIC_SCRATCH_BYTES;
__device__ __align__(256) uint8_t scratch[kStaticScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex scratch_mutex;
void *scratch_ptr[kMaxDevices];

// The current device's scratch address, or nullptr with *status set.
void *device_scratch(uint32_t *status) {
    int device = -1;
    if (hipGetDevice(&device) != hipSuccess || device < 0 || device >= kMaxDevices) {
        *status = NXG_ERROR_HIP;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[device]) {
        void *ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            *status = NXG_ERROR_HIP;
            return nullptr;
        }
        scratch_ptr[device] = ptr;
    }
    return scratch_ptr[device];
}

}  // namespace nxg

extern "C" {

uint32_t nxg_evaluate_root_attributes(const nxg_device_asset *asset, const nxg_generation_params *params,
                                      const nxg_expr_set *set, const nxg_texel_map *maps, uint32_t map_count,
                                      float *attributes, uint64_t attribute_capacity, uint32_t *error_word,
                                      void *stream) {
    return nxg::evaluate_attributes(asset, params, set, maps, map_count, attributes, attribute_capacity,
                                    error_word, stream);
}

uint32_t nxg_verify_attributes(const nxg_device_asset *asset, const nxg_generation_params *params,
                               const nxg_expr_set *set, const nxg_texel_map *maps, uint32_t map_count,
                               float *attributes, uint64_t attribute_capacity, float relative_threshold,
                               uint32_t sample_stride, uint32_t *tile_flags, uint64_t tile_flag_capacity,
                               uint32_t *stats, void *stream) {
    if (!params || !tile_flags) return NXG_ERROR_NULL_ARGUMENT;
    const uint64_t tiles = (uint64_t{params->strand_count} + 63u) / 64u;
    if (tile_flag_capacity < tiles) return NXG_ERROR_BUFFER_TOO_SMALL;
    if (sample_stride == 0u || !(relative_threshold >= 0.0f)) return NXG_ERROR_INVALID_MLP_REQUEST;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(nxg::clear_words_kernel, dim3(static_cast<uint32_t>((tiles + 255u) / 256u)), dim3(256), 0, s,
                       tile_flags, tiles, stats);
    if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    uint32_t st = nxg::evaluate_attributes(asset, params, set, maps, map_count, attributes, attribute_capacity,
                                           nullptr, stream, nxg::kAttrCheck, sample_stride, relative_threshold,
                                           tile_flags, stats);
    if (st != NXG_OK) return st;
    return nxg::evaluate_attributes(asset, params, set, maps, map_count, attributes, attribute_capacity, nullptr,
                                    stream, nxg::kAttrFallback, 1u, 0.0f, tile_flags, stats);
}

uint32_t nxg_generate_packed_attributed(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                                        const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                                        const float *attributes, uint64_t attribute_capacity,
                                        const nxg_packed_output *output, const nxg_stencil_output *stencil,
                                        const nxg_launch_config *config, void *workspace,
                                        uint64_t workspace_bytes, void *stream) {
    if (!attributes) return NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL;
    return nxg::launch(asset, deformed, params, octaves, nullptr, attributes, attribute_capacity, output,
                       stencil, config, workspace, workspace_bytes, stream, nxg::kPrepare | nxg::kStrands);
}

uint32_t nxg_generate_packed_attributed_prepared(const nxg_device_asset *asset,
                                                 const nxg_deformed_geometry *deformed,
                                                 const nxg_generation_params *params,
                                                 const nxg_noise_octaves *octaves, const float *attributes,
                                                 uint64_t attribute_capacity, const nxg_packed_output *output,
                                                 const nxg_stencil_output *stencil,
                                                 const nxg_launch_config *config, void *workspace,
                                                 uint64_t workspace_bytes, void *stream) {
    if (!attributes) return NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL;
    return nxg::launch(asset, deformed, params, octaves, nullptr, attributes, attribute_capacity, output, stencil,
                       config, workspace, workspace_bytes, stream, nxg::kStrands);
}

uint32_t nxg_generate_packed(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                             const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                             const nxg_packed_output *output, const nxg_stencil_output *stencil,
                             const nxg_launch_config *config, void *workspace,
                             uint64_t workspace_bytes, void *stream) {
    return nxg::launch(asset, deformed, params, octaves, nullptr, nullptr, 0, output, stencil, config, workspace,
                       workspace_bytes, stream, nxg::kPrepare | nxg::kStrands);
}

uint32_t nxg_generate_packed_noalloc(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                                     const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                                     const nxg_packed_output *output, const nxg_stencil_output *stencil,
                                     const nxg_launch_config *config, void *stream) {
    // the request is checked in full before the scratch is looked up or used
    if (uint32_t s = nxg_validate_packed_generation(asset, deformed, params, output, config); s != NXG_OK) return s;
    const uint64_t need = nxg_packed_generation_workspace_size(&asset->header, params);
    if (need > nxg::kStaticScratchBytes) return NXG_ERROR_SCRATCH_LIMIT;
    uint32_t status = NXG_OK;
    void *ws = nxg::device_scratch(&status);
    if (!ws) return status;
    // the packed workspace needs no clearing: the guide pass and the strand
    // order overwrite every word the strand kernel reads
    return nxg::launch(asset, deformed, params, octaves, nullptr, nullptr, 0, output, stencil, config, ws,
                       nxg::kStaticScratchBytes, stream, nxg::kPrepare | nxg::kStrands);
}

uint32_t nxg_prepare_packed_generation(const nxg_device_asset *asset,
                                       const nxg_deformed_geometry *deformed,
                                       const nxg_generation_params *params, void *workspace,
                                       uint64_t workspace_bytes, void *stream) {
    return nxg::launch(asset, deformed, params, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr, workspace,
                       workspace_bytes, stream, nxg::kPrepare);
}

uint32_t nxg_generate_packed_prepared(const nxg_device_asset *asset,
                                      const nxg_deformed_geometry *deformed,
                                      const nxg_generation_params *params,
                                      const nxg_noise_octaves *octaves,
                                      const nxg_packed_output *output,
                                      const nxg_stencil_output *stencil,
                                      const nxg_launch_config *config, void *workspace,
                                      uint64_t workspace_bytes, void *stream) {
    return nxg::launch(asset, deformed, params, octaves, nullptr, nullptr, 0, output, stencil, config, workspace,
                       workspace_bytes, stream, nxg::kStrands);
}

uint32_t nxg_generate_packed_from_roots(const nxg_device_asset *asset,
                                        const nxg_generation_params *params,
                                        const nxg_noise_octaves *octaves,
                                        const nxg_root_sample *roots,
                                        const nxg_packed_output *output,
                                        const nxg_launch_config *config, void *workspace,
                                        uint64_t workspace_bytes, void *stream) {
    if (!roots) return NXG_ERROR_INVALID_ROOTS;
    return nxg::launch(asset, nullptr, params, octaves, roots, nullptr, 0, output, nullptr, config, workspace,
                       workspace_bytes, stream, nxg::kPrepare | nxg::kStrands);
}

uint64_t nxg_cdf_sampler_size(const nxg_asset_header *header) {
    if (!header || header->triangle_count == 0u) return 0u;
    return nxg::cdf_layout(header->triangle_count).total_bytes;
}

uint32_t nxg_cdf_sampler_build(const nxg_device_asset *asset, void *sampler, uint64_t sampler_bytes,
                               void *stream) {
    return nxg::cdf_build(asset, sampler, sampler_bytes, static_cast<hipStream_t>(stream));
}

uint32_t nxg_generate_packed_cdf(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                                 const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                                 const void *sampler, uint64_t sampler_bytes, const nxg_packed_output *output,
                                 const nxg_stencil_output *stencil, const nxg_launch_config *config,
                                 void *workspace, uint64_t workspace_bytes, void *stream) {
    return nxg::generate_cdf(asset, deformed, params, octaves, sampler, sampler_bytes, output, stencil, config,
                             workspace, workspace_bytes, static_cast<hipStream_t>(stream));
}

void nxg_batch_default_tuning(nxg_batch_tuning *tuning) {
    if (tuning) *tuning = nxg::default_tuning();
}

uint32_t nxg_batch_validate(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning) {
    static_assert(sizeof(nxg::BatchArgs) <= 4096u, "batch kernel arguments exceed 4 KiB");
    uint32_t counter_stub = 0u;  // validation only, never dereferenced
    nxg::BatchLaunch b;
    return nxg::make_batch(items, count, tuning, &counter_stub, b);
}

uint32_t nxg_generate_batch(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning,
                            uint32_t *counter, void *stream) {
    nxg::BatchLaunch b;
    if (uint32_t s = nxg::make_batch(items, count, tuning, counter, b); s != NXG_OK) return s;
    return nxg::launch_batch(b, static_cast<hipStream_t>(stream));
}

struct nxg_batch_graph_handle {
    hipGraph_t graph;
    hipGraphExec_t exec;
};

uint32_t nxg_batch_graph_create(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning,
                                uint32_t *counter, void **graph) {
    if (!graph) return NXG_ERROR_NULL_ARGUMENT;
    *graph = nullptr;
    nxg::BatchLaunch b;
    if (uint32_t s = nxg::make_batch(items, count, tuning, counter, b); s != NXG_OK) return s;
    // the capture stream and the side streams, all at the default priority
    hipStream_t s = nullptr;
    nxg::BatchSides sides{};
    bool made = hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
                hipEventCreateWithFlags(&sides.fork, hipEventDisableTiming) == hipSuccess;
    for (uint32_t i = 0; made && i < nxg::BatchSides::kCount; ++i)
        made = hipStreamCreateWithFlags(&sides.stream[i], hipStreamNonBlocking) == hipSuccess &&
               hipEventCreateWithFlags(&sides.join[i], hipEventDisableTiming) == hipSuccess;
    hipGraph_t g = nullptr;
    uint32_t status = made ? NXG_OK : NXG_ERROR_HIP;
    if (status == NXG_OK && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        status = NXG_ERROR_HIP;
    } else if (status == NXG_OK) {
        status = nxg::launch_batch(b, s, &sides);
        if (hipStreamEndCapture(s, &g) != hipSuccess) status = NXG_ERROR_HIP;
    }
    for (uint32_t i = 0; i < nxg::BatchSides::kCount; ++i) {
        if (sides.join[i]) (void)hipEventDestroy(sides.join[i]);
        if (sides.stream[i]) (void)hipStreamDestroy(sides.stream[i]);
    }
    if (sides.fork) (void)hipEventDestroy(sides.fork);
    if (s) (void)hipStreamDestroy(s);
    hipGraphExec_t exec = nullptr;
    if (status == NXG_OK && hipGraphInstantiate(&exec, g, nullptr, nullptr, 0) != hipSuccess) status = NXG_ERROR_HIP;
    if (status != NXG_OK) {
        if (g) (void)hipGraphDestroy(g);
        return status;
    }
    *graph = new nxg_batch_graph_handle{g, exec};
    return NXG_OK;
}

uint32_t nxg_batch_graph_launch(void *graph, void *stream) {
    if (!graph) return NXG_ERROR_NULL_ARGUMENT;
    auto *h = static_cast<nxg_batch_graph_handle *>(graph);
    return hipGraphLaunch(h->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

uint32_t nxg_batch_graph_destroy(void *graph) {
    if (!graph) return NXG_OK;
    auto *h = static_cast<nxg_batch_graph_handle *>(graph);
    uint32_t status = NXG_OK;
    if (hipGraphExecDestroy(h->exec) != hipSuccess) status = NXG_ERROR_HIP;
    if (hipGraphDestroy(h->graph) != hipSuccess) status = NXG_ERROR_HIP;
    delete h;
    return status;
}

uint32_t nxg_batch_tuning_candidates(nxg_batch_tuning *candidates, uint32_t capacity) {
    uint32_t n = 0u;
    for (const auto &wc : nxg::kBatchWaveChunk)
        for (uint32_t claim : {1u, 2u, 4u, 8u})
            for (uint32_t order : {NXG_BATCH_ORDER_GIVEN, NXG_BATCH_ORDER_HEAVY_FIRST}) {
                if (candidates && n < capacity) candidates[n] = nxg_batch_tuning{wc[0], wc[1], claim, 0u, order};
                ++n;
            }
    return n;
}

uint32_t nxg_batch_autotune(const nxg_batch_item *items, uint32_t count, uint32_t *counter, void *stream,
                            const nxg_batch_tuning *candidates, uint32_t candidate_count, uint32_t repeats,
                            nxg_batch_tuning *best, float *times_ms) {
    if (!candidates || !best || !times_ms || candidate_count == 0u || repeats == 0u || repeats > 1000u)
        return NXG_ERROR_INVALID_BATCH;
    hipStream_t s = static_cast<hipStream_t>(stream);
    hipEvent_t start = nullptr, stop = nullptr;
    if (hipEventCreate(&start) != hipSuccess) return NXG_ERROR_HIP;
    if (hipEventCreate(&stop) != hipSuccess) {
        (void)hipEventDestroy(start);
        return NXG_ERROR_HIP;
    }
    uint32_t status = NXG_OK;
    float best_ms = 0.0f;
    bool any = false;
    float *samples = new float[repeats];
    for (uint32_t c = 0; c < candidate_count && status == NXG_OK; ++c) {
        nxg::BatchLaunch b;
        times_ms[c] = -1.0f;  // invalid candidate
        if (nxg::make_batch(items, count, &candidates[c], counter, b) != NXG_OK) continue;
        status = nxg::launch_batch(b, s);  // warm-up
        for (uint32_t r = 0; r < repeats && status == NXG_OK; ++r) {
            if (hipEventRecord(start, s) != hipSuccess) status = NXG_ERROR_HIP;
            if (status == NXG_OK) status = nxg::launch_batch(b, s);
            if (status == NXG_OK && (hipEventRecord(stop, s) != hipSuccess || hipEventSynchronize(stop) != hipSuccess ||
                                     hipEventElapsedTime(&samples[r], start, stop) != hipSuccess))
                status = NXG_ERROR_HIP;
        }
        if (status != NXG_OK) break;
        std::sort(samples, samples + repeats);
        times_ms[c] = samples[repeats / 2u];
        if (!any || times_ms[c] < best_ms) {
            best_ms = times_ms[c];
            *best = candidates[c];
            any = true;
        }
    }
    delete[] samples;
    (void)hipEventDestroy(start);
    (void)hipEventDestroy(stop);
    if (status == NXG_OK && !any) status = NXG_ERROR_INVALID_BATCH;
    return status;
}

uint32_t nxg_generate_motion(const nxg_device_asset *asset, const nxg_motion_sample *samples,
                             uint32_t sample_count, const nxg_generation_params *params,
                             const nxg_noise_octaves *octaves, const nxg_motion_output *output,
                             const nxg_launch_config *config, void *workspace, uint64_t workspace_bytes,
                             void *stream) {
    return nxg::launch_motion(asset, samples, sample_count, params, octaves, output, config, workspace,
                              workspace_bytes, static_cast<hipStream_t>(stream));
}

const char *nxg_build_info(void) {
    return "gfx950; -O3 -ffp-contract=off -fhip-fp32-correctly-rounded-divide-sqrt; "
           "no packed-fp32; wave64 strand tiles";
}

}  // extern "C"
