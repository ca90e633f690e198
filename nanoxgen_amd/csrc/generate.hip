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

#include <cstdint>

#include "../../include/nanoxgen_amd.h"
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
    seed ^= (seed << 7u) & 0x9d2c5680u;
    seed ^= (seed << 15u) & 0xefc60000u;
    seed ^= seed >> 18u;
    return (((seed & 0x00ff0000u) >> 4u) + (seed & 0xffu)) & 0xffu;
}

__device__ __forceinline__ float s_curve(float t) {  // generate.h:366-368
    return t * t * t * (t * (t * 6.0f - 15.0f) + 10.0f);
}

__device__ __forceinline__ float lerp_ref(float a, float b, float alpha) {
    return (1.0f - alpha) * a + alpha * b;  // generate.h:397-398
}

// One lattice coordinate of the fixed (non-moving) axes of a noise sample.
struct Axis {
    int32_t i;
    float w, alpha;
};
__device__ __forceinline__ Axis make_axis(float v) {
    const float f = floorf(v);
    Axis a;
    a.i = static_cast<int32_t>(f);
    a.w = v - f;
    a.alpha = s_curve(a.w);
    return a;
}

// Per-strand constants of the three noise samples (x, y or z moving).
struct NoiseDomain {
    float d[3];
    Axis fixed[3];
    uint32_t base[3][4];  // per moving axis: LCG constant of the 4 fixed corners
};

__device__ __forceinline__ NoiseDomain make_domain(V3 d) {
    NoiseDomain n;
    n.d[0] = d.x;
    n.d[1] = d.y;
    n.d[2] = d.z;
    n.fixed[0] = make_axis(d.x);
    n.fixed[1] = make_axis(d.y);
    n.fixed[2] = make_axis(d.z);
    const uint32_t X0 = static_cast<uint32_t>(n.fixed[0].i), Y0 = static_cast<uint32_t>(n.fixed[1].i),
                   Z0 = static_cast<uint32_t>(n.fixed[2].i);
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const uint32_t o1 = c & 1, o2 = (c >> 1) & 1;
        n.base[0][c] = (Y0 + o1) * kLcgA + (Z0 + o2) + kLcgK;           // x moves: (oy, oz)
        n.base[1][c] = (X0 + o1) * kLcgA2 + (Z0 + o2) + kLcgK;          // y moves: (ox, oz)
        n.base[2][c] = (X0 + o1) * kLcgA2 + (Y0 + o2) * kLcgA + kLcgK;  // z moves: (ox, oy)
    }
    return n;
}

// gradient_noise (generate.h:370-402) at the domain point displaced by
// `distance` along axis A, with the gradient table in LDS.
template <int A>
__device__ __forceinline__ float gradient_noise_axis(const NoiseDomain &n, float distance,
                                                     const float4 *grad) {
    const Axis m = make_axis(n.d[A] + distance);
    const uint32_t mult = A == 0 ? kLcgA2 : (A == 1 ? kLcgA : 1u);
    const uint32_t h0 = static_cast<uint32_t>(m.i) * mult;
    // weights[axis] - offset and alpha per axis, in x, y, z order
    float w[3][2], alpha[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
        const Axis a = ax == A ? m : n.fixed[ax];
        w[ax][0] = a.w - 0.0f;
        w[ax][1] = a.w - 1.0f;
        alpha[ax] = a.alpha;
    }
    float v[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int ox = c & 1, oy = (c >> 1) & 1, oz = (c >> 2) & 1;
        const int om = A == 0 ? ox : (A == 1 ? oy : oz);
        const int f1 = A == 0 ? oy : ox;
        const int f2 = A == 2 ? oy : oz;
        const uint32_t seed = h0 + (om ? mult : 0u) + n.base[A][f1 | (f2 << 1)];
        const float4 g = grad[temper_index(seed)];
        v[c] = g.x * w[0][ox] + g.y * w[1][oy] + g.z * w[2][oz];
    }
    // x, then y, then z reduction, exactly as the reference's loop nest
    v[0] = lerp_ref(v[0], v[1], alpha[0]);
    v[2] = lerp_ref(v[2], v[3], alpha[0]);
    v[4] = lerp_ref(v[4], v[5], alpha[0]);
    v[6] = lerp_ref(v[6], v[7], alpha[0]);
    v[0] = lerp_ref(v[0], v[2], alpha[1]);
    v[4] = lerp_ref(v[4], v[6], alpha[1]);
    v[0] = lerp_ref(v[0], v[4], alpha[2]);
    return 0.5f * v[0] + 0.5f;
}

__device__ __forceinline__ V3 rotate_by(V3 v, V3 a, float angle) {  // generate.h:404-418
    const float c = cosf(angle), s = sinf(angle), omc = 1.0f - c;
    return {dot(v, v3(a.x * a.x * omc + c, a.x * a.y * omc - a.z * s, a.x * a.z * omc + a.y * s)),
            dot(v, v3(a.x * a.y * omc + a.z * s, a.y * a.y * omc + c, a.y * a.z * omc - a.x * s)),
            dot(v, v3(a.x * a.z * omc - a.y * s, a.y * a.z * omc + a.x * s, a.z * a.z * omc + c))};
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
    uint32_t strand_count, cvs, seed, triangle_count, guide_count, has_normals;
    float support_scale, weight_power, rejection_cos, length_scale;
    float root_width, tip_width, radius_scale;
    float noise_mask, decorrelation_scale, preserve;
    float amp[NXG_MAX_NOISE_OCTAVES];
    float freq[NXG_MAX_NOISE_OCTAVES];
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
    uint32_t nearest;
};

struct Root {
    V3 position, normal;
    float u, v;
    uint32_t tri;
    float b1, b2;
};

// sample_root (generate.h:101-135)
__device__ __forceinline__ Root sample_root(const Args &a, uint32_t strand) {
    const float choose = random01(a.seed, strand, 0u) * static_cast<float>(a.triangle_count);
    uint32_t bucket = static_cast<uint32_t>(choose);
    if (bucket >= a.triangle_count) bucket = a.triangle_count - 1u;
    const nxg_alias_entry e = a.alias[bucket];
    uint32_t tri = random01(a.seed, strand, 1u) < e.probability ? bucket : e.alias;
    // validate_asset does not range-check alias targets; a corrupt entry
    // falls back to its own bucket instead of reading past the triangles
    if (tri >= a.triangle_count) tri = bucket;
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
        const float w = powf(1.0f - sqrtf(d2) / radius, fmax_ref(a.weight_power, 0.0f));
        b.idx[j] = gi;
        b.w[j] = w;
        b.sum += w;
    }
    return b;
}

// evaluate_base_strand_cv (generate.h:246-267, 300-310) from the offset table
__device__ __forceinline__ V3 base_cv(const Args &a, const Root &r, const Blend &b, uint32_t cv) {
    V3 off = v3(0.0f, 0.0f, 0.0f);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        if (b.idx[j] == NXG_INVALID_INDEX) continue;
        const float4 o = a.guide_offsets[uint64_t{b.idx[j]} * a.cvs + cv];
        off = off + v3(o.x, o.y, o.z) * b.w[j];
    }
    if (b.sum > 1.0e-12f) {
        off = off / b.sum;
    } else if (b.nearest != NXG_INVALID_INDEX) {
        const float4 o = a.guide_offsets[uint64_t{b.nearest} * a.cvs + cv];
        off = v3(o.x, o.y, o.z);
    } else {
        const float t = static_cast<float>(cv) / static_cast<float>(a.cvs - 1u);
        off = r.normal * t;
    }
    return r.position + off * a.length_scale;
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
    V3 prev, cur, next, tangent, tn;
    float arc;
};

template <int OCT>
struct NoiseState {
    NoiseDomain dom;
    float freq[OCT];
    float mag[OCT];  // noise_mask * amplitude
};

// evaluate_and_advance_noise_cursor (generate.h:459-515); OCT octaves share
// the transported frame and the arc length and their displacements add.
template <int OCT>
__device__ __forceinline__ V3 advance(const Args &a, const Root &r, const Blend &b,
                                      const NoiseState<OCT> &ns, Cursor &c, const float4 *grad) {
    if (c.cv > 0u) c.arc += sqrtf(len2(c.cur - c.prev));
    V3 nt = c.tangent;
    if (c.cv + 1u < a.cvs) {
        const V3 seg = c.next - c.cur;
        if (len2(seg) > 0.0f) nt = normalize(seg);
    }
    const V3 axis = cross(c.tangent, nt);
    if (len2(axis) > 0.0f) {
        const float angle = acosf(clamp_ref(dot(c.tangent, nt), -1.0f, 1.0f));
        c.tn = normalize(rotate_by(c.tn, axis, angle));
    }
    const V3 normal = c.tn;
    const V3 binormal = cross(normal, nt);
    const V3 tangent = cross(binormal, normal);
    V3 point = c.cur;
    if (c.cv > 0u) {
        const float t = static_cast<float>(c.cv) / static_cast<float>(a.cvs - 1u);
        float lx = 0.0f, ly = 0.0f, lz = 0.0f;
#pragma unroll
        for (int k = 0; k < OCT; ++k) {
            const float distance = c.arc * ns.freq[k];
            const float magnitude = ns.mag[k] * t;
            const float nx = (gradient_noise_axis<0>(ns.dom, distance, grad) - 0.5f) * magnitude;
            const float ny = (gradient_noise_axis<1>(ns.dom, distance, grad) - 0.5f) * magnitude;
            const float nz = (gradient_noise_axis<2>(ns.dom, distance, grad) - 0.5f) * magnitude;
            if (k == 0) {
                lx = nx;
                ly = ny;
                lz = nz;
            } else {
                lx += nx;
                ly += ny;
                lz += nz;
            }
        }
        point = point + normal * lx + binormal * ly + tangent * lz;
    }
    c.tangent = nt;
    c.prev = c.cur;
    if (c.cv + 1u < a.cvs) {
        c.cur = c.next;
        if (c.cv + 2u < a.cvs) c.next = base_cv(a, r, b, c.cv + 2u);
    }
    ++c.cv;
    return point;
}

template <int OCT>
__device__ __forceinline__ Cursor make_cursor(const Args &a, const Root &r, const Blend &b, V3 surface_u) {
    Cursor c;
    c.cv = 0u;
    c.cur = base_cv(a, r, b, 0u);
    c.prev = c.cur;
    c.next = base_cv(a, r, b, 1u);
    c.tangent = r.normal;
    c.tn = surface_u;
    c.arc = 0.0f;
    return c;
}

// guide state and offset table: one lane per (guide, cv)
__global__ void __launch_bounds__(256) guide_prep_kernel(
    const nxg_guide_record *guides, const nxg_vec3 *guide_cvs, const nxg_uint3 *triangles,
    const nxg_vec3 *d_positions, const nxg_vec3 *d_normals, const nxg_vec3 *d_guide_cvs,
    const nxg_vec3 *frame_positions, uint32_t guide_count, uint32_t triangle_count,
    uint32_t cvs, float4 *state, float4 *offsets) {
    const uint64_t i = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x;
    if (i >= uint64_t{guide_count} * cvs) return;
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

// Writes the wave's staged chunk [cv0, cv0 + n) for its 64 strands as
// contiguous runs of n float4 per strand.
__device__ __forceinline__ void flush_chunk(const Args &a, const float *stage, uint64_t wave_first,
                                            uint32_t cv0, uint32_t n, uint32_t lane) {
    for (uint32_t it = 0; it < n; ++it) {
        const uint32_t f = it * kWave + lane;
        const uint32_t s = f / n, k = f % n;
        const uint64_t strand = wave_first + s;
        if (strand < a.strand_count) {
            float4 p;
            p.x = stage[(k * 4u + 0u) * kStageStride + s];
            p.y = stage[(k * 4u + 1u) * kStageStride + s];
            p.z = stage[(k * 4u + 2u) * kStageStride + s];
            p.w = stage[(k * 4u + 3u) * kStageStride + s];
            reinterpret_cast<float4 *>(a.points)[strand * a.cvs + cv0 + k] = p;
        }
    }
}

template <bool NOISE, int OCT>
__global__ void __launch_bounds__(256) strand_kernel(Args a) {
    __shared__ float4 grad[NXG_NOISE_GRADIENT_COUNT];
    __shared__ float stage_all[256 / kWave][kChunk * 4 * kStageStride];
    for (uint32_t i = threadIdx.x; i < NXG_NOISE_GRADIENT_COUNT; i += blockDim.x) {
        const nxg_vec3 g = a.gradients[i];
        grad[i] = make_float4(g.x, g.y, g.z, 0.0f);
    }
    __syncthreads();
    const uint32_t lane = threadIdx.x % kWave;
    const uint32_t wave = threadIdx.x / kWave;
    float *stage = stage_all[wave];
    const uint64_t wave_first = uint64_t{blockIdx.x} * blockDim.x + wave * kWave;
    const uint64_t strand64 = wave_first + lane;
    const bool live = strand64 < a.strand_count;
    // lanes past the end compute the last strand and store nothing, so every
    // lane runs the same CV loop and reaches every barrier
    const uint32_t strand = live ? static_cast<uint32_t>(strand64) : a.strand_count - 1u;

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
    const Blend b = make_blend(a, r);
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
        if (a.point_counts) a.point_counts[strand] = a.cvs;
        if (a.stencil_indices) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                a.stencil_indices[uint64_t{strand} * 8u + j] = b.idx[j];
                a.stencil_weights[uint64_t{strand} * 8u + j] = b.w[j];
            }
        }
    }

    NoiseState<OCT> ns;
    V3 surface_u = v3(0.0f, 0.0f, 0.0f);
    float base_length = 0.0f;
    if constexpr (NOISE) {
        surface_u = root_surface_u(a, r);
        // base_strand_length (generate.h:420-433)
        V3 prev = base_cv(a, r, b, 0u);
        for (uint32_t cv = 1u; cv < a.cvs; ++cv) {
            const V3 c = base_cv(a, r, b, cv);
            base_length += sqrtf(len2(c - prev));
            prev = c;
        }
#pragma unroll
        for (int k = 0; k < OCT; ++k) {
            ns.freq[k] = base_length > 0.0f ? fmax_ref(0.5f / base_length, a.freq[k]) : a.freq[k];
            ns.mag[k] = a.noise_mask * a.amp[k];
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
        if constexpr (NOISE) cur = make_cursor<OCT>(a, r, b, surface_u);
        V3 last = v3(0.0f, 0.0f, 0.0f);
        float noisy = 0.0f;
        for (uint32_t cv0 = 0; cv0 < a.cvs; cv0 += kChunk) {
            const uint32_t n = a.cvs - cv0 < kChunk ? a.cvs - cv0 : kChunk;
#pragma unroll 1
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t cv = cv0 + k;
                V3 p;
                if constexpr (NOISE) {
                    p = advance<OCT>(a, r, b, ns, cur, grad);
                } else {
                    p = base_cv(a, r, b, cv);
                }
                if (!emit) {
                    if (cv > 0u) noisy += sqrtf(len2(p - last));
                    last = p;
                    continue;
                }
                if (preserve_scale != 1.0f) p = r.position + (p - r.position) * preserve_scale;
                // evaluate_strand_width (generate.h:577-581) and the packed radius
                const float t = static_cast<float>(cv) / static_cast<float>(a.cvs - 1u);
                const float width = a.root_width * (1.0f - t) + a.tip_width * t;
                stage[(k * 4u + 0u) * kStageStride + lane] = p.x;
                stage[(k * 4u + 1u) * kStageStride + lane] = p.y;
                stage[(k * 4u + 2u) * kStageStride + lane] = p.z;
                stage[(k * 4u + 3u) * kStageStride + lane] = 0.5f * width * a.radius_scale;
            }
            if (emit) {
                __syncthreads();
                flush_chunk(a, stage, wave_first, cv0, n, lane);
                __syncthreads();
            }
        }
        if (!emit && noisy > 0.0f) {  // generate.h:553-563
            const float target = base_length * a.preserve + noisy * (1.0f - a.preserve);
            const float diff = noisy - target;
            if ((diff < 0.0f ? -diff : diff) >= 0.0001f) preserve_scale = target / noisy;
        }
    }
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

uint32_t launch(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                const nxg_root_sample *explicit_roots, const nxg_packed_output *output,
                const nxg_stencil_output *stencil, const nxg_launch_config *config, void *workspace,
                uint64_t workspace_bytes, void *stream_ptr, uint32_t stages) {
    if (uint32_t s = validate_common(asset, deformed, params, config); s != NXG_OK) return s;
    if (stages & kStrands) {
        if (uint32_t s = validate_packed_output(*params, output); s != NXG_OK) return s;
    }
    if (!valid_octaves(octaves)) return NXG_ERROR_INVALID_OCTAVES;
    if (stencil && (stencil->indices || stencil->weights)) {
        if (!stencil->indices || !stencil->weights ||
            stencil->capacity < uint64_t{params->strand_count} * 8u)
            return NXG_ERROR_STENCIL_CAPACITY_TOO_SMALL;
    }
    const nxg_asset_header &h = asset->header;
    const WorkspaceLayout L = workspace_layout(h.guide_count, params->cvs_per_strand);
    if (!workspace || workspace_bytes < L.total) return NXG_ERROR_WORKSPACE_TOO_SMALL;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);

    const nxg_vec3 *d_pos = deformed ? deformed->positions : nullptr;
    const nxg_vec3 *d_nrm = deformed ? deformed->normals : nullptr;
    const nxg_vec3 *d_gcv = deformed ? deformed->guide_cvs : nullptr;
    const nxg_vec3 *rest_pos = section<nxg_vec3>(*asset, h.positions_offset);
    const nxg_vec3 *frame_pos = d_pos ? d_pos : rest_pos;
    float4 *state = reinterpret_cast<float4 *>(static_cast<uint8_t *>(workspace) + L.guide_state_offset);
    float4 *offsets = reinterpret_cast<float4 *>(static_cast<uint8_t *>(workspace) + L.guide_offsets_offset);

    if (stages & kPrepare) {
        const uint64_t prep_lanes = uint64_t{h.guide_count} * params->cvs_per_strand;
        hipLaunchKernelGGL(guide_prep_kernel, dim3(static_cast<uint32_t>((prep_lanes + 255u) / 256u)),
                           dim3(256), 0, stream, section<nxg_guide_record>(*asset, h.guides_offset),
                           section<nxg_vec3>(*asset, h.guide_cvs_offset),
                           section<nxg_uint3>(*asset, h.triangles_offset), d_pos, d_nrm, d_gcv,
                           frame_pos, h.guide_count, h.triangle_count, params->cvs_per_strand, state,
                           offsets);
        if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    }
    if (!(stages & kStrands)) return NXG_OK;

    Args a{};
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
    a.radius_scale = output->radius_scale;
    a.noise_mask = params->noise_mask;
    const float decorrelation = 1.0f - params->noise_correlation;
    a.decorrelation_scale = 100.0f * decorrelation * decorrelation;
    a.preserve = params->noise_preserve_length;
    uint32_t oct = 1u;
    if (octaves && octaves->count > 1u) {
        oct = octaves->count;
        for (uint32_t k = 0; k < oct; ++k) {
            a.amp[k] = octaves->amplitude[k];
            a.freq[k] = octaves->frequency[k];
        }
    } else {
        a.amp[0] = params->noise_amplitude;
        a.freq[0] = params->noise_frequency;
    }
    a.points = output->points;
    a.roots = output->roots;
    a.root_uvs = output->root_uvs;
    a.point_counts = output->point_counts;
    a.stencil_indices = stencil ? stencil->indices : nullptr;
    a.stencil_weights = stencil ? stencil->weights : nullptr;

    const dim3 grid((params->strand_count + 255u) / 256u), block(256);
    if (!noise_enabled(*params, octaves)) {
        hipLaunchKernelGGL((strand_kernel<false, 1>), grid, block, 0, stream, a);
    } else {
        switch (oct) {
            case 1: hipLaunchKernelGGL((strand_kernel<true, 1>), grid, block, 0, stream, a); break;
            case 2: hipLaunchKernelGGL((strand_kernel<true, 2>), grid, block, 0, stream, a); break;
            case 3: hipLaunchKernelGGL((strand_kernel<true, 3>), grid, block, 0, stream, a); break;
            default: hipLaunchKernelGGL((strand_kernel<true, 4>), grid, block, 0, stream, a); break;
        }
    }
    return hipGetLastError() == hipSuccess ? NXG_OK : NXG_ERROR_HIP;
}

}  // namespace
}  // namespace nxg

extern "C" {

uint32_t nxg_generate_packed(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                             const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                             const nxg_packed_output *output, const nxg_stencil_output *stencil,
                             const nxg_launch_config *config, void *workspace,
                             uint64_t workspace_bytes, void *stream) {
    return nxg::launch(asset, deformed, params, octaves, nullptr, output, stencil, config, workspace,
                       workspace_bytes, stream, nxg::kPrepare | nxg::kStrands);
}

uint32_t nxg_prepare_packed_generation(const nxg_device_asset *asset,
                                       const nxg_deformed_geometry *deformed,
                                       const nxg_generation_params *params, void *workspace,
                                       uint64_t workspace_bytes, void *stream) {
    return nxg::launch(asset, deformed, params, nullptr, nullptr, nullptr, nullptr, nullptr, workspace,
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
    return nxg::launch(asset, deformed, params, octaves, nullptr, output, stencil, config, workspace,
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
    return nxg::launch(asset, nullptr, params, octaves, roots, output, nullptr, config, workspace,
                       workspace_bytes, stream, nxg::kPrepare | nxg::kStrands);
}

const char *nxg_build_info(void) {
    return "gfx950; -O3 -ffp-contract=off -fhip-fp32-correctly-rounded-divide-sqrt; "
           "no packed-fp32; wave64 strand tiles";
}

}  // extern "C"
