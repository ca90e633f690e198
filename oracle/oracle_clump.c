/*
 * ORACLE (test infrastructure only): the Classic ClumpingFX module of the
 * reference's float runtime restated in C, for a world-space guide binding
 * (no root-relative axes) and module expressions that are constants (the
 * amount optionally times a per-strand multiplier):
 *   prepare_xgen_classic_clump_runtime_data  src/xgen_classic_runtime.cpp:839-890
 *                                            (render axes, guide spline lengths)
 *   cut_to_length_and_rebuild                src/xgen_classic_runtime.cpp:750-799
 *   build_clump_noise_axis                   src/xgen_classic_runtime.cpp:1041-1186
 *   xgen_classic_noise_float                 src/xgen_classic_runtime.cpp:802-837
 *   per-strand clump step                    src/xgen_classic_runtime.cpp:1680-1793
 * Float operations in the reference's order, built with -ffp-contract=off;
 * acosf/cosf/sinf are the C library's, as the reference's std:: calls are.
 */
#include <math.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

typedef struct { float x, y, z; } v3;

static v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static v3 mul(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static v3 divs(v3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
static float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float len2(v3 v) { return dot(v, v); }
static v3 cross(v3 a, v3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static v3 normalize(v3 v) { /* generate.h:67-70 */
    const float l2 = len2(v);
    return l2 > 1.0e-20f ? divs(v, sqrtf(l2)) : mk(0.0f, 1.0f, 0.0f);
}
static v3 fallback_surface_u(v3 n) { /* generate.h:312-316 */
    const v3 axis = n.z > -0.999f && n.z < 0.999f ? mk(0.0f, 0.0f, 1.0f) : mk(0.0f, 1.0f, 0.0f);
    return normalize(cross(n, axis));
}
static float max_std(float a, float b) { return a < b ? b : a; }             /* std::max */
static float clamp_std(float x, float lo, float hi) { return x < lo ? lo : (hi < x ? hi : x); } /* std::clamp */

/* xgen_curve_eval over n points (src/xgen_classic_runtime.cpp:587-612) */
static v3 curve_eval(const v3 *p, uint32_t n, float t) {
    if (t < 1.0e-7f) return p[0];
    if (t > 1.0f - 1.0e-7f) return p[n - 1u];
    const uint32_t spans = n - 1u;
    const float scaled = t * (float)spans;
    const uint32_t span = (uint32_t)scaled;
    const float f = scaled - (float)span;
    const float f2 = f * f;
    const float f3 = f2 * f;
    const float omf = 1.0f - f;
    const float b0 = omf * omf * omf;
    const float b1 = 3.0f * f3 - 6.0f * f2 + 4.0f;
    const float b2 = -3.0f * f3 + 3.0f * f2 + 3.0f * f + 1.0f;
    const float b3 = f3;
    const v3 p0 = span == 0u ? sub(mul(p[0], 2.0f), p[1]) : p[span - 1u];
    const v3 p1 = p[span];
    const v3 p2 = p[span + 1u];
    const v3 p3 = span + 2u == n ? sub(mul(p[n - 1u], 2.0f), p[n - 2u]) : p[span + 2u];
    return mul(add(add(add(mul(p0, b0), mul(p1, b1)), mul(p2, b2)), mul(p3, b3)), 1.0f / 6.0f);
}

/* curve_spline_length over Vec3 points (src/xgen_classic_runtime.cpp:665-690) */
static float spline_length(const v3 *p, uint32_t n) {
    const uint32_t intervals = n * 2u + 4u;
    const float step = 1.0f / (float)intervals;
    v3 previous = p[0];
    float result = 0.0f;
    for (uint32_t k = 1u; k < intervals; ++k) {
        const v3 current = curve_eval(p, n, step * (float)k);
        result += sqrtf(len2(sub(current, previous)));
        previous = current;
    }
    result += sqrtf(len2(sub(p[n - 1u], previous)));
    return result;
}

void orc_clump_prepare(const float *axes3, uint32_t guides, uint32_t cvs, float *render3, float *lengths) {
    for (uint32_t g = 0; g < guides; ++g) {
        const v3 *src = (const v3 *)(axes3 + (uint64_t)g * cvs * 3u);
        v3 *dst = (v3 *)(render3 + (uint64_t)g * cvs * 3u);
        dst[0] = src[0];
        for (uint32_t cv = 1u; cv + 1u < cvs; ++cv) dst[cv] = curve_eval(src, cvs, (float)cv / (float)(cvs - 1u));
        dst[cvs - 1u] = src[cvs - 1u];
        lengths[g] = spline_length(src, cvs);
    }
}

/* rotate_between (src/xgen_classic_runtime.cpp:1109-1122) */
static v3 rotate_between(v3 value, v3 from, v3 to, float fraction) {
    v3 axis = cross(from, to);
    const float als = len2(axis);
    if (!(als >= 1.0e-12f)) return value;
    axis = mul(axis, 1.0f / sqrtf(als));
    const float angle = acosf(clamp_std(dot(from, to), -1.0f, 1.0f)) * fraction;
    const float c = cosf(angle);
    const float s = sinf(angle);
    return add(add(mul(value, c), mul(cross(axis, value), s)), mul(axis, dot(axis, value) * (1.0f - c)));
}

/* build_clump_noise_axis (src/xgen_classic_runtime.cpp:1041-1186) from the
 * evaluated noise frequency and correlation and the per-CV magnitudes
 * mask * noise * max(noiseScale(t), 0); magnitudes NULL: the zero axis (the
 * noise or mask test failed).  out[0..n). */
static void noise_axis_general(const v3 *axis, const v3 *distance_axis, uint32_t n, float raw_frequency,
                               float raw_correlation, v3 domain_position, v3 normal, v3 tangent,
                               const float *magnitudes, const nxg_vec3 *gradients, v3 *out) {
    for (uint32_t i = 0; i < n; ++i) out[i] = mk(0.0f, 0.0f, 0.0f);
    if (!magnitudes) return;
    const float frequency = max_std(raw_frequency, 0.0f);
    const float correlation = clamp_std(raw_correlation * 0.01f, 0.0f, 1.0f);
    float polyline = 0.0f;
    for (uint32_t cv = 1u; cv < n; ++cv) polyline += sqrtf(len2(sub(distance_axis[cv], distance_axis[cv - 1u])));
    const float effective = polyline > 0.0f ? max_std(0.5f / polyline, frequency) : frequency;
    const float decorrelation = 1.0f - correlation;
    const float domain_scale = 100.0f * decorrelation * decorrelation;
    const v3 domain = mul(add(domain_position, mk(0.419276f, 0.184247f, 0.805721f)), domain_scale);
    normal = !(len2(normal) > 1.0e-20f) ? mk(0.0f, 1.0f, 0.0f) : normalize(normal);
    tangent = !(len2(tangent) > 1.0e-20f) ? fallback_surface_u(normal) : normalize(tangent);
    v3 tu = tangent, tv = cross(normal, tangent), current = normal;
    const v3 first_segment = sub(axis[1], axis[0]);
    if (len2(first_segment) > 1.0e-20f) current = normalize(first_segment);
    if (dot(normal, current) < 0.99999f) {
        tu = rotate_between(tu, normal, current, 1.0f);
        tv = rotate_between(tv, normal, current, 1.0f);
    }
    float travelled = 0.0f;
    for (uint32_t cv = 1u; cv < n; ++cv) {
        travelled += sqrtf(len2(sub(distance_axis[cv], distance_axis[cv - 1u])));
        v3 su = tu, sv = tv, next = current;
        int advance = 0;
        if (cv + 1u < n) {
            const v3 segment = sub(axis[cv + 1u], axis[cv]);
            if (len2(segment) > 1.0e-20f) {
                next = normalize(segment);
                const float td = dot(current, next);
                if (fabsf(td) < 0.99999f) {
                    su = rotate_between(tu, current, next, 0.5f);
                    sv = rotate_between(tv, current, next, 0.5f);
                    advance = 1;
                } else {
                    current = next;
                }
            }
        }
        const float magnitude = magnitudes[cv];
        const float distance = travelled * effective;
        const float first = orc_gradient_noise_f32(gradients, domain.x + distance, domain.y, domain.z) - 0.5f;
        const float second = orc_gradient_noise_f32(gradients, domain.x, domain.y, domain.z + distance) - 0.5f;
        out[cv] = mul(add(mul(su, first), mul(sv, second)), magnitude);
        if (advance) {
            tu = rotate_between(tu, current, next, 1.0f);
            tv = rotate_between(tv, current, next, 1.0f);
            current = next;
        }
    }
}

/* build_clump_noise_axis with constant noise expressions: out[0..n) */
static void noise_axis(const v3 *axis, const v3 *distance_axis, uint32_t n, const orc_clump_params *p, v3 domain_position,
                       v3 normal, v3 tangent, float mask, const nxg_vec3 *gradients, v3 *out) {
    float magnitudes[256];
    const float noise = max_std(p->noise, 0.0f);
    const int on = noise > 1.0e-5f && mask > 1.0e-4f;
    for (uint32_t cv = 0; cv < n; ++cv) magnitudes[cv] = mask * noise * max_std(p->noise_scale, 0.0f);
    noise_axis_general(axis, distance_axis, n, p->noise_frequency, p->noise_correlation, domain_position, normal,
                       tangent, on ? magnitudes : NULL, gradients, out);
}

void orc_clump_noise_axis(const float *axis3, const float *distance3, uint32_t n, float raw_frequency,
                          float raw_correlation, const float *domain3, const float *normal3, const float *tangent3,
                          const float *magnitudes, const nxg_vec3 *gradients, float *out3) {
    noise_axis_general((const v3 *)axis3, (const v3 *)distance3, n, raw_frequency, raw_correlation,
                       *(const v3 *)domain3, *(const v3 *)normal3, *(const v3 *)tangent3, magnitudes, gradients,
                       (v3 *)out3);
}

/* cut_to_length_and_rebuild: points in place, n <= 256 */
static void cut_to_length(v3 *points, uint32_t n, float target);
void orc_cut_to_length(float *points3, uint32_t n, float target) { cut_to_length((v3 *)points3, n, target); }

static void cut_to_length(v3 *points, uint32_t n, float target) {
    v3 source[256];
    memcpy(source, points, sizeof(v3) * n);
    if (target <= 1.0e-10f) {
        for (uint32_t i = 0; i < n; ++i) points[i] = source[0];
        return;
    }
    const uint32_t intervals = n * 2u + 4u;
    const float step = 1.0f / (float)intervals;
    float previous_parameter = 0.0f, accumulated = 0.0f, cut = 1.0f;
    v3 previous = source[0];
    for (uint32_t sample = 1u; sample <= intervals; ++sample) {
        const float parameter = sample == intervals ? 1.0f : step * (float)sample;
        const v3 current = sample == intervals ? source[n - 1u] : curve_eval(source, n, parameter);
        const float segment = sqrtf(len2(sub(current, previous)));
        accumulated += segment;
        if (accumulated >= target && segment > 0.0f) {
            cut = parameter - ((accumulated - target) / segment) * (parameter - previous_parameter);
            break;
        }
        previous_parameter = parameter;
        previous = current;
    }
    const float rebuild = cut / (float)(n - 1u);
    points[0] = source[0];
    for (uint32_t i = 1u; i < n; ++i) points[i] = curve_eval(source, n, rebuild * (float)i);
}

typedef struct {
    float *points;
    uint32_t cvs;
    const uint32_t *strand_guides;
    const float *axes3, *render3, *lengths, *normals3, *tangents3, *reference3;
    uint32_t guides;
    const nxg_vec3 *gradients;
    const orc_clump_params *p;
    const float *amount_multiplier;
    uint32_t multiplier_stride;
    uint8_t *failed;
} clump_ctx;

static void clump_one(void *c, uint32_t s) {
    const clump_ctx *x = (const clump_ctx *)c;
    const orc_clump_params *p = x->p;
    const uint32_t cvs = x->cvs;
    v3 pts[256], noise[256];
    x->failed[s] = 0u;
    const uint32_t g = x->strand_guides[s];
    if (g == 0xffffffffu) return;
    if (g >= x->guides) {
        x->failed[s] = 1u;
        return;
    }
    const float mask = clamp_std(p->mask, 0.0f, 1.0f);
    if (mask <= 0.0f) return;
    float *io = x->points + (uint64_t)s * cvs * 4u;
    const v3 *render = (const v3 *)(x->render3 + (uint64_t)g * cvs * 3u);
    const v3 *source = (const v3 *)(x->axes3 + (uint64_t)g * cvs * 3u);
    const v3 root = source[0];
    const v3 domain_position = x->reference3 ? *(const v3 *)(x->reference3 + 3u * g) : root;
    noise_axis(render, source, cvs, p, domain_position, *(const v3 *)(x->normals3 + 3u * g),
               *(const v3 *)(x->tangents3 + 3u * g), mask, x->gradients, noise);
    for (uint32_t i = 0; i < cvs; ++i) pts[i] = mk(io[4 * i], io[4 * i + 1], io[4 * i + 2]);
    cut_to_length(pts, cvs, x->lengths[g]);
    float raw_amount = p->amount;
    if (x->amount_multiplier) raw_amount = raw_amount * x->amount_multiplier[(uint64_t)s * x->multiplier_stride];
    if (!isfinite(raw_amount)) {
        x->failed[s] = 1u;
        return;
    }
    const float amount = clamp_std(raw_amount, 0.0f, 1.0f);
    for (uint32_t cv = 1u; cv < cvs; ++cv) {
        const v3 goal = sub(render[cv], mk(0.0f, 0.0f, 0.0f));
        pts[cv] = add(pts[cv], mul(sub(goal, pts[cv]), mask * amount * (1.0f - 2.0f * p->clump_scale)));
    }
    for (uint32_t cv = 1u; cv < cvs; ++cv) pts[cv] = add(pts[cv], noise[cv]);
    for (uint32_t i = 0; i < cvs; ++i) {
        io[4 * i] = pts[i].x;
        io[4 * i + 1] = pts[i].y;
        io[4 * i + 2] = pts[i].z;
    }
}

/* Strands are independent: they run in parallel over `workers` threads
 * (0 = all cores).  Returns ~0u (points then undefined) where the
 * reference throws: a guide index out of range or a non-finite amount. */
uint32_t orc_clump_apply(float *points, uint32_t strand_count, uint32_t cvs, const uint32_t *strand_guides,
                         const float *axes3, const float *render3, const float *lengths, const float *normals3,
                         const float *tangents3, const float *reference3, uint32_t guides,
                         const nxg_vec3 *gradients, const orc_clump_params *p, const float *amount_multiplier,
                         uint32_t multiplier_stride, uint32_t workers) {
    if (cvs < 3u || cvs > 256u) return ~0u;
    uint8_t *failed = (uint8_t *)calloc(strand_count ? strand_count : 1u, 1u);
    if (!failed) return ~0u;
    clump_ctx c = {points, cvs, strand_guides, axes3, render3, lengths, normals3, tangents3, reference3,
                   guides, gradients, p, amount_multiplier, multiplier_stride, failed};
    orc_run_tiles(strand_count, workers, 64u, clump_one, &c);
    uint32_t bad = 0u;
    for (uint32_t s = 0; s < strand_count; ++s) bad |= failed[s];
    free(failed);
    return bad ? ~0u : 0u;
}

/* apply_noise's geometry (src/xgen_classic_runtime.cpp:914-1039) on n points
 * (xyz, in place) once its expressions are evaluated: the frame transported
 * from the surface normal along the segments, per CV the displacement of
 * magnitudes[cv] along normal, binormal and tangent by the classic noise of
 * the domain moved along each axis, then the preserve-length rescale. */
void orc_noise_fx(float *points3, uint32_t n, float frequency, float correlation, float preserve,
                  const float *domain_position3, const float *surface_normal3, const float *surface_tangent3,
                  const float *magnitudes, const nxg_vec3 *gradients) {
    v3 *pts = (v3 *)points3;
    float original = 0.0f; /* curve_polyline_length */
    for (uint32_t cv = 1u; cv < n; ++cv) original += sqrtf(len2(sub(pts[cv], pts[cv - 1u])));
    const float effective = original > 0.0f ? max_std(0.5f / original, frequency) : frequency;
    const float decorrelation = 1.0f - correlation;
    const float domain_scale = 100.0f * decorrelation * decorrelation;
    const v3 root = pts[0];
    const v3 dp = domain_position3 ? mk(domain_position3[0], domain_position3[1], domain_position3[2]) : root;
    const v3 domain = mul(add(dp, mk(0.419276f, 0.184247f, 0.805721f)), domain_scale);
    const v3 surface_normal = mk(surface_normal3[0], surface_normal3[1], surface_normal3[2]);
    v3 surface_tangent = mk(surface_tangent3[0], surface_tangent3[1], surface_tangent3[2]);
    surface_tangent = !(len2(surface_tangent) >= 1.0e-12f) ? fallback_surface_u(surface_normal)
                                                           : normalize(surface_tangent);
    v3 previous_base = root, current_base = root, next_base = pts[1];
    v3 prior = len2(surface_normal) >= 1.0e-12f ? normalize(surface_normal) : mk(0.0f, 1.0f, 0.0f);
    v3 transported = surface_tangent;
    float travelled = 0.0f;
    for (uint32_t cv = 0u; cv < n; ++cv) {
        if (cv > 0u) travelled += sqrtf(len2(sub(current_base, previous_base)));
        v3 next = prior;
        if (cv + 1u < n) {
            const v3 segment = sub(next_base, current_base);
            if (len2(segment) >= 1.0e-12f) next = normalize(segment);
        }
        const v3 axis = cross(prior, next);
        const float als = len2(axis);
        const float tangent_dot = clamp_std(dot(prior, next), -1.0f, 1.0f);
        if (als >= 1.0e-12f) {
            const v3 ra = mul(axis, 1.0f / sqrtf(als));
            const float angle = acosf(tangent_dot);
            const float cosine = cosf(angle), sine = sinf(angle);
            const v3 rotated = add(add(mul(transported, cosine), mul(cross(ra, transported), sine)),
                                   mul(ra, dot(ra, transported) * (1.0f - cosine)));
            transported = len2(rotated) >= 1.0e-12f ? normalize(rotated) : mk(0.0f, 0.0f, 0.0f);
        }
        const v3 normal = transported;
        v3 binormal = cross(normal, next);
        binormal = len2(binormal) >= 1.0e-12f ? normalize(binormal) : mk(0.0f, 0.0f, 0.0f);
        const v3 tangent = cross(binormal, normal);
        if (cv > 0u) {
            const float distance = travelled * effective;
            const float m = magnitudes[cv];
            const v3 local = mk((orc_gradient_noise_f32(gradients, domain.x + distance, domain.y, domain.z) - 0.5f) * m,
                                (orc_gradient_noise_f32(gradients, domain.x, domain.y + distance, domain.z) - 0.5f) * m,
                                (orc_gradient_noise_f32(gradients, domain.x, domain.y, domain.z + distance) - 0.5f) * m);
            pts[cv] = add(add(add(current_base, mul(normal, local.x)), mul(binormal, local.y)), mul(tangent, local.z));
        }
        prior = next;
        previous_base = current_base;
        if (cv + 1u < n) {
            current_base = next_base;
            if (cv + 2u < n) next_base = pts[cv + 2u];
        }
    }
    if (preserve > 0.001f) {
        float noisy = 0.0f;
        for (uint32_t cv = 1u; cv < n; ++cv) noisy += sqrtf(len2(sub(pts[cv], pts[cv - 1u])));
        if (noisy > 0.0f) {
            const float target = original * preserve + noisy * (1.0f - preserve);
            if (fabsf(noisy - target) >= 0.0001f) { /* scale_curve (:576-585) */
                const float scale = target / noisy;
                const v3 r0 = pts[0];
                for (uint32_t cv = 0u; cv < n; ++cv) pts[cv] = add(r0, mul(sub(pts[cv], r0), scale));
            }
        }
    }
}
