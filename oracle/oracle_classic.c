/*
 * ORACLE (test infrastructure only): the Classic CutFX chain of the
 * reference's float runtime restated in C, for the cut-from-tip form with
 * amount = amount_scale * cLength + amount_offset (times an optional
 * per-strand multiplier), no length or width expressions:
 *   xgen_curve_eval        src/xgen_classic_runtime.cpp:587-612, 614-642
 *   curve_spline_length    src/xgen_classic_runtime.cpp:644-663
 *   cut_from_tip_and_rebuild src/xgen_classic_runtime.cpp:692-743
 *   per-strand chain       src/xgen_classic_runtime.cpp:1648-1678 (cLength,
 *                          default width, 1e-4 culling), 1806-1821 (CutFX,
 *                          culling below cut parameter 1e-4)
 *   compaction             src/xgen_classic_runtime.cpp:1893-1918 (stable)
 *   renderer endpoints     src/xgen_classic_runtime.cpp:1921-1956
 * Float operations in the reference's order, built with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

typedef struct { float x, y, z; } cv3;

static cv3 c3(float x, float y, float z) { cv3 r = {x, y, z}; return r; }
static cv3 add3(cv3 a, cv3 b) { return c3(a.x + b.x, a.y + b.y, a.z + b.z); }
static cv3 sub3(cv3 a, cv3 b) { return c3(a.x - b.x, a.y - b.y, a.z - b.z); }
static cv3 mul3(cv3 a, float s) { return c3(a.x * s, a.y * s, a.z * s); }
static float len2(cv3 v) { return v.x * v.x + v.y * v.y + v.z * v.z; }
static float max_std(float a, float b) { return a < b ? b : a; } /* std::max */

/* xgen_curve_eval over n control points (uniform cubic B-spline with
 * reflected end points), src/xgen_classic_runtime.cpp:587-612 */
cv3 orc_curve_eval_v(const cv3 *p, uint32_t n, float t) {
    const float epsilon = 1.0e-7f;
    if (t < epsilon) return p[0];
    if (t > 1.0f - epsilon) return p[n - 1u];
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
    const cv3 p0 = span == 0u ? sub3(mul3(p[0], 2.0f), p[1]) : p[span - 1u];
    const cv3 p1 = p[span];
    const cv3 p2 = p[span + 1u];
    const cv3 p3 = span + 2u == n ? sub3(mul3(p[n - 1u], 2.0f), p[n - 2u]) : p[span + 2u];
    return mul3(add3(add3(add3(mul3(p0, b0), mul3(p1, b1)), mul3(p2, b2)), mul3(p3, b3)), 1.0f / 6.0f);
}

/* curve_spline_length: 2n+4 fixed intervals, raw end CVs
 * (src/xgen_classic_runtime.cpp:644-663) */
float orc_curve_spline_length_v(const cv3 *p, uint32_t n) {
    const uint32_t intervals = n * 2u + 4u;
    const float step = 1.0f / (float)intervals;
    cv3 previous = p[0];
    float result = 0.0f;
    for (uint32_t sample = 1u; sample < intervals; ++sample) {
        const cv3 current = orc_curve_eval_v(p, n, step * (float)sample);
        result += sqrtf(len2(sub3(current, previous)));
        previous = current;
    }
    result += sqrtf(len2(sub3(p[n - 1u], previous)));
    return result;
}

/* cut_from_tip_and_rebuild (src/xgen_classic_runtime.cpp:692-743): walks
 * back from the tip along 2n+4 spline steps until `amount` of length is
 * removed, then rebuilds n CVs uniformly on [0, cut]; returns the cut
 * parameter (0: everything cut, every CV at the root). */
float orc_cut_from_tip_v(cv3 *points, uint32_t n, float amount, cv3 *source) {
    memcpy(source, points, sizeof(cv3) * n);
    float cut = 1.0f;
    if (amount >= 1.0e-10f) {
        const float step = 1.0f / (2.0f * (float)n + 4.0f);
        float previous_parameter = 1.0f;
        cv3 previous = source[n - 1u];
        float accumulated = 0.0f;
        for (;;) {
            const float parameter = max_std(previous_parameter - step, 0.0f);
            const cv3 current = orc_curve_eval_v(source, n, parameter);
            const float segment = sqrtf(len2(sub3(previous, current)));
            accumulated += segment;
            if (accumulated >= amount && segment > 0.0f) {
                cut = parameter + ((accumulated - amount) / segment) * (previous_parameter - parameter);
                break;
            }
            if (parameter <= 1.0e-10f) {
                cut = 0.0f;
                for (uint32_t i = 0; i < n; ++i) points[i] = source[0];
                break;
            }
            previous_parameter = parameter;
            previous = current;
        }
    }
    if (cut > 0.0f) {
        const float step = cut / (float)(n - 1u);
        points[0] = source[0];
        for (uint32_t i = 1; i < n; ++i) points[i] = orc_curve_eval_v(source, n, step * (float)i);
    }
    return cut;
}

float orc_curve_eval(const float *points4, uint32_t n, float t, float *out3) {
    cv3 p[256];
    if (n > 256u) return NAN;
    for (uint32_t i = 0; i < n; ++i) p[i] = c3(points4[4 * i], points4[4 * i + 1], points4[4 * i + 2]);
    const cv3 r = orc_curve_eval_v(p, n, t);
    out3[0] = r.x;
    out3[1] = r.y;
    out3[2] = r.z;
    return 0.0f;
}

float orc_curve_spline_length(const float *points4, uint32_t n) {
    cv3 p[256];
    if (n > 256u) return NAN;
    for (uint32_t i = 0; i < n; ++i) p[i] = c3(points4[4 * i], points4[4 * i + 1], points4[4 * i + 2]);
    return orc_curve_spline_length_v(p, n);
}

/* The cut chain over `strand_count` packed strands of `cvs` points (float4
 * position + radius), compacting the kept strands, in order, into the
 * outputs; returns the kept count (or ~0u on a non-finite width or amount,
 * which the reference reports as an error).  multiplier may be NULL. */
uint32_t orc_classic_cut(const float *points, const uint32_t *counts, const nxg_root_sample *roots,
                         const nxg_vec2 *uvs, uint32_t strand_count, uint32_t cvs, float amount_scale,
                         float amount_offset, const float *multiplier, float radius_scale, float *out_points,
                         uint32_t *out_counts, nxg_root_sample *out_roots, nxg_vec2 *out_uvs, uint8_t *keep) {
    cv3 p[256], source[256];
    if (cvs < 2u || cvs > 256u) return ~0u;
    uint32_t write = 0u;
    for (uint32_t s = 0; s < strand_count; ++s) {
        const float *in = points + (uint64_t)s * cvs * 4u;
        for (uint32_t i = 0; i < cvs; ++i) p[i] = c3(in[4 * i], in[4 * i + 1], in[4 * i + 2]);
        const float c_length = orc_curve_spline_length_v(p, cvs);
        const float raw_width = radius_scale > 0.0f ? 2.0f * in[3] / radius_scale : 0.0f;
        if (!isfinite(raw_width)) return ~0u;
        uint8_t k = 1u;
        if (c_length < 1.0e-4f || raw_width < 1.0e-4f) {
            k = 0u;
        } else {
            float amount = amount_scale * c_length + amount_offset;
            if (multiplier) amount = amount * multiplier[s];
            if (!isfinite(amount)) return ~0u;
            const float cut = orc_cut_from_tip_v(p, cvs, max_std(amount, 0.0f), source);
            if (cut < 1.0e-4f) k = 0u;
        }
        if (keep) keep[s] = k;
        if (!k) continue;
        float *out = out_points + (uint64_t)write * cvs * 4u;
        for (uint32_t i = 0; i < cvs; ++i) {
            out[4 * i] = p[i].x;
            out[4 * i + 1] = p[i].y;
            out[4 * i + 2] = p[i].z;
            out[4 * i + 3] = in[4 * i + 3];
        }
        if (out_counts) out_counts[write] = counts ? counts[s] : cvs;
        if (out_roots && roots) out_roots[write] = roots[s];
        if (out_uvs && uvs) out_uvs[write] = uvs[s];
        ++write;
    }
    return write;
}

/* add_xgen_classic_renderer_endpoints (src/xgen_classic_runtime.cpp:1921-
 * 1956): each strand of cvs points becomes cvs + 2, a phantom point
 * 2 p0 - p1 before the root and 2 p[n-1] - p[n-2] after the tip, each with
 * its neighbour's radius.  out holds strand_count * (cvs + 2) float4s. */
void orc_classic_endpoints(const float *points, uint32_t strand_count, uint32_t cvs, float *out) {
    const uint32_t m = cvs + 2u;
    for (uint32_t s = 0; s < strand_count; ++s) {
        const float *in = points + (uint64_t)s * cvs * 4u;
        float *o = out + (uint64_t)s * m * 4u;
        memcpy(o + 4, in, sizeof(float) * 4u * cvs);
        for (int c = 0; c < 3; ++c) {
            o[c] = 2.0f * in[c] - in[4 + c];
            o[4u * (m - 1u) + c] = 2.0f * in[4u * (cvs - 1u) + c] - in[4u * (cvs - 2u) + c];
        }
        o[3] = in[3];
        o[4u * (m - 1u) + 3] = in[4u * (cvs - 1u) + 3];
    }
}
