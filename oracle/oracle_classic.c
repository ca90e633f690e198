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
#include <stdlib.h>
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

/* Stable compaction of the strands whose flags[s] is 1 (kept, no error):
 * strand s's points were written at slot s of out_points, and each kept one
 * moves to the next free slot in strand order (src/xgen_classic_runtime.cpp:
 * 1893-1918); counts, roots and uvs follow it.  Returns the kept count. */
static uint32_t compact_slots(uint32_t strand_count, uint32_t cvs, const uint8_t *flags, const uint32_t *counts,
                              const nxg_root_sample *roots, const nxg_vec2 *uvs, float *out_points,
                              uint32_t *out_counts, nxg_root_sample *out_roots, nxg_vec2 *out_uvs) {
    uint32_t write = 0u;
    for (uint32_t s = 0; s < strand_count; ++s) {
        if (flags[s] != 1u) continue;
        if (write != s)
            memmove(out_points + (uint64_t)write * cvs * 4u, out_points + (uint64_t)s * cvs * 4u,
                    sizeof(float) * 4u * cvs);
        if (out_counts) out_counts[write] = counts ? counts[s] : cvs;
        if (out_roots && roots) out_roots[write] = roots[s];
        if (out_uvs && uvs) out_uvs[write] = uvs[s];
        ++write;
    }
    return write;
}

typedef struct {
    const float *points, *multiplier;
    uint32_t cvs;
    float amount_scale, amount_offset, radius_scale;
    float *out_points;
    uint8_t *flags; /* 0 culled, 1 kept, 2 non-finite width or amount */
} cut_ctx;

static void cut_one(void *c, uint32_t s) {
    const cut_ctx *x = (const cut_ctx *)c;
    cv3 p[256], source[256];
    const uint32_t cvs = x->cvs;
    const float *in = x->points + (uint64_t)s * cvs * 4u;
    for (uint32_t i = 0; i < cvs; ++i) p[i] = c3(in[4 * i], in[4 * i + 1], in[4 * i + 2]);
    const float c_length = orc_curve_spline_length_v(p, cvs);
    const float raw_width = x->radius_scale > 0.0f ? 2.0f * in[3] / x->radius_scale : 0.0f;
    x->flags[s] = 0u;
    if (!isfinite(raw_width)) {
        x->flags[s] = 2u;
        return;
    }
    if (c_length < 1.0e-4f || raw_width < 1.0e-4f) return;
    float amount = x->amount_scale * c_length + x->amount_offset;
    if (x->multiplier) amount = amount * x->multiplier[s];
    if (!isfinite(amount)) {
        x->flags[s] = 2u;
        return;
    }
    const float cut = orc_cut_from_tip_v(p, cvs, max_std(amount, 0.0f), source);
    if (cut < 1.0e-4f) return;
    float *out = x->out_points + (uint64_t)s * cvs * 4u;
    for (uint32_t i = 0; i < cvs; ++i) {
        out[4 * i] = p[i].x;
        out[4 * i + 1] = p[i].y;
        out[4 * i + 2] = p[i].z;
        out[4 * i + 3] = in[4 * i + 3];
    }
    x->flags[s] = 1u;
}

/* The cut chain over `strand_count` packed strands of `cvs` points (float4
 * position + radius), compacting the kept strands, in order, into the
 * outputs; returns the kept count (or ~0u on a non-finite width or amount,
 * which the reference reports as an error).  multiplier may be NULL.
 * out_points holds strand_count * cvs float4s (each strand is computed in
 * its own slot, in parallel over `workers` threads, 0 = all cores, before
 * the stable compaction); keep (strand_count flags) is required. */
uint32_t orc_classic_cut(const float *points, const uint32_t *counts, const nxg_root_sample *roots,
                         const nxg_vec2 *uvs, uint32_t strand_count, uint32_t cvs, float amount_scale,
                         float amount_offset, const float *multiplier, float radius_scale, float *out_points,
                         uint32_t *out_counts, nxg_root_sample *out_roots, nxg_vec2 *out_uvs, uint8_t *keep,
                         uint32_t workers) {
    if (cvs < 2u || cvs > 256u || !keep) return ~0u;
    cut_ctx c = {points, multiplier, cvs, amount_scale, amount_offset, radius_scale, out_points, keep};
    orc_run_tiles(strand_count, workers, 256u, cut_one, &c);
    for (uint32_t s = 0; s < strand_count; ++s)
        if (keep[s] == 2u) return ~0u;
    return compact_slots(strand_count, cvs, keep, counts, roots, uvs, out_points, out_counts, out_roots, out_uvs);
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

/* ---- Classic primitive expressions -------------------------------------
 * apply_xgen_classic_float_runtime_plan_cpu (src/xgen_classic_runtime.cpp:
 * 1608-1918) for a plan of SplinePrimitive expressions (length, width,
 * taper, taperStart, widthRamp) and an effect list of ClumpingFX, NoiseFX
 * and CutFX (rebuildType 1) modules.  Per strand, in the reference's order:
 *   fixed point counts                                          :1619-1622
 *   context: id, u, v, face seed, random prefix, input row      :1628-1647
 *   cLength; length expression, scale_curve about the root      :1648-1657
 *   raw width (expression, else 2 radius[0] / radius_scale)     :1658-1665
 *   cull below 1e-4 (the strand is skipped)                     :1666-1676
 *   the effects in list order                                   :1679-1826
 *     ClumpingFX (root-relative goals, local guide axes)        :1680-1793
 *     NoiseFX (apply_noise), cLength again                      :1794-1805
 *     CutFX: amount, cut_from_tip_and_rebuild, culled when the
 *            cut < 1e-4, cLength of the rebuilt curve           :1806-1821
 *   taper, taperStart; per-CV width profile -> radius           :1827-1858
 *   stable compaction                                           :1893-1918
 * Expression inputs bind as input_value does (:504-553): id, cLength,
 * cWidth, then the input row's columns. */

static int classic_eval(const nxg_expr_program *prog, orc_expr_ctx *ctx, float *vars, float c_length,
                        float c_width, float t, float *out) {
    float scratch[NXG_EXPR_MAX_VALUES];
    vars[ORC_CLASSIC_VAR_CLENGTH] = c_length;
    vars[ORC_CLASSIC_VAR_CWIDTH] = c_width;
    ctx->t = t;
    return orc_expr_eval_ctx(prog, ctx, scratch, out);
}

static uint32_t classic_strand(const float *in, uint32_t s, uint32_t cvs, const orc_classic_plan *plan,
                               const nxg_root_sample *roots, const uint32_t *primitive_ids,
                               const uint32_t *random_prefixes, const float *columns, uint32_t column_count,
                               float radius_scale, float *out4, uint8_t *keep) {
    cv3 p[256], source[256];
    float vars[3 + ORC_CLASSIC_MAX_COLUMNS];
    for (uint32_t i = 0; i < cvs; ++i) p[i] = c3(in[4 * i], in[4 * i + 1], in[4 * i + 2]);
    const nxg_root_sample *root = &roots[s];
    orc_expr_ctx ctx;
    memset(&ctx, 0, sizeof(ctx));
    vars[ORC_CLASSIC_VAR_ID] = (float)(primitive_ids ? primitive_ids[s] : s);
    vars[ORC_CLASSIC_VAR_CLENGTH] = 0.0f;
    vars[ORC_CLASSIC_VAR_CWIDTH] = 0.0f;
    for (uint32_t k = 0; k < column_count; ++k) vars[3u + k] = columns[(uint64_t)s * column_count + k];
    ctx.vars = vars;
    ctx.var_count = 3u + column_count;
    ctx.u = root->uv.x;
    ctx.v = root->uv.y;
    {
        const uint32_t face = root->surface_face_id == NXG_INVALID_INDEX ? root->triangle_index
                                                                          : root->surface_face_id;
        const float args[3] = {(float)plan->description_id, (float)plan->name_seed, (float)face};
        ctx.face_seed = orc_runtime_hash(args, 3u);
    }
    ctx.has_random_prefix = random_prefixes != NULL;
    ctx.random_prefix = random_prefixes ? random_prefixes[s] : 0u;

    *keep = 0u;
    float c_length = orc_curve_spline_length_v(p, cvs);
    if (plan->length) {
        float scale;
        if (classic_eval(plan->length, &ctx, vars, c_length, 0.0f, 0.0f, &scale)) return ORC_CLASSIC_EXPRESSION;
        if (!isfinite(scale) || scale < 0.0f) return ORC_CLASSIC_NEGATIVE_LENGTH;
        const cv3 r = p[0]; /* scale_curve, :576-585 */
        for (uint32_t i = 0; i < cvs; ++i) p[i] = add3(r, mul3(sub3(p[i], r), scale));
        c_length *= scale;
    }
    float raw_width;
    if (plan->width) {
        if (classic_eval(plan->width, &ctx, vars, c_length, 0.0f, 0.0f, &raw_width)) return ORC_CLASSIC_EXPRESSION;
    } else {
        raw_width = radius_scale > 0.0f ? 2.0f * in[3] / radius_scale : 0.0f;
    }
    if (!isfinite(raw_width)) return ORC_CLASSIC_NONFINITE_WIDTH;
    if (c_length < 1.0e-4f || raw_width < 1.0e-4f) return 0u;
    const float c_width = raw_width;
    /* the context's t after the effects: a module with a guide sets it to 0
     * for its mask, and one that clumps (or a NoiseFX that runs) leaves it
     * at its last CV's (1); the cut amounts, taper and taperStart read it */
    float t_after = 0.0f;
    uint8_t k = 1u;
    for (uint32_t e = 0; e < plan->effect_count; ++e) { /* plan.effects in order, :1679-1826 */
        const uint32_t type = plan->effect_type[e], module = plan->effect_module[e];
        if (type == ORC_EFFECT_NOISE) { /* apply_noise, :897-1039 */
            const nxg_expr_program *const *ex = plan->noise[module];
            float raw_mask, f, c, pl, magnitudes[256];
            if (classic_eval(ex[0], &ctx, vars, c_length, c_width, t_after, &raw_mask)) return ORC_CLASSIC_EXPRESSION;
            const float mask = raw_mask < 0.0f ? 0.0f : (1.0f < raw_mask ? 1.0f : raw_mask);
            if (mask > 1.0e-6f) {
                if (classic_eval(ex[3], &ctx, vars, c_length, c_width, t_after, &f) ||
                    classic_eval(ex[4], &ctx, vars, c_length, c_width, t_after, &c) ||
                    classic_eval(ex[5], &ctx, vars, c_length, c_width, t_after, &pl))
                    return ORC_CLASSIC_EXPRESSION;
                const float frequency = max_std(f, 0.0f);
                const float cc = c * 0.01f, pp = pl * 0.01f;
                const float correlation = cc < 0.0f ? 0.0f : (1.0f < cc ? 1.0f : cc);
                const float preserve = pp < 0.0f ? 0.0f : (1.0f < pp ? 1.0f : pp);
                if (!isfinite(frequency) || !isfinite(correlation) || !isfinite(preserve))
                    return ORC_CLASSIC_NONFINITE_NOISE_PARAMETER;
                /* the magnitudes depend on t alone along the strand; the
                 * reference evaluates them in CV order inside its loop */
                for (uint32_t cv = 0u; cv < cvs; ++cv) {
                    float mg, ms;
                    t_after = (float)cv / (float)(cvs - 1u);
                    if (classic_eval(ex[1], &ctx, vars, c_length, c_width, t_after, &mg) ||
                        classic_eval(ex[2], &ctx, vars, c_length, c_width, t_after, &ms))
                        return ORC_CLASSIC_EXPRESSION;
                    magnitudes[cv] = max_std(mg, 0.0f) * max_std(ms, 0.0f) * mask;
                    if (!isfinite(magnitudes[cv])) return ORC_CLASSIC_NONFINITE_NOISE;
                }
                const float normal3[3] = {root->normal.x, root->normal.y, root->normal.z};
                orc_noise_fx(&p[0].x, cvs, frequency, correlation, preserve,
                             plan->noise_domain3 ? plan->noise_domain3 + 3u * s : NULL, normal3,
                             plan->surface_tangents3 + 3u * s, magnitudes, plan->gradients);
            }
            c_length = orc_curve_spline_length_v(p, cvs); /* whatever the mask (:1805) */
            continue;
        }
        if (type == ORC_EFFECT_CUT) { /* :1806-1821 */
            float amount;
            if (classic_eval(plan->cut_amount[module], &ctx, vars, c_length, c_width, t_after, &amount))
                return ORC_CLASSIC_EXPRESSION;
            if (!isfinite(amount)) return ORC_CLASSIC_NONFINITE_AMOUNT;
            const float cut = orc_cut_from_tip_v(p, cvs, max_std(amount, 0.0f), source);
            if (cut < 1.0e-4f) k = 0u; /* culled; the later effects still run */
            c_length = orc_curve_spline_length_v(p, cvs);
            continue;
        }
        const orc_clump_binding *b = plan->binding[module]; /* ClumpingFX, :1680-1793 */
        const nxg_expr_program *const *ex = plan->clump[module];
        const uint32_t g = b->strand_guides[s];
        if (g == NXG_INVALID_INDEX) continue;
        if (g >= b->guide_count) return ORC_CLASSIC_INVALID_GUIDE;
        t_after = 0.0f;
        float raw_mask;
        if (classic_eval(ex[0], &ctx, vars, c_length, c_width, 0.0f, &raw_mask)) return ORC_CLASSIC_EXPRESSION;
        const float mask = raw_mask < 0.0f ? 0.0f : (1.0f < raw_mask ? 1.0f : raw_mask);
        if (!(mask > 0.0f)) continue;
        /* build_clump_noise_axis in the guide's context (:1041-1186) */
        orc_expr_ctx gctx = ctx;
        gctx.u = b->guide_uvs[g].x;
        gctx.v = b->guide_uvs[g].y;
        gctx.face_seed = 0.0f; /* read only by rand() without a prefix */
        gctx.random_prefix = b->guide_prefixes[g];
        gctx.has_random_prefix = 1;
        float raw_noise, magnitudes[256], noise3[3 * 256];
        if (classic_eval(ex[3], &gctx, vars, c_length, c_width, 0.0f, &raw_noise)) return ORC_CLASSIC_EXPRESSION;
        const float noise = max_std(raw_noise, 0.0f);
        const int noisy = noise > 1.0e-5f && mask > 1.0e-4f;
        float frequency = 0.0f, correlation = 0.0f;
        if (noisy) {
            if (classic_eval(ex[5], &gctx, vars, c_length, c_width, 0.0f, &frequency) ||
                classic_eval(ex[6], &gctx, vars, c_length, c_width, 0.0f, &correlation))
                return ORC_CLASSIC_EXPRESSION;
            for (uint32_t cv = 1u; cv < cvs; ++cv) {
                float raw_scale;
                const float t = (float)cv / (float)(cvs - 1u);
                if (classic_eval(ex[4], &gctx, vars, c_length, c_width, t, &raw_scale)) return ORC_CLASSIC_EXPRESSION;
                magnitudes[cv] = mask * noise * max_std(raw_scale, 0.0f);
                if (!isfinite(magnitudes[cv])) return ORC_CLASSIC_NONFINITE_CLUMP;
            }
        }
        /* root-relative evaluation with local guide axes (:1697-1721): the
         * source and render axes are the local copies, the goal their render
         * axis moved to the world guide root, about the strand's root */
        const int local = plan->root_relative && b->local3 != NULL;
        const float *world = b->axes3 + (uint64_t)g * cvs * 3u;
        const float *axes = local ? b->local3 + (uint64_t)g * cvs * 3u : world;
        const float *render = (local ? b->local_render3 : b->render3) + (uint64_t)g * cvs * 3u;
        orc_clump_noise_axis(render, axes, cvs, frequency, correlation,
                             b->reference3 ? b->reference3 + 3u * g : world, b->normals3 + 3u * g,
                             b->tangents3 + 3u * g, noisy ? magnitudes : NULL, b->gradients, noise3);
        orc_cut_to_length(&p[0].x, cvs, b->lengths[g]);
        c_length = orc_curve_spline_length_v(p, cvs);
        float raw_amount;
        if (classic_eval(ex[1], &ctx, vars, c_length, c_width, 0.0f, &raw_amount)) return ORC_CLASSIC_EXPRESSION;
        const float amount = raw_amount < 0.0f ? 0.0f : (1.0f < raw_amount ? 1.0f : raw_amount);
        const cv3 origin = plan->root_relative ? c3(root->position.x, root->position.y, root->position.z)
                                               : c3(0.0f, 0.0f, 0.0f);
        const cv3 guide_root = c3(world[0], world[1], world[2]);
        for (uint32_t cv = 1u; cv < cvs; ++cv) {
            float scale;
            t_after = (float)cv / (float)(cvs - 1u);
            if (classic_eval(ex[2], &ctx, vars, c_length, c_width, t_after, &scale)) return ORC_CLASSIC_EXPRESSION;
            const cv3 r = c3(render[3 * cv], render[3 * cv + 1], render[3 * cv + 2]);
            const cv3 goal = local ? add3(r, sub3(guide_root, origin)) : sub3(r, origin);
            p[cv] = add3(p[cv], mul3(sub3(goal, p[cv]), mask * amount * (1.0f - 2.0f * scale)));
        }
        for (uint32_t cv = 1u; cv < cvs; ++cv)
            p[cv] = add3(p[cv], c3(noise3[3 * cv], noise3[3 * cv + 1], noise3[3 * cv + 2]));
        c_length = orc_curve_spline_length_v(p, cvs);
    }
    float taper = 0.0f, taper_start = 0.0f;
    if (plan->taper && classic_eval(plan->taper, &ctx, vars, c_length, c_width, t_after, &taper))
        return ORC_CLASSIC_EXPRESSION;
    if (plan->taper_start && classic_eval(plan->taper_start, &ctx, vars, c_length, c_width, t_after, &taper_start))
        return ORC_CLASSIC_EXPRESSION;
    if (!isfinite(taper) || !isfinite(taper_start)) return ORC_CLASSIC_NONFINITE_TAPER;
    const int profile = plan->width || plan->taper || plan->taper_start || plan->width_ramp;
    for (uint32_t i = 0; i < cvs; ++i) {
        out4[4 * i] = p[i].x;
        out4[4 * i + 1] = p[i].y;
        out4[4 * i + 2] = p[i].z;
        out4[4 * i + 3] = in[4 * i + 3];
        if (!profile) continue;
        const float t = (float)i / (float)(cvs - 1u);
        float scale = 1.0f;
        if (t > taper_start && taper_start < 1.0f) scale *= 1.0f - taper * ((t - taper_start) / (1.0f - taper_start));
        if (plan->width_ramp) {
            float ramp;
            if (classic_eval(plan->width_ramp, &ctx, vars, c_length, c_width, t, &ramp)) return ORC_CLASSIC_EXPRESSION;
            scale *= ramp;
        }
        const float raw_diameter = c_width * scale;
        if (!isfinite(raw_diameter)) return ORC_CLASSIC_NONFINITE_PROFILE;
        out4[4 * i + 3] = 0.5f * max_std(raw_diameter, 0.0f) * radius_scale;
    }
    *keep = k;
    return 0u;
}

typedef struct {
    const float *points;
    const uint32_t *counts;
    const nxg_root_sample *roots;
    uint32_t cvs;
    const orc_classic_plan *plan;
    const uint32_t *primitive_ids, *random_prefixes;
    const float *columns;
    uint32_t column_count;
    float radius_scale;
    float *out_points;
    uint8_t *flags; /* 0 culled, 1 kept; keep[] values */
    uint32_t *errors; /* per strand ORC_CLASSIC_* bits */
} runtime_ctx;

static void runtime_one(void *c, uint32_t s) {
    const runtime_ctx *x = (const runtime_ctx *)c;
    const uint32_t cvs = x->cvs;
    x->flags[s] = 0u;
    if (x->counts && x->counts[s] != cvs) {
        x->errors[s] = ORC_CLASSIC_VARIABLE_COUNTS;
        return;
    }
    uint8_t k = 0u;
    x->errors[s] = classic_strand(x->points + (uint64_t)s * cvs * 4u, s, cvs, x->plan, x->roots, x->primitive_ids,
                                  x->random_prefixes, x->columns, x->column_count, x->radius_scale,
                                  x->out_points + (uint64_t)s * cvs * 4u, &k);
    x->flags[s] = k;
}

/* out_points holds strand_count * cvs float4s: each strand is computed in
 * its own slot, in parallel over `workers` threads (0 = all cores), then
 * the kept ones are compacted in strand order. */
uint32_t orc_classic_runtime(const float *points, const uint32_t *counts, const nxg_root_sample *roots,
                             const nxg_vec2 *uvs, uint32_t strand_count, uint32_t cvs,
                             const orc_classic_plan *plan, const uint32_t *primitive_ids,
                             const uint32_t *random_prefixes, const float *columns, uint32_t column_count,
                             float radius_scale, float *out_points, uint32_t *out_counts,
                             nxg_root_sample *out_roots, nxg_vec2 *out_uvs, uint8_t *keep, uint32_t *error_bits,
                             uint32_t workers) {
    if (cvs < 2u || cvs > 256u || column_count > ORC_CLASSIC_MAX_COLUMNS || !roots || !keep) {
        *error_bits = ~0u;
        return ~0u;
    }
    uint32_t *errors = (uint32_t *)calloc(strand_count ? strand_count : 1u, sizeof(uint32_t));
    if (!errors) {
        *error_bits = ~0u;
        return ~0u;
    }
    runtime_ctx c = {points, counts, roots, cvs, plan, primitive_ids, random_prefixes, columns, column_count,
                     radius_scale, out_points, keep, errors};
    orc_run_tiles(strand_count, workers, 64u, runtime_one, &c);
    uint32_t bits = 0u;
    for (uint32_t s = 0; s < strand_count; ++s) bits |= errors[s];
    free(errors);
    *error_bits = bits;
    if (bits) return ~0u;
    return compact_slots(strand_count, cvs, keep, NULL, roots, uvs, out_points, out_counts, out_roots, out_uvs);
}
