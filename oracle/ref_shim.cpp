// ORACLE (test infrastructure only): C entry points over the reference's own
// CPU core, compiled from the reference sources where they lie (see Makefile
// target `ref`).  Used only in this container to pin the C restatement and to
// write golden fixtures; nothing here ships or runs on the GPU.
#include "nanoxgen/asset.h"
#include "nanoxgen/curve_cache.h"
#include "nanoxgen/curve_payload.h"
#include "nanoxgen/generate.h"
#include "nanoxgen/xgen_classic.h"
#include "nanoxgen/xgen_classic_runtime.h"
#include "nanoxgen/xgen_expression.h"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <span>
#include <sstream>
#include <string>
#include <vector>

#include "../include/nanoxgen_amd.h"

using namespace nanoxgen;

namespace {

Asset asset_from(const void *blob, uint64_t size) {
    AlignedByteVector bytes(size);
    std::memcpy(bytes.data(), blob, size);
    return Asset{std::move(bytes)};
}

int fail(char *err, uint64_t cap, const char *msg) {
    if (err && cap) std::snprintf(err, cap, "%s", msg);
    return 1;
}

GenerationParams params_from(const nxg_generation_params *p) {
    static_assert(sizeof(GenerationParams) == sizeof(nxg_generation_params));
    GenerationParams out{};
    std::memcpy(&out, p, sizeof(out));
    return out;
}

}  // namespace

extern "C" {

int ref_build_asset(const nxg_asset_build_input *in, void **blob, uint64_t *size, char *err,
                    uint64_t cap) {
    *blob = nullptr;
    *size = 0;
    try {
        AssetBuildInput input{};
        input.positions.assign(reinterpret_cast<const Vec3 *>(in->positions),
                               reinterpret_cast<const Vec3 *>(in->positions) + in->vertex_count);
        if (in->normals)
            input.normals.assign(reinterpret_cast<const Vec3 *>(in->normals),
                                 reinterpret_cast<const Vec3 *>(in->normals) + in->vertex_count);
        if (in->texcoords)
            input.texcoords.assign(reinterpret_cast<const Vec2 *>(in->texcoords),
                                   reinterpret_cast<const Vec2 *>(in->texcoords) + in->vertex_count);
        input.triangles.assign(reinterpret_cast<const UInt3 *>(in->triangles),
                               reinterpret_cast<const UInt3 *>(in->triangles) + in->triangle_count);
        uint64_t cv = 0;
        for (uint32_t g = 0; g < in->guide_count; ++g) {
            GuideInput guide{};
            const uint32_t n = in->guide_cv_counts[g];
            guide.cvs.assign(reinterpret_cast<const Vec3 *>(in->guide_cvs) + cv,
                             reinterpret_cast<const Vec3 *>(in->guide_cvs) + cv + n);
            cv += n;
            if (in->guide_root_normals) {
                guide.root_normal = {in->guide_root_normals[g].x, in->guide_root_normals[g].y,
                                     in->guide_root_normals[g].z};
            }
            if (in->guide_root_uvs) guide.root_uv = {in->guide_root_uvs[g].x, in->guide_root_uvs[g].y};
            if (in->guide_triangles) guide.triangle_index = in->guide_triangles[g];
            if (in->guide_barycentrics)
                guide.barycentric = {in->guide_barycentrics[g].x, in->guide_barycentrics[g].y};
            if (in->guide_support_radii) guide.support_radius = in->guide_support_radii[g];
            input.guides.push_back(std::move(guide));
        }
        const Asset asset = build_asset(input);
        void *out = nullptr;
        if (posix_memalign(&out, 64, asset.bytes().size()) != 0) return fail(err, cap, "oom");
        std::memcpy(out, asset.bytes().data(), asset.bytes().size());
        *blob = out;
        *size = asset.bytes().size();
        return 0;
    } catch (const std::exception &e) {
        return fail(err, cap, e.what());
    }
}

void ref_free(void *p) { std::free(p); }

int ref_validate_asset(const void *blob, uint64_t size, char *msg, uint64_t cap) {
    const std::string m = validate_asset({static_cast<const std::byte *>(blob), size});
    if (msg && cap) std::snprintf(msg, cap, "%s", m.c_str());
    return m.empty() ? 0 : 1;
}

// generate_packed_cpu / generate_packed_deformed_cpu / generate_packed_roots_cpu
int ref_generate_packed(const void *blob, uint64_t size, const nxg_generation_params *p,
                        const nxg_deformed_geometry *deformed, const nxg_root_sample *explicit_roots,
                        float radius_scale, uint32_t workers, uint32_t block,
                        nxg_packed_point *points, nxg_root_sample *roots, nxg_vec2 *uvs,
                        uint32_t *counts, char *err, uint64_t cap) {
    try {
        const Asset asset = asset_from(blob, size);
        const GenerationParams params = params_from(p);
        const CpuGenerationOptions options{workers, block};
        PackedGeneratedCurves curves;
        if (explicit_roots) {
            std::vector<RootSample> rs(p->strand_count);
            std::memcpy(rs.data(), explicit_roots, rs.size() * sizeof(RootSample));
            curves = generate_packed_roots_cpu(asset, params, rs, radius_scale, options);
        } else if (deformed) {
            const uint32_t v = asset.view().header().vertex_count;
            const uint32_t gcv = asset.view().header().guide_cv_count;
            DeformedGeometryView view{};
            if (deformed->positions)
                view.positions = {reinterpret_cast<const Vec3 *>(deformed->positions), v};
            if (deformed->normals) view.normals = {reinterpret_cast<const Vec3 *>(deformed->normals), v};
            if (deformed->guide_cvs)
                view.guide_cvs = {reinterpret_cast<const Vec3 *>(deformed->guide_cvs), gcv};
            curves = generate_packed_deformed_cpu(asset, params, view, radius_scale, options);
        } else {
            curves = generate_packed_cpu(asset, params, radius_scale, options);
        }
        std::memcpy(points, curves.points.data(), curves.points.size() * sizeof(PackedCurvePoint));
        if (roots) std::memcpy(roots, curves.roots.data(), curves.roots.size() * sizeof(RootSample));
        if (uvs) std::memcpy(uvs, curves.root_uvs.data(), curves.root_uvs.size() * sizeof(Vec2));
        if (counts) std::memcpy(counts, curves.point_counts.data(), curves.point_counts.size() * 4u);
        return 0;
    } catch (const std::exception &e) {
        return fail(err, cap, e.what());
    }
}

float ref_gradient_noise(const void *blob, float x, float y, float z) {
    return gradient_noise(DeviceAssetView{blob}, {x, y, z});
}

// make_guide_blend_state (generate.h:214-244) for one strand's sampled root.
void ref_blend(const void *blob, const nxg_generation_params *p, uint32_t strand, uint32_t *indices,
               float *weights, uint32_t *nearest) {
    const GenerationParams params = params_from(p);
    const DeviceAssetView view{blob};
    const RootSample root = sample_root(view, {}, params, strand);
    const GuideBlendState b = make_guide_blend_state(view, {}, params, root);
    for (int j = 0; j < 8; ++j) {
        indices[j] = b.indices[j];
        weights[j] = b.weights[j];
    }
    *nearest = b.nearest;
}

// make_strand_generation_state (generate.h:534-566)
void ref_strand_state(const void *blob, const nxg_generation_params *p, uint32_t strand, float *out8) {
    const GenerationParams params = params_from(p);
    const DeviceAssetView view{blob};
    const StrandGenerationState s = make_strand_generation_state(view, {}, params, strand);
    out8[0] = s.noise_domain.x;
    out8[1] = s.noise_domain.y;
    out8[2] = s.noise_domain.z;
    out8[3] = s.effective_noise_frequency;
    out8[4] = s.original_length;
    out8[5] = s.preserve_scale;
    out8[6] = s.surface_u.x;
    out8[7] = s.surface_u.y;
}

uint32_t ref_hash32(uint32_t x) { return hash32(x); }

// evaluate_xgen_scalar_expression_float (xgen_expression.cpp:984-1148) on a
// flattened program (reference op set only); inputs[i] is program input i.
// Returns 0 and *out, or 1 where the reference throws.
int ref_expr_eval(const nxg_expr_instruction *ins, uint32_t n_ins, const uint32_t *operands, uint32_t n_ops,
                  const nxg_expr_ramp_point *points, uint32_t n_points, const nxg_expr_ramp *ramps,
                  uint32_t n_ramps, const float *inputs, uint32_t n_inputs, uint32_t result, float u, float v,
                  float face_seed, float t, float *out) {
    try {
        XgenFloatExpressionProgram p{};
        for (uint32_t i = 0; i < n_ins; ++i) {
            p.instructions.push_back({static_cast<XgenScalarOp>(ins[i].op), ins[i].operand_offset,
                                      ins[i].operand_count, ins[i].auxiliary, ins[i].immediate});
        }
        p.operands.assign(operands, operands + n_ops);
        for (uint32_t i = 0; i < n_inputs; ++i) p.inputs.push_back("in" + std::to_string(i));
        for (uint32_t i = 0; i < n_points; ++i)
            p.ramp_points.push_back({points[i].position, points[i].value,
                                     static_cast<std::uint8_t>(points[i].interpolation)});
        for (uint32_t i = 0; i < n_ramps; ++i)
            p.ramps.push_back({ramps[i].point_offset, static_cast<std::uint16_t>(ramps[i].point_count)});
        p.result = result;
        XgenExpressionFloatContext ctx{};
        ctx.inputs = std::span<const float>(inputs, n_inputs);
        ctx.u = u;
        ctx.v = v;
        ctx.face_seed = face_seed;
        ctx.t = t;
        *out = evaluate_xgen_scalar_expression_float(p, ctx);
        return 0;
    } catch (const std::exception &) {
        return 1;
    }
}

// apply_xgen_classic_float_runtime_plan_cpu (xgen_classic_runtime.cpp:1580-1918)
// with one CutFX module whose amount is amount_scale * $cLength +
// amount_offset (an IR program over the bound input "cLength"), no length,
// width or taper expressions.  Strands are packed float4 points; roots,
// uvs and counts are carried.  Returns the kept strand count (outputs
// compacted in place), or ~0u where the reference throws; *seconds (when
// given) receives the wall time of the reference call alone.
uint32_t ref_classic_cut(float *points, uint32_t *counts, nxg_root_sample *roots, nxg_vec2 *uvs,
                         uint32_t strand_count, uint32_t cvs, float amount_scale, float amount_offset,
                         float radius_scale, double *seconds) {
    try {
        PackedGeneratedCurves curves{};
        curves.strand_count = strand_count;
        curves.cvs_per_strand = cvs;
        curves.point_counts.assign(counts, counts + strand_count);
        curves.points.resize(std::size_t{strand_count} * cvs);
        std::memcpy(curves.points.data(), points, curves.points.size() * sizeof(PackedCurvePoint));
        curves.roots.resize(strand_count);
        std::memcpy(curves.roots.data(), roots, std::size_t{strand_count} * sizeof(RootSample));
        if (uvs) {
            curves.root_uvs.resize(strand_count);
            std::memcpy(curves.root_uvs.data(), uvs, std::size_t{strand_count} * sizeof(Vec2));
        }
        XgenFloatExpressionProgram amount{};
        amount.inputs = {"cLength"};
        amount.instructions = {{XgenScalarOp::input, 0u, 0u, 0u, 0.0f},
                               {XgenScalarOp::constant, 0u, 0u, 0u, amount_scale},
                               {XgenScalarOp::multiply, 0u, 2u, 0u, 0.0f},
                               {XgenScalarOp::constant, 0u, 0u, 0u, amount_offset},
                               {XgenScalarOp::add, 2u, 2u, 0u, 0.0f}};
        amount.operands = {1u, 0u, 2u, 3u};
        amount.result = 4u;
        ClassicFloatRuntimePlan plan{};
        plan.description_name = "cut";
        plan.fx_cv_count = cvs;
        ClassicFloatCutModule cut{};
        cut.amount.object_type = "CutFXModule";
        cut.amount.attribute = "amount";
        cut.amount.program = amount;
        cut.rebuild_type = 1u;
        plan.cuts.push_back(cut);
        plan.effects.push_back({ClassicFloatEffectType::Cut, 0u});
        const auto start = std::chrono::steady_clock::now();
        apply_xgen_classic_float_runtime_plan_cpu(curves, plan, radius_scale);
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::memcpy(points, curves.points.data(), curves.points.size() * sizeof(PackedCurvePoint));
        std::memcpy(counts, curves.point_counts.data(), curves.point_counts.size() * sizeof(uint32_t));
        std::memcpy(roots, curves.roots.data(), curves.roots.size() * sizeof(RootSample));
        if (uvs) std::memcpy(uvs, curves.root_uvs.data(), curves.root_uvs.size() * sizeof(Vec2));
        return curves.strand_count;
    } catch (const std::exception &) {
        return ~0u;
    }
}

// add_xgen_classic_renderer_endpoints (xgen_classic_runtime.cpp:1921-1956)
// on strand_count packed strands of cvs points; out holds (cvs + 2) per
// strand.  Returns 0, or 1 where the reference throws.
uint32_t ref_classic_endpoints(const float *points, uint32_t strand_count, uint32_t cvs, float *out) {
    try {
        PackedGeneratedCurves curves{};
        curves.strand_count = strand_count;
        curves.cvs_per_strand = cvs;
        curves.point_counts.assign(strand_count, cvs);
        curves.points.resize(std::size_t{strand_count} * cvs);
        std::memcpy(curves.points.data(), points, curves.points.size() * sizeof(PackedCurvePoint));
        add_xgen_classic_renderer_endpoints(curves);
        std::memcpy(out, curves.points.data(), curves.points.size() * sizeof(PackedCurvePoint));
        return 0u;
    } catch (const std::exception &) {
        return 1u;
    }
}

// apply_xgen_classic_float_runtime_plan_cpu with one ClumpingFX module whose
// expressions are the constants k[0..6] (mask, clump, clumpScale, noise,
// noiseScale, noiseFrequency, noiseCorrelation), bound to world-space guides
// (ClassicClumpRuntimeData, xgen_classic_runtime.h:59-88; uvs, face ids and
// random prefixes zero; reference positions optional).  Strands must survive
// the runtime's length/width culling.  Returns the kept strand count, or ~0u
// where the reference throws.
uint32_t ref_clump_apply(float *points, uint32_t strand_count, uint32_t cvs, const uint32_t *strand_guides,
                         const float *axes3, const float *normals3, const float *tangents3,
                         const float *reference3, uint32_t guides, const float *k) {
    try {
        PackedGeneratedCurves curves{};
        curves.strand_count = strand_count;
        curves.cvs_per_strand = cvs;
        curves.point_counts.assign(strand_count, cvs);
        curves.points.resize(std::size_t{strand_count} * cvs);
        std::memcpy(curves.points.data(), points, curves.points.size() * sizeof(PackedCurvePoint));
        curves.roots.resize(strand_count);
        const auto constant = [](const char *attribute, float value) {
            ClassicFloatRuntimeExpression e{};
            e.object_type = "ClumpingFXModule";
            e.attribute = attribute;
            e.program.instructions = {{XgenScalarOp::constant, 0u, 0u, 0u, value}};
            e.program.result = 0u;
            return e;
        };
        ClassicFloatClumpModule clump{};
        clump.name = "clump";
        clump.mask = constant("mask", k[0]);
        clump.clump = constant("clump", k[1]);
        clump.clump_scale = constant("clumpScale", k[2]);
        clump.noise = constant("noise", k[3]);
        clump.noise_scale = constant("noiseScale", k[4]);
        clump.noise_frequency = constant("noiseFrequency", k[5]);
        clump.noise_correlation = constant("noiseCorrelation", k[6]);
        ClassicFloatRuntimePlan plan{};
        plan.description_name = "clump";
        plan.fx_cv_count = cvs;
        plan.clumps.push_back(clump);
        plan.effects.push_back({ClassicFloatEffectType::Clump, 0u});
        ClassicClumpRuntimeData data{};
        data.module_name = "clump";
        data.cvs_per_guide = cvs;
        const std::size_t axis_count = std::size_t{guides} * cvs;
        data.guide_axes.resize(axis_count);
        std::memcpy(data.guide_axes.data(), axes3, axis_count * sizeof(Vec3));
        data.guide_normals.resize(guides);
        std::memcpy(data.guide_normals.data(), normals3, guides * sizeof(Vec3));
        data.guide_tangents.resize(guides);
        std::memcpy(data.guide_tangents.data(), tangents3, guides * sizeof(Vec3));
        if (reference3) {
            data.guide_reference_positions.resize(guides);
            std::memcpy(data.guide_reference_positions.data(), reference3, guides * sizeof(Vec3));
        }
        data.guide_uvs.assign(guides, Vec2{});
        data.guide_face_ids.assign(guides, 0u);
        data.guide_random_prefixes.assign(guides, 0u);
        data.strand_guide_indices.assign(strand_guides, strand_guides + strand_count);
        apply_xgen_classic_float_runtime_plan_cpu(curves, plan, 1.0f, {}, {}, {}, {&data, 1u});
        std::memcpy(points, curves.points.data(), curves.points.size() * sizeof(PackedCurvePoint));
        return curves.strand_count;
    } catch (const std::exception &) {
        return ~0u;
    }
}

// build_curve_cache (curve_cache.cpp:104-164) from flat host arrays (any
// optional one NULL); the bytes go to *out (free with ref_free).  Returns
// 0, or 1 with the message in err.
int ref_build_curve_cache(const uint32_t *counts, uint32_t strands, const float *points, uint32_t point_count,
                          const float *texcoords, const float *patch_uvs, const float *face_uvs,
                          const uint32_t *face_ids, const float *motion_times, uint32_t samples,
                          const float *motion_points, void **out, uint64_t *size, char *err, uint64_t cap) {
    try {
        CurveCacheBuildInput in{};
        in.point_counts = {counts, strands};
        in.points = {reinterpret_cast<const PackedCurvePoint *>(points), point_count};
        if (texcoords) in.texcoords = {reinterpret_cast<const Vec2 *>(texcoords), point_count};
        if (patch_uvs) in.patch_uvs = {reinterpret_cast<const Vec2 *>(patch_uvs), strands};
        if (face_uvs) in.face_uvs = {reinterpret_cast<const Vec2 *>(face_uvs), strands};
        if (face_ids) in.face_ids = {face_ids, strands};
        if (motion_times) in.motion_times = {motion_times, samples};
        if (motion_points) in.motion_points = {reinterpret_cast<const Vec3 *>(motion_points),
                                               std::size_t{samples} * point_count};
        const CurveCache cache = build_curve_cache(in);
        *size = cache.bytes().size();
        *out = std::malloc(*size);
        std::memcpy(*out, cache.bytes().data(), *size);
        return 0;
    } catch (const std::exception &e) {
        std::snprintf(err, cap, "%s", e.what());
        return 1;
    }
}

// validate_curve_cache (curve_cache.cpp:166-239) into msg ("" when valid).
void ref_validate_curve_cache(const void *bytes, uint64_t size, char *msg, uint64_t cap) {
    const std::string e = validate_curve_cache({static_cast<const std::byte *>(bytes), size});
    std::snprintf(msg, cap, "%s", e.c_str());
}

// build_curve_batches (curve_payload.cpp:150-301).  Curve set k (0: base,
// 1..S: motion samples) has strand_counts[k] curves, counts[k] (NULL:
// fixed[k] CVs each), points[k]
// (npoints[k] Vec3), widths[k] (optional), constant_width[k]; the base may
// have root_uvs.  The batches are concatenated into out_points (packed),
// out_counts, out_uvs and out_motion (each batch's motion block in turn);
// batch_first[b] = first strand.  Returns the batch count, or ~0u with the
// message in err.
uint32_t ref_curve_payload(const uint32_t *strand_counts, uint32_t sets, const uint32_t *const *counts,
                           const uint32_t *fixed,
                           const float *const *points, const uint64_t *npoints, const float *const *widths,
                           const float *constant_width, const float *times, const float *root_uvs,
                           uint32_t max_strands, uint32_t resample, float radius_scale, const float *transform,
                           float *out_points, uint32_t *out_counts, float *out_uvs, float *out_motion,
                           uint32_t *batch_first, char *err, uint64_t cap) {
    try {
        const uint32_t strands = strand_counts[0];
        auto view = [&](uint32_t k) {
            CurveSetView v{};
            v.strand_count = strand_counts[k];
            v.fixed_cvs_per_strand = fixed[k];
            if (counts[k]) v.point_counts = {counts[k], strand_counts[k]};
            v.points = {reinterpret_cast<const Vec3 *>(points[k]), npoints[k]};
            if (widths[k]) v.widths = {widths[k], npoints[k]};
            v.constant_width = constant_width[k];
            return v;
        };
        CurveSetView base = view(0u);
        if (root_uvs) base.root_uvs = {reinterpret_cast<const Vec2 *>(root_uvs), strands};
        std::vector<MotionSampleView> motion;
        for (uint32_t k = 1; k < sets; ++k) motion.push_back({times[k - 1u], view(k)});
        CurvePayloadOptions options{};
        options.max_strands_per_batch = max_strands;
        options.resample_cv_count = resample;
        options.radius_scale = radius_scale;
        for (int i = 0; i < 12; ++i) options.transform.values[i] = transform[i];
        const std::vector<CurveBatch> batches = build_curve_batches(base, motion, {}, {}, options);
        std::size_t p = 0u, c = 0u, u = 0u, m = 0u;
        for (std::size_t b = 0; b < batches.size(); ++b) {
            const CurveBatch &batch = batches[b];
            batch_first[b] = batch.first_strand;
            std::memcpy(out_points + 4u * p, batch.points.data(), batch.points.size() * sizeof(PackedCurvePoint));
            p += batch.points.size();
            std::memcpy(out_counts + c, batch.point_counts.data(), batch.point_counts.size() * 4u);
            c += batch.point_counts.size();
            if (out_uvs && !batch.root_uvs.empty()) {
                std::memcpy(out_uvs + 2u * u, batch.root_uvs.data(), batch.root_uvs.size() * sizeof(Vec2));
                u += batch.root_uvs.size();
            }
            if (out_motion && !batch.motion_points.empty()) {
                std::memcpy(out_motion + 3u * m, batch.motion_points.data(), batch.motion_points.size() * sizeof(Vec3));
                m += batch.motion_points.size();
            }
        }
        return static_cast<uint32_t>(batches.size());
    } catch (const std::exception &e) {
        std::snprintf(err, cap, "%s", e.what());
        return ~0u;
    }
}

// compile_xgen_scalar_expression + make_xgen_float_expression_program
// (xgen_expression.cpp:771-818).  The float program goes to caller arrays of
// at least `cap` entries each (instructions as nxg_expr_instruction, ramp
// points as nxg_expr_ramp_point, ramps as two u32), input names NUL-separated
// into `names`; counts[0..7] = instructions, operands, ramp points, ramps,
// inputs, result, expression seed, rand calls.  Returns 0, or 1 with the
// message in err.
int ref_expr_compile(const char *source, uint64_t size, const char *name, const char *type, const uint64_t *limits,
                     nxg_expr_instruction *ins, uint32_t *operands, nxg_expr_ramp_point *points, uint32_t *ramps,
                     char *names, uint32_t cap, uint32_t *counts, char *err, uint64_t errcap) {
    try {
        XgenExpressionCompileOptions o{};
        o.expression_name = name;
        o.object_type = type;
        if (limits[0]) o.max_source_bytes = limits[0];
        if (limits[1]) o.max_tokens = limits[1];
        if (limits[2]) o.max_instructions = limits[2];
        if (limits[3]) o.max_locals = limits[3];
        if (limits[4]) o.max_ramp_points = limits[4];
        const XgenFloatExpressionProgram p =
            make_xgen_float_expression_program(compile_xgen_scalar_expression({source, size}, o));
        if (p.instructions.size() > cap || p.operands.size() > cap || p.ramp_points.size() > cap ||
            p.ramps.size() > cap)
            throw std::runtime_error("shim capacity");
        for (std::size_t i = 0; i < p.instructions.size(); ++i) {
            const XgenFloatScalarInstruction &x = p.instructions[i];
            ins[i] = {static_cast<uint8_t>(x.op), 0u, x.operand_count, x.operand_offset, x.auxiliary, x.immediate};
        }
        std::memcpy(operands, p.operands.data(), p.operands.size() * 4u);
        for (std::size_t i = 0; i < p.ramp_points.size(); ++i)
            points[i] = {p.ramp_points[i].position, p.ramp_points[i].value, p.ramp_points[i].interpolation};
        for (std::size_t i = 0; i < p.ramps.size(); ++i) {
            ramps[2 * i] = p.ramps[i].point_offset;
            ramps[2 * i + 1] = p.ramps[i].point_count;
        }
        char *n = names;
        for (const std::string &s : p.inputs) {
            std::memcpy(n, s.c_str(), s.size() + 1u);
            n += s.size() + 1u;
        }
        const uint32_t c[8] = {static_cast<uint32_t>(p.instructions.size()), static_cast<uint32_t>(p.operands.size()),
                               static_cast<uint32_t>(p.ramp_points.size()), static_cast<uint32_t>(p.ramps.size()),
                               static_cast<uint32_t>(p.inputs.size()), p.result, p.expression_seed,
                               p.random_call_count};
        std::memcpy(counts, c, sizeof(c));
        return 0;
    } catch (const std::exception &e) {
        std::snprintf(err, errcap, "%s", e.what());
        return 1;
    }
}

// The Classic float runtime as the reference runs a description: parse the
// .xgen text (parse_xgen_classic_collection), compile the first
// description's plan (compile_xgen_classic_float_runtime_plan, which binds
// map()/custom/$Prefg inputs to the input row) and apply it
// (apply_xgen_classic_float_runtime_plan_cpu) to packed curves with optional
// primitive ids, SeExpr prefixes and a row-major input table.  Returns the
// kept count (outputs compacted in place) or ~0u with the reference's
// message in err (a plan that needs a fallback reports its reasons).
// guide_count > 0 binds the description's ClumpingFX module `clump_name` to
// world-space guides (ClassicClumpRuntimeData: axes, normals, tangents,
// optional reference positions, uvs and SeExpr prefixes, a guide per strand).
uint32_t ref_classic_runtime(const char *xgen_text, float *points, uint32_t *counts, nxg_root_sample *roots,
                             nxg_vec2 *uvs, uint32_t strand_count, uint32_t cvs, const uint32_t *primitive_ids,
                             const uint32_t *prefixes, const float *inputs, uint64_t input_count,
                             float radius_scale, uint32_t clump_count, const char *const *clump_names,
                             const uint32_t *guide_counts, const float *guide_axes3, const float *guide_normals3,
                             const float *guide_tangents3, const float *guide_reference3, const float *guide_uvs2,
                             const uint32_t *guide_prefixes, const uint32_t *strand_guides,
                             const float *surface_tangents3, const float *noise_domain3,
                             const uint32_t *local_flags, const float *guide_local3, int root_relative, char *err,
                             uint64_t errcap, double *seconds) {
    try {
        std::istringstream text(xgen_text);
        const ClassicCollection collection = parse_xgen_classic_collection(text);
        if (collection.descriptions.empty()) throw std::runtime_error("no description");
        const ClassicFloatRuntimePlan plan =
            compile_xgen_classic_float_runtime_plan(collection.descriptions.front(), collection.palette_attributes);
        if (!plan.lowering_complete()) {
            std::string reasons = "fallback:";
            for (const std::string &r : plan.fallback_reasons) reasons += " " + r + ";";
            throw std::runtime_error(reasons);
        }
        PackedGeneratedCurves curves{};
        curves.strand_count = strand_count;
        curves.cvs_per_strand = cvs;
        curves.point_counts.assign(counts, counts + strand_count);
        curves.points.resize(std::size_t{strand_count} * cvs);
        std::memcpy(curves.points.data(), points, curves.points.size() * sizeof(PackedCurvePoint));
        curves.roots.resize(strand_count);
        std::memcpy(curves.roots.data(), roots, std::size_t{strand_count} * sizeof(RootSample));
        if (uvs) {
            curves.root_uvs.resize(strand_count);
            std::memcpy(curves.root_uvs.data(), uvs, std::size_t{strand_count} * sizeof(Vec2));
        }
        const std::span<const std::uint32_t> ids =
            primitive_ids ? std::span<const std::uint32_t>(primitive_ids, strand_count)
                          : std::span<const std::uint32_t>();
        const std::span<const std::uint32_t> pre =
            prefixes ? std::span<const std::uint32_t>(prefixes, strand_count) : std::span<const std::uint32_t>();
        const std::span<const float> rows =
            inputs ? std::span<const float>(inputs, input_count) : std::span<const float>();
        // one binding per module, the guide arrays of the modules back to back
        std::vector<ClassicClumpRuntimeData> clumps;
        std::size_t first = 0u;
        for (uint32_t m = 0; m < clump_count; ++m) {
            const uint32_t guide_count = guide_counts[m];
            ClassicClumpRuntimeData d{};
            d.module_name = clump_names[m];
            d.cvs_per_guide = cvs;
            const std::size_t axis_count = std::size_t{guide_count} * cvs;
            d.guide_axes.resize(axis_count);
            std::memcpy(d.guide_axes.data(), guide_axes3 + 3u * first * cvs, axis_count * sizeof(Vec3));
            d.guide_normals.resize(guide_count);
            std::memcpy(d.guide_normals.data(), guide_normals3 + 3u * first, guide_count * sizeof(Vec3));
            d.guide_tangents.resize(guide_count);
            std::memcpy(d.guide_tangents.data(), guide_tangents3 + 3u * first, guide_count * sizeof(Vec3));
            if (guide_reference3) {
                d.guide_reference_positions.resize(guide_count);
                std::memcpy(d.guide_reference_positions.data(), guide_reference3 + 3u * first,
                            guide_count * sizeof(Vec3));
            }
            d.guide_uvs.resize(guide_count);
            std::memcpy(d.guide_uvs.data(), guide_uvs2 + 2u * first, guide_count * sizeof(Vec2));
            d.guide_face_ids.assign(guide_count, 0u);
            d.guide_random_prefixes.assign(guide_prefixes + first, guide_prefixes + first + guide_count);
            d.strand_guide_indices.assign(strand_guides + std::size_t{m} * strand_count,
                                          strand_guides + std::size_t{m + 1u} * strand_count);
            if (local_flags && local_flags[m] && guide_local3) {
                d.guide_local_axes.resize(axis_count);
                std::memcpy(d.guide_local_axes.data(), guide_local3 + 3u * first * cvs, axis_count * sizeof(Vec3));
            }
            clumps.push_back(std::move(d));
            first += guide_count;
        }
        const auto start = std::chrono::steady_clock::now();
        std::vector<Vec3> tangents, domain;
        if (surface_tangents3) {
            tangents.resize(strand_count);
            std::memcpy(tangents.data(), surface_tangents3, std::size_t{strand_count} * sizeof(Vec3));
        }
        if (noise_domain3) {
            domain.resize(strand_count);
            std::memcpy(domain.data(), noise_domain3, std::size_t{strand_count} * sizeof(Vec3));
        }
        apply_xgen_classic_float_runtime_plan_cpu(curves, plan, radius_scale, tangents, pre, ids, clumps, rows,
                                                  domain, root_relative != 0);
        if (seconds) *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start).count();
        std::memcpy(points, curves.points.data(), curves.points.size() * sizeof(PackedCurvePoint));
        std::memcpy(counts, curves.point_counts.data(), curves.point_counts.size() * sizeof(uint32_t));
        std::memcpy(roots, curves.roots.data(), curves.roots.size() * sizeof(RootSample));
        if (uvs) std::memcpy(uvs, curves.root_uvs.data(), curves.root_uvs.size() * sizeof(Vec2));
        return curves.strand_count;
    } catch (const std::exception &e) {
        std::snprintf(err, errcap, "%s", e.what());
        return ~0u;
    }
}

double ref_seexpr_hash(const double *args, uint32_t n) { return xgen_seexpr_hash({args, n}); }
double ref_face_seed(uint32_t id, const char *patch, uint32_t face) { return xgen_face_seed(id, patch, face); }

uint32_t ref_runtime_hash_component(float x) { return xgen_runtime_hash_component(x); }
uint32_t ref_seexpr_hash_finish(uint32_t s) { return xgen_seexpr_hash_finish(s); }
float ref_random01(uint32_t seed, uint32_t strand, uint32_t dim) { return random01(seed, strand, dim); }

}  // extern "C"
