// Checked device contract of the C ABI: status messages and request
// validation, restating include/nanoxgen/device_contract.h:58-254 of the
// reference in C-ABI terms (status codes 0..20 are its DeviceGenerationError
// values; 21.. are extensions).
#include <cmath>
#include <cstdint>
#include <limits>

#include "internal.h"

namespace nxg {

namespace {
constexpr uint64_t kHeaderBytes = sizeof(nxg_asset_header);

uint64_t align256(uint64_t v) { return (v + 255u) & ~uint64_t{255u}; }
}  // namespace

bool valid_generation_parameters(const nxg_generation_params &p) {
    using std::isfinite;
    return p.strand_count != 0u && p.cvs_per_strand >= 2u && isfinite(p.guide_support_scale) &&
           p.guide_support_scale >= 0.0f && isfinite(p.guide_weight_power) &&
           p.guide_weight_power >= 0.0f && isfinite(p.normal_rejection_cos) &&
           p.normal_rejection_cos >= -1.0f && p.normal_rejection_cos <= 1.0f &&
           isfinite(p.length_scale) && p.length_scale >= 0.0f && isfinite(p.root_width) &&
           p.root_width >= 0.0f && isfinite(p.tip_width) && p.tip_width >= 0.0f &&
           isfinite(p.noise_amplitude) && p.noise_amplitude >= 0.0f &&
           isfinite(p.noise_frequency) && p.noise_frequency >= 0.0f && isfinite(p.noise_mask) &&
           p.noise_mask >= 0.0f && p.noise_mask <= 1.0f && isfinite(p.noise_correlation) &&
           p.noise_correlation >= 0.0f && p.noise_correlation <= 1.0f &&
           isfinite(p.noise_preserve_length) && p.noise_preserve_length >= 0.0f &&
           p.noise_preserve_length <= 1.0f;
}

bool valid_octaves(const nxg_noise_octaves *o) {
    if (!o) return true;
    if (o->count > NXG_MAX_NOISE_OCTAVES) return false;
    for (uint32_t k = 0; k < o->count; ++k) {
        if (!std::isfinite(o->amplitude[k]) || o->amplitude[k] < 0.0f ||
            !std::isfinite(o->frequency[k]) || o->frequency[k] < 0.0f)
            return false;
    }
    return true;
}

uint32_t validate_asset_view(const nxg_device_asset *asset) {
    if (!asset) return NXG_ERROR_NULL_ARGUMENT;
    if (!asset->data) return NXG_ERROR_NULL_ASSET;
    if (reinterpret_cast<uintptr_t>(asset->data) % NXG_BLOB_ALIGNMENT != 0u)
        return NXG_ERROR_MISALIGNED_ASSET;
    const nxg_asset_header &h = asset->header;
    if (h.magic != NXG_MAGIC || h.version_major != NXG_VERSION_MAJOR ||
        h.version_minor != NXG_VERSION_MINOR || h.byte_size < kHeaderBytes ||
        h.vertex_count == 0u || h.triangle_count == 0u || h.guide_count == 0u ||
        h.guide_cv_count == 0u || h.guide_stencil_size != NXG_GUIDE_STENCIL_SIZE ||
        h.noise_gradient_count != NXG_NOISE_GRADIENT_COUNT ||
        h.noise_gradients_offset < kHeaderBytes || h.noise_gradients_offset > h.byte_size ||
        uint64_t{h.noise_gradient_count} * sizeof(nxg_vec3) > h.byte_size - h.noise_gradients_offset)
        return NXG_ERROR_INVALID_ASSET_METADATA;
    if (asset->byte_capacity < h.byte_size) return NXG_ERROR_ASSET_CAPACITY_TOO_SMALL;
    return NXG_OK;
}

uint32_t validate_common(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                         const nxg_generation_params *params, const nxg_launch_config *config) {
    if (!asset || !params) return NXG_ERROR_NULL_ARGUMENT;
    if (uint32_t s = validate_asset_view(asset); s != NXG_OK) return s;
    const nxg_asset_header &h = asset->header;
    if (!valid_generation_parameters(*params)) return NXG_ERROR_INVALID_GENERATION_PARAMETERS;
    const uint32_t block = config ? config->block_size : 128u;
    if (block == 0u || block > 1024u) return NXG_ERROR_INVALID_BLOCK_SIZE;
    if (deformed) {
        const nxg_deformed_geometry &d = *deformed;
        if ((d.positions == nullptr) != (d.position_count == 0u) ||
            (d.positions && d.position_count != h.vertex_count))
            return NXG_ERROR_DEFORMED_POSITION_COUNT_MISMATCH;
        if ((d.normals == nullptr) != (d.normal_count == 0u) ||
            (d.normals && d.normal_count != h.vertex_count))
            return NXG_ERROR_DEFORMED_NORMAL_COUNT_MISMATCH;
        if ((d.guide_cvs == nullptr) != (d.guide_cv_count == 0u) ||
            (d.guide_cvs && d.guide_cv_count != h.guide_cv_count))
            return NXG_ERROR_DEFORMED_GUIDE_CV_COUNT_MISMATCH;
    }
    return NXG_OK;
}

uint32_t validate_packed_output(const nxg_generation_params &p, const nxg_packed_output *out) {
    if (!out || !out->points) return NXG_ERROR_MISSING_OUTPUT;
    if (!std::isfinite(out->radius_scale) || out->radius_scale < 0.0f)
        return NXG_ERROR_INVALID_RADIUS_SCALE;
    const uint64_t points = uint64_t{p.strand_count} * p.cvs_per_strand;
    if (out->point_capacity < points) return NXG_ERROR_POINT_CAPACITY_TOO_SMALL;
    auto optional_ok = [](const void *ptr, uint64_t cap, uint64_t need) {
        return ptr ? cap >= need : cap == 0u;
    };
    if (!optional_ok(out->roots, out->root_capacity, p.strand_count))
        return NXG_ERROR_ROOT_CAPACITY_TOO_SMALL;
    if (!optional_ok(out->root_uvs, out->root_uv_capacity, p.strand_count))
        return NXG_ERROR_ROOT_UV_CAPACITY_TOO_SMALL;
    if (!optional_ok(out->point_counts, out->point_count_capacity, p.strand_count))
        return NXG_ERROR_POINT_COUNT_CAPACITY_TOO_SMALL;
    return NXG_OK;
}

WorkspaceLayout workspace_layout(uint32_t guide_count, uint32_t cvs_per_strand, uint32_t strand_count,
                                 uint32_t triangle_count) {
    (void)triangle_count;
    WorkspaceLayout l{};
    l.guide_state_offset = 0;
    l.guide_offsets_offset = align256(uint64_t{guide_count} * 32u);
    l.order_offset = align256(l.guide_offsets_offset + uint64_t{guide_count} * cvs_per_strand * 16u);
    l.counts_offset = align256(l.order_offset + uint64_t{strand_count} * 4u);
    const uint64_t counts = uint64_t{kOrderBins} * order_tiles(strand_count);
    l.partials_offset = align256(l.counts_offset + counts * 4u);
    l.keys_offset = align256(l.partials_offset + (counts / kScanBlock + 1u) * 4u);
    l.total = align256(l.keys_offset + uint64_t{strand_count} * 4u);
    return l;
}

}  // namespace nxg

extern "C" {

const char *nxg_status_message(uint32_t status) {
    switch (status) {
        case NXG_OK: return "";
        case NXG_ERROR_NULL_ASSET: return "device asset pointer is null";
        case NXG_ERROR_MISALIGNED_ASSET: return "device asset pointer is not 64-byte aligned";
        case NXG_ERROR_INVALID_ASSET_METADATA: return "device asset metadata is invalid";
        case NXG_ERROR_ASSET_CAPACITY_TOO_SMALL: return "device asset allocation is too small";
        case NXG_ERROR_INVALID_GENERATION_PARAMETERS: return "generation parameters are invalid";
        case NXG_ERROR_INVALID_BLOCK_SIZE: return "GPU block size must be in [1, 1024]";
        case NXG_ERROR_DEFORMED_POSITION_COUNT_MISMATCH:
            return "deformed position count does not match the asset";
        case NXG_ERROR_DEFORMED_NORMAL_COUNT_MISMATCH:
            return "deformed normal count does not match the asset";
        case NXG_ERROR_DEFORMED_GUIDE_CV_COUNT_MISMATCH:
            return "deformed guide-CV count does not match the asset";
        case NXG_ERROR_MISSING_OUTPUT: return "no device output was supplied";
        case NXG_ERROR_POINT_CAPACITY_TOO_SMALL: return "device point output is too small";
        case NXG_ERROR_WIDTH_CAPACITY_TOO_SMALL: return "device width output is too small";
        case NXG_ERROR_ROOT_CAPACITY_TOO_SMALL: return "device root output is too small";
        case NXG_ERROR_ROOT_UV_CAPACITY_TOO_SMALL: return "device root-UV output is too small";
        case NXG_ERROR_POINT_COUNT_CAPACITY_TOO_SMALL:
            return "device point-count output is too small";
        case NXG_ERROR_INVALID_RADIUS_SCALE:
            return "device curve radius scale must be finite and non-negative";
        case NXG_ERROR_MISSING_MOTION_SAMPLES: return "no motion samples were supplied";
        case NXG_ERROR_INVALID_MOTION_TIMES:
            return "motion sample times must be finite and strictly increasing";
        case NXG_ERROR_MOTION_POINT_CAPACITY_TOO_SMALL: return "device motion-point output is too small";
        case NXG_ERROR_OUTPUT_SIZE_OVERFLOW: return "device output size overflows uint64";
        case NXG_ERROR_HIP: return "a HIP runtime call failed";
        case NXG_ERROR_WORKSPACE_TOO_SMALL: return "device workspace is too small";
        case NXG_ERROR_INVALID_OCTAVES:
            return "noise octaves must number at most 4 with finite non-negative amplitude and frequency";
        case NXG_ERROR_NULL_ARGUMENT: return "a required argument pointer is null";
        case NXG_ERROR_STENCIL_CAPACITY_TOO_SMALL: return "device stencil output is too small";
        case NXG_ERROR_INVALID_ROOTS: return "explicit roots are missing";
        case NXG_ERROR_INVALID_BUILD_INPUT: return "asset build input is invalid";
        case NXG_ERROR_BUFFER_TOO_SMALL: return "output buffer is too small";
        case NXG_ERROR_INVALID_ASSET: return "asset is invalid";
        case NXG_ERROR_INVALID_PROGRAM: return "expression program or texel map is invalid";
        case NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL: return "device attribute buffer is too small";
        case NXG_ERROR_INVALID_MLP_REQUEST: return "surrogate request is invalid";
        case NXG_ERROR_INVALID_BATCH: return "batch item count or tuning is invalid";
        case NXG_ERROR_INVALID_CLASSIC_REQUEST:
            return "Classic cut curves, parameters or outputs are invalid (CV count, alignment, overlap)";
        case NXG_ERROR_SCRATCH_LIMIT:
            return "request needs more workspace than the library's static scratch (NXG_STATIC_SCRATCH_BYTES)";
        case NXG_ERROR_INVALID_SAMPLER:
            return "inverse-CDF sampler buffer is missing, misaligned or smaller than nxg_cdf_sampler_size";
    }
    return "unknown device generation error";
}

uint32_t nxg_validate_packed_generation(const nxg_device_asset *asset,
                                        const nxg_deformed_geometry *deformed,
                                        const nxg_generation_params *params,
                                        const nxg_packed_output *output,
                                        const nxg_launch_config *config) {
    const uint32_t common = nxg::validate_common(asset, deformed, params, config);
    if (common != NXG_OK) return common;
    return nxg::validate_packed_output(*params, output);
}

uint64_t nxg_packed_generation_workspace_size(const nxg_asset_header *header,
                                              const nxg_generation_params *params) {
    if (!header || !params) return 0;
    return nxg::workspace_layout(header->guide_count, params->cvs_per_strand, params->strand_count,
                                 header->triangle_count)
        .total;
}

uint32_t nxg_validate_motion_generation(const nxg_device_asset *asset, const nxg_motion_sample *samples,
                                        uint32_t sample_count, const nxg_generation_params *params,
                                        const nxg_motion_output *output, const nxg_launch_config *config) {
    if (!samples || sample_count == 0u) return NXG_ERROR_MISSING_MOTION_SAMPLES;
    float previous = -std::numeric_limits<float>::infinity();
    for (uint32_t s = 0; s < sample_count; ++s) {
        if (!std::isfinite(samples[s].time) || samples[s].time <= previous) return NXG_ERROR_INVALID_MOTION_TIMES;
        previous = samples[s].time;
        if (uint32_t e = nxg::validate_common(asset, &samples[s].deformation, params, config); e != NXG_OK) return e;
    }
    const uint64_t per_sample = uint64_t{params->strand_count} * params->cvs_per_strand;
    if (per_sample > std::numeric_limits<uint64_t>::max() / sample_count) return NXG_ERROR_OUTPUT_SIZE_OVERFLOW;
    if (!output || !output->points || output->point_capacity < per_sample * sample_count)
        return NXG_ERROR_MOTION_POINT_CAPACITY_TOO_SMALL;
    return NXG_OK;
}

uint64_t nxg_motion_generation_workspace_size(const nxg_asset_header *header, const nxg_generation_params *params,
                                              uint32_t sample_count) {
    return nxg_packed_generation_workspace_size(header, params) * sample_count;
}

}  // extern "C"

namespace {
bool overlaps(const void *a, uint64_t a_bytes, const void *b, uint64_t b_bytes) {
    if (!a || !b || a_bytes == 0u || b_bytes == 0u) return false;
    const uintptr_t x = reinterpret_cast<uintptr_t>(a), y = reinterpret_cast<uintptr_t>(b);
    return x < y + b_bytes && y < x + a_bytes;
}
}  // namespace

uint32_t nxg_validate_classic_cut(const nxg_classic_curves *curves, const nxg_classic_cut_params *cut,
                                  const nxg_classic_output *output) {
    if (!curves || !cut || !output || !output->kept_count) return NXG_ERROR_NULL_ARGUMENT;
    const uint64_t n = curves->strand_count, cvs = curves->cvs_per_strand;
    if (cvs < 2u || cvs > NXG_CLASSIC_MAX_CVS || n > NXG_CLASSIC_MAX_STRANDS)
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (cut->multiplier && cut->multiplier_stride == 0u) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (n == 0u) return NXG_OK;
    if (!curves->points || !output->points) return NXG_ERROR_NULL_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(curves->points) % 16u != 0u ||
        reinterpret_cast<uintptr_t>(output->points) % 16u != 0u)
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (output->point_capacity < n * cvs) return NXG_ERROR_POINT_CAPACITY_TOO_SMALL;
    if (output->point_counts && output->point_count_capacity < n) return NXG_ERROR_POINT_COUNT_CAPACITY_TOO_SMALL;
    if (output->roots && !curves->roots) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (output->roots && output->root_capacity < n) return NXG_ERROR_ROOT_CAPACITY_TOO_SMALL;
    if (output->root_uvs && !curves->root_uvs) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (output->root_uvs && output->root_uv_capacity < n) return NXG_ERROR_ROOT_UV_CAPACITY_TOO_SMALL;
    // the kernels read every input strand while others are being written
    const uint64_t point_bytes = n * cvs * sizeof(nxg_packed_point);
    const void *outs[] = {output->points, output->point_counts, output->roots, output->root_uvs,
                          output->kept_count, output->status};
    const uint64_t out_bytes[] = {point_bytes, output->point_counts ? n * 4u : 0u,
                                  output->roots ? n * sizeof(nxg_root_sample) : 0u,
                                  output->root_uvs ? n * sizeof(nxg_vec2) : 0u, 4u, 4u};
    const void *ins[] = {curves->points, curves->point_counts, curves->roots, curves->root_uvs};
    const uint64_t in_bytes[] = {point_bytes, n * 4u, n * sizeof(nxg_root_sample), n * sizeof(nxg_vec2)};
    for (int o = 0; o < 6; ++o)
        for (int i = 0; i < 4; ++i)
            if (overlaps(outs[o], out_bytes[o], ins[i], in_bytes[i])) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    return NXG_OK;
}

uint64_t nxg_classic_cut_workspace_size(uint32_t strand_count) { return nxg::classic_layout(strand_count).total; }

uint32_t nxg_validate_classic_runtime(const nxg_classic_curves *curves, const nxg_classic_runtime_inputs *inputs,
                                      const nxg_classic_output *output) {
    if (!curves || !inputs || !output || !inputs->plan) return NXG_ERROR_NULL_ARGUMENT;
    const nxg_classic_plan_info &plan = *inputs->plan;
    const float rs = inputs->radius_scale;
    if (!std::isfinite(rs) || rs < 0.0f) return NXG_ERROR_INVALID_RADIUS_SCALE;
    const nxg_classic_cut_params common{0.0f, 0.0f, rs, 0u, nullptr};
    if (uint32_t st = nxg_validate_classic_cut(curves, &common, output); st != NXG_OK) return st;
    // the effect list (the packer's report; the device indexes modules by it)
    if (plan.effect_count > NXG_CLASSIC_MAX_EFFECTS || plan.clump_count > NXG_CLASSIC_MAX_CLUMPS ||
        plan.noise_count > NXG_CLASSIC_MAX_NOISES || plan.cut_count > NXG_CLASSIC_MAX_CUTS ||
        plan.has_cut != (plan.cut_count ? 1u : 0u) || plan.has_noise != (plan.noise_count ? 1u : 0u))
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    {
        const uint32_t counts[3] = {plan.clump_count, plan.noise_count, plan.cut_count};
        uint32_t seen[3] = {0u, 0u, 0u};
        for (uint32_t e = 0; e < plan.effect_count; ++e) {
            const nxg_classic_effect &x = plan.effects[e];
            if (x.type > NXG_CLASSIC_EFFECT_CUT || x.module >= counts[x.type] || ((seen[x.type] >> x.module) & 1u))
                return NXG_ERROR_INVALID_CLASSIC_REQUEST;
            seen[x.type] |= 1u << x.module;
        }
        for (uint32_t t = 0; t < 3u; ++t)
            if (seen[t] != (1u << counts[t]) - 1u) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
        const uint32_t trailing =
            plan.effect_count && plan.effects[plan.effect_count - 1u].type == NXG_CLASSIC_EFFECT_CUT ? 1u : 0u;
        if (plan.trailing_cut != trailing || plan.rewrites != (plan.effect_count > trailing ? 1u : 0u))
            return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    }
    // every module's programs present, none of an absent module
    {
        uint64_t need = 0u, allowed = (uint64_t{1} << NXG_CLASSIC_EXPR_CLUMP_MASK) - 1u;
        allowed &= ~(uint64_t{1} << NXG_CLASSIC_EXPR_CUT_AMOUNT);
        for (uint32_t m = 0; m < plan.clump_count; ++m)
            for (uint32_t k = 0; k < NXG_CLASSIC_CLUMP_EXPRS; ++k)
                need |= uint64_t{1} << NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_MASK + k);
        for (uint32_t n = 0; n < plan.noise_count; ++n)
            for (uint32_t k = 0; k < NXG_CLASSIC_NOISE_EXPRS; ++k)
                need |= uint64_t{1} << NXG_CLASSIC_EXPR_NOISE_OF(n, NXG_CLASSIC_EXPR_NOISE_MASK + k);
        for (uint32_t c = 0; c < plan.cut_count; ++c) need |= uint64_t{1} << NXG_CLASSIC_EXPR_CUT_AMOUNT_OF(c);
        allowed |= need;
        if ((plan.present & need) != need || (plan.present & ~allowed)) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    }
    // effects that change the CVs run on staged strands
    if (plan.rewrites && curves->cvs_per_strand > NXG_CLUMP_MAX_CVS) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (plan.noise_count && curves->strand_count && !inputs->surface_tangents)
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (plan.clump_count) {
        if (!inputs->clumps) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
        uint64_t guides = 0u;
        for (uint32_t m = 0; m < plan.clump_count; ++m) {
            const nxg_classic_clump_binding &b = inputs->clumps[m];
            const nxg_clump_guides *g = b.guides;
            if (!g || g->guide_count == 0u || !g->axes || !g->normals || !g->tangents || !b.guide_uvs ||
                !b.guide_random_prefixes || (curves->strand_count && !b.strand_guides))
                return NXG_ERROR_INVALID_CLASSIC_REQUEST;
            guides += g->guide_count;
        }
        if (guides > 0xFFFFFFFFu) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    }
    if (inputs->column_count > NXG_CLASSIC_MAX_COLUMNS || inputs->column_count < plan.columns_required ||
        (inputs->column_count != 0u) != (inputs->columns != nullptr))
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (curves->strand_count == 0u) return NXG_OK;
    // the reference's context reads every strand's root and, when a program
    // calls rand(), its SeExpr prefix (src/xgen_classic_runtime.cpp:1432-1491)
    if (!curves->roots) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (plan.needs_random_prefix && !inputs->random_prefixes) return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    if (plan.present && (!inputs->plan_image || reinterpret_cast<uintptr_t>(inputs->plan_image) % 16u != 0u))
        return NXG_ERROR_INVALID_CLASSIC_REQUEST;
    return NXG_OK;
}

uint64_t nxg_classic_runtime_workspace_size(const nxg_classic_plan_info *plan, uint32_t strand_count,
                                            uint32_t cvs_per_strand, uint32_t guide_count) {
    return nxg::classic_runtime_layout(strand_count, cvs_per_strand, guide_count, plan && plan->rewrites).total;
}
