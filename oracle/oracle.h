/*
 * ORACLE (test infrastructure only): C restatement of the reference's
 * strand-generation path, used by tests/, __graft_entry__.smoke() and
 * bench.py's CPU-baseline leg as the checker.  The shipped library
 * (nanoxgen_amd/) never links or calls it.
 *
 * Layout and types come from include/nanoxgen_amd.h (the .nxg format of
 * include/nanoxgen/types.h:77-101).
 */
#ifndef NXG_ORACLE_H
#define NXG_ORACLE_H

#include <stdint.h>
#include "../include/nanoxgen_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One generation request (host pointers).  The f32 driver writes float
 * points (4 per CV), nxg_root_sample roots and root_uvs; the f64 twin
 * writes double points (4 per CV) and 12 doubles per root. */
typedef struct orc_job {
    const void *blob;
    const nxg_deformed_geometry *deformed;
    const nxg_generation_params *params;
    const nxg_noise_octaves *octaves;
    const nxg_root_sample *explicit_roots;
    float radius_scale;
    void *points;
    void *roots;
    nxg_vec2 *root_uvs;
    uint32_t *point_counts;
    uint32_t *stencil_indices;
    float *stencil_weights;
    /* optional: generate only strands strand_list[0..list_count) into
     * output slots 0..list_count (explicit_roots then index by slot) */
    const uint32_t *strand_list;
    uint32_t list_count;
    /* optional per-root attribute multipliers, 4 floats per output slot */
    const float *attributes;
} orc_job;

/* generate.h:72-84 */
uint32_t orc_hash32(uint32_t x);
float orc_random01(uint32_t seed, uint32_t strand, uint32_t dimension);
/* generate.h:353-364 */
uint32_t orc_noise_hash(int32_t x, int32_t y, int32_t z);
/* asset.cpp:34-41 */
uint64_t orc_fnv1a(const void *bytes, uint64_t size);

float orc_gradient_noise_f32(const nxg_vec3 *gradients, float x, float y, float z);
double orc_gradient_noise_f64(const nxg_vec3 *gradients, double x, double y, double z);
void orc_generate_strand_f32(const orc_job *job, uint32_t strand, uint64_t slot);
void orc_generate_strand_f64(const orc_job *job, uint32_t strand, uint64_t slot);
void orc_strand_state_f32(const orc_job *job, uint32_t strand, float *out8);
void orc_strand_state_f64(const orc_job *job, uint32_t strand, double *out8);

/* asset.cpp:160-202: persistent 128-strand tiles claimed from an atomic
 * counter by `workers` threads (0 = min(hardware, ceil(tiles/4))).
 * precision: 32 runs the restatement, 64 the FP64 twin. */
int orc_generate(const orc_job *job, uint32_t precision, uint32_t workers,
                 uint32_t strands_per_block);
/* The same persistent-tile pool for any per-item function. */
void orc_run_tiles(uint32_t count, uint32_t workers, uint32_t block, void (*fn)(void *, uint32_t), void *ctx);

/* asset.cpp:229-293 (build_asset) into a malloc'd 64-aligned blob the
 * caller frees with orc_free.  Returns 0, or 1 with the reference's
 * exception message in err. */
int orc_build_asset(const nxg_asset_build_input *in, void **blob, uint64_t *size,
                    char *err, uint64_t err_capacity);
void orc_free(void *p);
/* asset.cpp:295-338: 0 and "" when valid, else 1 and the message. */
int orc_validate_asset(const void *blob, uint64_t size, char *msg, uint64_t msg_capacity);

void orc_sample_root_f32(const void *blob, uint32_t seed, uint32_t strand, nxg_root_sample *out);
void orc_root_on_triangle_f32(const void *blob, uint32_t seed, uint32_t strand, uint32_t tri,
                              nxg_root_sample *out);
/* The opt-in inverse-CDF sampler (oracle_cdf.c): prefix sums (T u64), one
 * strand's triangle, and the roots of strands 0..strand_count-1. */
void orc_cdf_prefix(const void *blob, uint64_t *prefix);
uint32_t orc_cdf_triangle(const uint64_t *prefix, uint32_t T, uint32_t seed, uint32_t strand);
void orc_cdf_roots(const void *blob, const uint64_t *prefix, uint32_t seed, uint32_t strand_count,
                   nxg_root_sample *roots);

/* float expression runtime (xgen_expression.cpp:637-722, 984-1148) */
uint32_t orc_seexpr_hash_finish(uint32_t state);
uint32_t orc_runtime_hash_component(float argument);
float orc_runtime_hash(const float *args, uint32_t n);
typedef struct orc_expr_ctx {
    const float *vars;  /* program input i reads vars[input_bindings[i]] */
    uint32_t var_count;
    float u, v, face_seed, t;
    uint32_t random_prefix;
    int has_random_prefix;
    const nxg_vec3 *gradients;  /* noise3 (NULL: noise3 is an error) */
    const nxg_texel_map *maps;
    uint32_t map_count;
} orc_expr_ctx;
int orc_expr_eval_ctx(const nxg_expr_program *prog, const orc_expr_ctx *ctx, float *scratch, float *out);
int orc_expr_eval(const nxg_expr_program *prog, const float *root_vars, float u, float v, float face_seed,
                  float t, const nxg_vec3 *gradients, const nxg_texel_map *maps, uint32_t map_count,
                  float *scratch, float *out);
/* Attributes of strands (all, or job->strand_list) into attributes[4*slot+k]:
 * the root variables and context of nxg_evaluate_root_attributes.
 * Returns the number of roots whose evaluation failed (their slots get NaN). */
uint32_t orc_root_attributes(const orc_job *job, const nxg_expr_program *programs, uint32_t program_count,
                             const nxg_texel_map *maps, uint32_t map_count, float *attributes,
                             uint32_t workers);

/* Classic CutFX chain (oracle_classic.c). */
float orc_curve_eval(const float *points4, uint32_t n, float t, float *out3);
float orc_curve_spline_length(const float *points4, uint32_t n);
void orc_classic_endpoints(const float *points, uint32_t strand_count, uint32_t cvs, float *out);

/* Classic ClumpingFX with constant module expressions (oracle_clump.c). */
typedef struct orc_clump_params {
    float mask, amount, clump_scale, noise, noise_scale, noise_frequency, noise_correlation;
} orc_clump_params;
void orc_clump_prepare(const float *axes3, uint32_t guides, uint32_t cvs, float *render3, float *lengths);
uint32_t orc_clump_apply(float *points, uint32_t strand_count, uint32_t cvs, const uint32_t *strand_guides,
                         const float *axes3, const float *render3, const float *lengths, const float *normals3,
                         const float *tangents3, const float *reference3, uint32_t guides,
                         const nxg_vec3 *gradients, const orc_clump_params *p, const float *amount_multiplier,
                         uint32_t multiplier_stride, uint32_t workers);
/* Classic primitive expressions and a CutFX amount expression
 * (oracle_classic.c, orc_classic_runtime).  Program inputs bind to
 * ORC_CLASSIC_VAR_* (3 + k: column k of the per-strand input row). */
#define ORC_CLASSIC_VAR_ID 0u
#define ORC_CLASSIC_VAR_CLENGTH 1u
#define ORC_CLASSIC_VAR_CWIDTH 2u
#define ORC_CLASSIC_MAX_COLUMNS 183u
#define ORC_CLASSIC_NONFINITE_AMOUNT 1u
#define ORC_CLASSIC_NONFINITE_WIDTH 2u
#define ORC_CLASSIC_NEGATIVE_LENGTH 8u
#define ORC_CLASSIC_NONFINITE_TAPER 16u
#define ORC_CLASSIC_NONFINITE_PROFILE 32u
#define ORC_CLASSIC_EXPRESSION 64u
#define ORC_CLASSIC_VARIABLE_COUNTS 128u
#define ORC_CLASSIC_INVALID_GUIDE 4u
#define ORC_CLASSIC_NONFINITE_CLUMP 256u
#define ORC_CLASSIC_NONFINITE_NOISE_PARAMETER 512u
#define ORC_CLASSIC_NONFINITE_NOISE 1024u
/* The clump's geometry: world-space guides, render axes and spline lengths
 * (orc_clump_prepare), the guides' uv and SeExpr prefixes, a guide per
 * strand (0xffffffff: none). */
typedef struct orc_clump_binding {
    const float *axes3, *render3, *lengths, *normals3, *tangents3, *reference3; /* reference3 optional */
    const nxg_vec2 *guide_uvs;
    const uint32_t *guide_prefixes;
    const uint32_t *strand_guides;
    uint32_t guide_count;
    const nxg_vec3 *gradients;
    /* optional translation-free guide CVs and their render axes
     * (guide_local_axes / guide_local_render_axes); lengths are theirs then */
    const float *local3, *local_render3;
} orc_clump_binding;
#define ORC_CLASSIC_MAX_CLUMPS 3u
#define ORC_CLASSIC_MAX_NOISES 2u
#define ORC_CLASSIC_MAX_CUTS 2u
#define ORC_CLASSIC_MAX_EFFECTS 8u
#define ORC_EFFECT_CLUMP 0u /* ClassicFloatEffectType order */
#define ORC_EFFECT_NOISE 1u
#define ORC_EFFECT_CUT 2u
typedef struct orc_classic_plan {
    const nxg_expr_program *length, *width, *taper, *taper_start, *width_ramp; /* NULL: absent */
    uint32_t description_id;  /* face seed: runtime hash of (id, name seed, face) */
    uint32_t name_seed;
    /* plan.effects: type (ORC_EFFECT_*) and module index of each, in order */
    uint32_t effect_count;
    uint32_t effect_type[ORC_CLASSIC_MAX_EFFECTS], effect_module[ORC_CLASSIC_MAX_EFFECTS];
    /* per ClumpingFX module: mask, clump, clumpScale, noise, noiseScale, noiseFrequency, noiseCorrelation */
    const nxg_expr_program *clump[ORC_CLASSIC_MAX_CLUMPS][7];
    const orc_clump_binding *binding[ORC_CLASSIC_MAX_CLUMPS];
    /* per NoiseFX module: mask, magnitude, magnitudeScale, frequency, correlation, preserveLength */
    const nxg_expr_program *noise[ORC_CLASSIC_MAX_NOISES][6];
    const nxg_expr_program *cut_amount[ORC_CLASSIC_MAX_CUTS]; /* per CutFX module */
    int root_relative;                              /* clump goals about each strand's root */
    const float *surface_tangents3, *noise_domain3; /* per strand; noise_domain3 optional */
    const nxg_vec3 *gradients;                      /* the SeExpr gradient table (with a NoiseFX) */
} orc_classic_plan;
void orc_noise_fx(float *points3, uint32_t n, float frequency, float correlation, float preserve,
                  const float *domain_position3, const float *surface_normal3, const float *surface_tangent3,
                  const float *magnitudes, const nxg_vec3 *gradients);
void orc_seexpr_gradients(float *out); /* 256 x 3 */
void orc_clump_noise_axis(const float *axis3, const float *distance3, uint32_t n, float raw_frequency,
                          float raw_correlation, const float *domain3, const float *normal3, const float *tangent3,
                          const float *magnitudes, const nxg_vec3 *gradients, float *out3);
void orc_cut_to_length(float *points3, uint32_t n, float target);
/* Returns the kept count (outputs compacted in order), or ~0u with the
 * ORC_CLASSIC_* bits of every strand that raised one in *error_bits.
 * out_points and keep hold strand_count entries; strands run in parallel
 * over `workers` threads (0 = all cores), with identical results. */
uint32_t orc_classic_runtime(const float *points, const uint32_t *counts, const nxg_root_sample *roots,
                             const nxg_vec2 *uvs, uint32_t strand_count, uint32_t cvs,
                             const orc_classic_plan *plan, const uint32_t *primitive_ids,
                             const uint32_t *random_prefixes, const float *columns, uint32_t column_count,
                             float radius_scale, float *out_points, uint32_t *out_counts,
                             nxg_root_sample *out_roots, nxg_vec2 *out_uvs, uint8_t *keep, uint32_t *error_bits,
                             uint32_t workers);
uint32_t orc_classic_cut(const float *points, const uint32_t *counts, const nxg_root_sample *roots,
                         const nxg_vec2 *uvs, uint32_t strand_count, uint32_t cvs, float amount_scale,
                         float amount_offset, const float *multiplier, float radius_scale, float *out_points,
                         uint32_t *out_counts, nxg_root_sample *out_roots, nxg_vec2 *out_uvs, uint8_t *keep,
                         uint32_t workers);

#ifdef __cplusplus
}
#endif
#endif
