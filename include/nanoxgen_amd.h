/*
 * nanoxgen_amd.h -- C ABI of the MI355X (gfx950) strand generator.
 *
 * Every type and entry point here replaces one piece of the reference's C++
 * interface for the strand-generation path; the reference file:line is given
 * beside each.  All functions are plain C: pointers, sizes and POD structs,
 * no exceptions, no hidden allocation, no hidden synchronisation.  Device
 * pointers are owned by the caller; work is enqueued on the caller's stream.
 *
 * Status codes 0..20 are the reference's DeviceGenerationError values
 * (include/nanoxgen/device_contract.h:58-80) verbatim; 21.. are extensions.
 */
#ifndef NANOXGEN_AMD_H
#define NANOXGEN_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- .nxg blob format (include/nanoxgen/types.h:17-22, 77-101) ---------- */
#define NXG_MAGIC 0x4e4547584f4e414eull /* "NANOXGEN", little endian */
#define NXG_VERSION_MAJOR 0u
#define NXG_VERSION_MINOR 2u
#define NXG_GUIDE_STENCIL_SIZE 8u
#define NXG_NOISE_GRADIENT_COUNT 256u
#define NXG_INVALID_INDEX 0xffffffffu
#define NXG_BLOB_ALIGNMENT 64u
#define NXG_ASSET_HAS_NORMALS 1u   /* AssetFlags::HasNormals   types.h:71 */
#define NXG_ASSET_HAS_TEXCOORDS 2u /* AssetFlags::HasTexcoords types.h:72 */
/* Extension: at most this many stacked noise octaves per strand. */
#define NXG_MAX_NOISE_OCTAVES 4u

typedef struct nxg_vec2 { float x, y; } nxg_vec2;         /* types.h:24-27 */
typedef struct nxg_vec3 { float x, y, z; } nxg_vec3;      /* types.h:29-33 */
typedef struct nxg_uint3 { uint32_t x, y, z; } nxg_uint3; /* types.h:35-39 */

/* Walker alias bucket (types.h:53-56). */
typedef struct nxg_alias_entry {
    float probability;
    uint32_t alias;
} nxg_alias_entry;

/* 56-byte guide record (types.h:58-68). */
typedef struct nxg_guide_record {
    uint32_t first_cv;
    uint16_t cv_count;
    uint16_t flags;
    uint32_t triangle_index;
    float support_radius;
    nxg_vec2 barycentric; /* weights of triangle vertices y and z */
    nxg_vec2 root_uv;
    nxg_vec3 root_position;
    nxg_vec3 root_normal;
} nxg_guide_record;

/* 128-byte blob header (types.h:77-101).  The blob base must be 64-byte
 * aligned in memory; this ABI struct itself carries no over-alignment so it
 * can be embedded by value in descriptors. */
typedef struct nxg_asset_header {
    uint64_t magic;
    uint16_t version_major;
    uint16_t version_minor;
    uint32_t flags;
    uint64_t byte_size;
    uint64_t content_hash;
    uint32_t vertex_count;
    uint32_t triangle_count;
    uint32_t guide_count;
    uint32_t guide_cv_count;
    uint32_t guide_stencil_size;
    uint32_t noise_gradient_count;
    uint64_t positions_offset;
    uint64_t normals_offset;
    uint64_t texcoords_offset;
    uint64_t triangles_offset;
    uint64_t alias_table_offset;
    uint64_t guides_offset;
    uint64_t guide_cvs_offset;
    uint64_t triangle_guides_offset;
    uint64_t noise_gradients_offset;
} nxg_asset_header;

/* GenerationParams (types.h:108-124), field for field. */
typedef struct nxg_generation_params {
    uint32_t strand_count;
    uint32_t cvs_per_strand;
    uint32_t seed;
    uint32_t reserved;
    float guide_support_scale;
    float guide_weight_power;
    float normal_rejection_cos;
    float length_scale;
    float root_width;
    float tip_width;
    float noise_amplitude;
    float noise_frequency;
    float noise_mask;
    float noise_correlation;
    float noise_preserve_length;
} nxg_generation_params;

/* Extension (no reference counterpart): stacked noise octaves.  count 0 or 1
 * is the reference's single octave (noise_amplitude, noise_frequency).  For
 * count k > 1, amplitude[i]/frequency[i] (i < k) replace the single pair and
 * the per-octave displacements are summed in the one transported frame. */
typedef struct nxg_noise_octaves {
    uint32_t count;
    float amplitude[NXG_MAX_NOISE_OCTAVES];
    float frequency[NXG_MAX_NOISE_OCTAVES];
} nxg_noise_octaves;

/* RootSample (types.h:126-135): 48 bytes. */
typedef struct nxg_root_sample {
    nxg_vec3 position;
    nxg_vec3 normal;
    nxg_vec2 uv;
    uint32_t triangle_index;
    nxg_vec2 barycentric;
    uint32_t surface_face_id;
} nxg_root_sample;

/* PackedCurvePoint (types.h:139-147): float4(position, radius). */
typedef struct nxg_packed_point { float x, y, z, radius; } nxg_packed_point;

/* ---- checked device descriptors (device_contract.h:14-56) -------------- */

/* DeviceAssetDescriptor (device_contract.h:14-18): device pointer to the
 * uploaded blob, the validated host mirror of its header, and the device
 * allocation size in bytes. */
typedef struct nxg_device_asset {
    const void *data;
    nxg_asset_header header;
    uint64_t byte_capacity;
} nxg_device_asset;

/* DeviceDeformedGeometryDescriptor (device_contract.h:20-25, types.h:158-162).
 * A null pointer selects the asset's rest array; its count must then be 0. */
typedef struct nxg_deformed_geometry {
    const nxg_vec3 *positions;
    const nxg_vec3 *normals;
    const nxg_vec3 *guide_cvs;
    uint64_t position_count;
    uint64_t normal_count;
    uint64_t guide_cv_count;
} nxg_deformed_geometry;

/* DevicePackedCurveOutputDescriptor (device_contract.h:34-40, types.h:164-170).
 * Capacities are in elements.  points is required; the others are optional
 * and a null pointer must come with capacity 0. */
typedef struct nxg_packed_output {
    nxg_packed_point *points;
    nxg_root_sample *roots;
    nxg_vec2 *root_uvs;
    float radius_scale;
    uint32_t *point_counts;
    uint64_t point_capacity;
    uint64_t root_capacity;
    uint64_t root_uv_capacity;
    uint64_t point_count_capacity;
} nxg_packed_output;

/* Extension: per-strand guide-blend record (GuideBlendState.indices/weights,
 * generate.h:207-212), strand-major, NXG_GUIDE_STENCIL_SIZE entries per
 * strand.  Slots the blend rejected hold NXG_INVALID_INDEX and weight 0. */
typedef struct nxg_stencil_output {
    uint32_t *indices;
    float *weights;
    uint64_t capacity; /* entries, >= strand_count * 8 */
} nxg_stencil_output;

/* DeviceLaunchConfig (device_contract.h:54-56).  block_size is validated to
 * [1, 1024] as in the reference; the gfx950 kernels use their own tile of
 * four 64-lane wavefronts per workgroup whatever its value. */
typedef struct nxg_launch_config {
    uint32_t block_size;
} nxg_launch_config;

/* ---- status codes ------------------------------------------------------ */
enum {
    NXG_OK = 0,                              /* DeviceGenerationError::None */
    NXG_ERROR_NULL_ASSET = 1,
    NXG_ERROR_MISALIGNED_ASSET = 2,
    NXG_ERROR_INVALID_ASSET_METADATA = 3,
    NXG_ERROR_ASSET_CAPACITY_TOO_SMALL = 4,
    NXG_ERROR_INVALID_GENERATION_PARAMETERS = 5,
    NXG_ERROR_INVALID_BLOCK_SIZE = 6,
    NXG_ERROR_DEFORMED_POSITION_COUNT_MISMATCH = 7,
    NXG_ERROR_DEFORMED_NORMAL_COUNT_MISMATCH = 8,
    NXG_ERROR_DEFORMED_GUIDE_CV_COUNT_MISMATCH = 9,
    NXG_ERROR_MISSING_OUTPUT = 10,
    NXG_ERROR_POINT_CAPACITY_TOO_SMALL = 11,
    NXG_ERROR_WIDTH_CAPACITY_TOO_SMALL = 12,
    NXG_ERROR_ROOT_CAPACITY_TOO_SMALL = 13,
    NXG_ERROR_ROOT_UV_CAPACITY_TOO_SMALL = 14,
    NXG_ERROR_POINT_COUNT_CAPACITY_TOO_SMALL = 15,
    NXG_ERROR_INVALID_RADIUS_SCALE = 16,
    NXG_ERROR_MISSING_MOTION_SAMPLES = 17,
    NXG_ERROR_INVALID_MOTION_TIMES = 18,
    NXG_ERROR_MOTION_POINT_CAPACITY_TOO_SMALL = 19,
    NXG_ERROR_OUTPUT_SIZE_OVERFLOW = 20,
    /* extensions */
    NXG_ERROR_HIP = 21,                /* a HIP runtime call failed */
    NXG_ERROR_WORKSPACE_TOO_SMALL = 22,
    NXG_ERROR_INVALID_OCTAVES = 23,
    NXG_ERROR_NULL_ARGUMENT = 24,
    NXG_ERROR_STENCIL_CAPACITY_TOO_SMALL = 25,
    NXG_ERROR_INVALID_ROOTS = 26,
    NXG_ERROR_INVALID_BUILD_INPUT = 27,
    NXG_ERROR_BUFFER_TOO_SMALL = 28,
    NXG_ERROR_INVALID_ASSET = 29,
    NXG_ERROR_INVALID_PROGRAM = 30,
    NXG_ERROR_ATTRIBUTE_CAPACITY_TOO_SMALL = 31,
    NXG_ERROR_INVALID_MLP_REQUEST = 32,
    NXG_ERROR_INVALID_BATCH = 33,
    NXG_ERROR_INVALID_CLASSIC_REQUEST = 34,
    NXG_ERROR_SCRATCH_LIMIT = 35,      /* request exceeds the static scratch of the no-alloc entry point */
    NXG_ERROR_INVALID_SAMPLER = 36     /* inverse-CDF sampler buffer missing, misaligned or too small */
};

/* device_generation_error_message (device_contract.h:82-116); extension codes
 * get messages of their own.  Never returns NULL. */
const char *nxg_status_message(uint32_t status);

/* ---- host asset builder (asset.h:14-44, 119-125; asset.cpp:229-378) --- */

/* Flattened AssetBuildInput/GuideInput (asset.h:14-44).  Optional arrays
 * may be NULL.  Guide i owns guide_cv_counts[i] consecutive CVs of
 * guide_cvs; its first CV is its root. */
typedef struct nxg_asset_build_input {
    const nxg_vec3 *positions;
    uint32_t vertex_count;
    uint32_t triangle_count;
    const nxg_vec3 *normals;          /* optional: vertex_count */
    const nxg_vec2 *texcoords;        /* optional: vertex_count */
    const nxg_uint3 *triangles;
    uint32_t guide_count;
    uint32_t stencil_k;               /* extension: 0 means 8 (reference) */
    const uint32_t *guide_cv_counts;  /* guide_count entries, each 2..65535 */
    const nxg_vec3 *guide_cvs;        /* sum(guide_cv_counts) entries */
    const nxg_vec3 *guide_root_normals;  /* optional: default (0,1,0) */
    const nxg_vec2 *guide_root_uvs;      /* optional */
    const uint32_t *guide_triangles;     /* optional: default invalid */
    const nxg_vec2 *guide_barycentrics;  /* optional */
    const float *guide_support_radii;    /* optional: 0 selects automatic */
    const float *triangle_weights;       /* extension, optional: alias
                                            weights are area * weight */
} nxg_asset_build_input;

/* Size in bytes of the blob nxg_asset_build would write (cheap). */
uint32_t nxg_asset_build_size(const nxg_asset_build_input *input, uint64_t *byte_size);

/* build_asset (asset.cpp:229-293) into caller memory `out` (64-byte
 * aligned, capacity bytes).  On a refused input, writes the reference's
 * std::invalid_argument message into err (if non-NULL). */
uint32_t nxg_asset_build(const nxg_asset_build_input *input, void *out, uint64_t capacity,
                         uint64_t *byte_size, char *err, uint64_t err_capacity);

/* build_asset (asset.cpp:75-293) with its O(triangles x guides) stencil
 * search (asset.cpp:252-269) run on the GPU: the host writes the blob into
 * `out` (as nxg_asset_build), copies it to `device_out` (64-byte aligned,
 * device_capacity bytes, caller-owned), the stencil kernel fills the
 * device copy's stencil section, and the section and the final content
 * hash come back so that both copies are the blob nxg_asset_build writes,
 * byte for byte.  Synchronises `stream`. */
uint32_t nxg_asset_build_gpu(const nxg_asset_build_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                             char *err, uint64_t err_capacity, void *device_out, uint64_t device_capacity,
                             void *stream);

/* validate_asset (asset.cpp:295-338): writes "" or the reference's message
 * into msg; returns NXG_OK or NXG_ERROR_INVALID_ASSET. */
uint32_t nxg_asset_validate(const void *blob, uint64_t byte_size, char *msg, uint64_t msg_capacity);

/* make_device_asset_descriptor (asset.cpp:364-378): validates the host blob
 * and fills the descriptor for its device copy. */
uint32_t nxg_make_device_asset(const void *host_blob, uint64_t host_byte_size,
                               const void *device_data, uint64_t device_byte_capacity,
                               nxg_device_asset *out);

/* ---- device generation (device_contract.h:221-254; generate.h:612-676) -- */

/* validate_device_packed_generation_request (device_contract.h:221-254).
 * deformed and config may be NULL (reference defaults). */
uint32_t nxg_validate_packed_generation(const nxg_device_asset *asset,
                                        const nxg_deformed_geometry *deformed,
                                        const nxg_generation_params *params,
                                        const nxg_packed_output *output,
                                        const nxg_launch_config *config);

/* Bytes of device workspace nxg_generate_packed needs for this request. */
uint64_t nxg_packed_generation_workspace_size(const nxg_asset_header *header,
                                              const nxg_generation_params *params);

/* generate_packed_cpu's device equivalent: runs generate_packed_strand
 * (generate.h:612-642) for every strand on `stream` (a hipStream_t; NULL is
 * the null stream).  Validates first, launches nothing on error, allocates
 * nothing and does not synchronise.  octaves, stencil, deformed and config
 * may be NULL. */
uint32_t nxg_generate_packed(const nxg_device_asset *asset,
                             const nxg_deformed_geometry *deformed,
                             const nxg_generation_params *params,
                             const nxg_noise_octaves *octaves,
                             const nxg_packed_output *output,
                             const nxg_stencil_output *stencil,
                             const nxg_launch_config *config,
                             void *workspace, uint64_t workspace_bytes,
                             void *stream);

/* Bytes of the library's static device scratch, per GPU (32 MiB): the
 * largest nxg_packed_generation_workspace_size nxg_generate_packed_noalloc
 * accepts. */
#define NXG_STATIC_SCRATCH_BYTES (32ull << 20)

/* nxg_generate_packed without a workspace argument and without any device
 * allocation: the workspace is the library's static device scratch of the
 * current GPU (a file-scope device array of NXG_STATIC_SCRATCH_BYTES, held
 * until the process ends).  No counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, serving loops
 * that forbid hipMalloc).  Requests whose workspace would exceed the scratch
 * return NXG_ERROR_SCRATCH_LIMIT before any launch.  Every call on one GPU
 * shares its scratch, and a call leaves its contents there: calls that
 * overlap on one GPU (different streams without ordering between them) race.
 * The first call on a device looks up that device's scratch address. */
uint32_t nxg_generate_packed_noalloc(const nxg_device_asset *asset,
                                     const nxg_deformed_geometry *deformed,
                                     const nxg_generation_params *params,
                                     const nxg_noise_octaves *octaves,
                                     const nxg_packed_output *output,
                                     const nxg_stencil_output *stencil,
                                     const nxg_launch_config *config,
                                     void *stream);

/* The two stream-ordered halves of nxg_generate_packed.  prepare fills the
 * workspace with the per-guide state and the guide offset table, which
 * depend only on the asset, the deformed geometry and cvs_per_strand (the
 * reference recomputes them inside interpolate_offset, generate.h:246-267),
 * and with the processing order of the strands (grouped by root triangle;
 * it depends on the asset, seed and strand_count and changes no output);
 * generate_packed_prepared runs the strand kernel from a workspace that a
 * prepare call with the same asset, deformed geometry, cvs_per_strand, seed
 * and strand_count filled earlier on the same stream (or one ordered before
 * it). */
uint32_t nxg_prepare_packed_generation(const nxg_device_asset *asset,
                                       const nxg_deformed_geometry *deformed,
                                       const nxg_generation_params *params,
                                       void *workspace, uint64_t workspace_bytes,
                                       void *stream);
uint32_t nxg_generate_packed_prepared(const nxg_device_asset *asset,
                                      const nxg_deformed_geometry *deformed,
                                      const nxg_generation_params *params,
                                      const nxg_noise_octaves *octaves,
                                      const nxg_packed_output *output,
                                      const nxg_stencil_output *stencil,
                                      const nxg_launch_config *config,
                                      void *workspace, uint64_t workspace_bytes,
                                      void *stream);

/* ---- opt-in inverse-CDF root sampler ------------------------------------
 * No reference counterpart: an alternative to the alias-table triangle draw
 * of sample_root (generate.h:107-112) that samples the same distribution
 * through a prefix sum and a binary search.  It sends a strand to a
 * different triangle than the alias draw (the barycentric draw and
 * everything after it are the alias mode's), so its outputs match the
 * reference only in distribution; use it only when that is what is wanted.
 * The distribution is read off the blob's alias table in integers: bucket j
 * keeps ceil(prob_j 2^24) 2^8 and hands the rest of its 2^32 to its alias;
 * the inclusive prefix sums of those masses (u64, T of them, summing to
 * T 2^32) are the table.  A strand's target is floor((h0 2^32 + h1) T / 2^32)
 * with h0, h1 the hashes of random01's dimensions 0 and 1, and its triangle
 * the first whose prefix sum exceeds the target.
 *
 * nxg_cdf_sampler_size: bytes of the sampler buffer for an asset (0 for a
 * null header or no triangles).  The buffer (device memory, 256-byte
 * aligned) starts with the T prefix sums.
 * nxg_cdf_sampler_build: fills the buffer from the device asset's alias
 * table (five kernels on `stream`).  It depends on the asset only.
 * nxg_generate_packed_cdf: nxg_generate_packed with the roots drawn from the
 * built sampler; output->roots is required and receives them. */
uint64_t nxg_cdf_sampler_size(const nxg_asset_header *header);
uint32_t nxg_cdf_sampler_build(const nxg_device_asset *asset, void *sampler, uint64_t sampler_bytes,
                               void *stream);
uint32_t nxg_generate_packed_cdf(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                                 const nxg_generation_params *params, const nxg_noise_octaves *octaves,
                                 const void *sampler, uint64_t sampler_bytes, const nxg_packed_output *output,
                                 const nxg_stencil_output *stencil, const nxg_launch_config *config,
                                 void *workspace, uint64_t workspace_bytes, void *stream);

/* generate_packed_roots_cpu's device equivalent (asset.cpp:451-501,
 * generate.h:644-676): strands start from caller-supplied device roots
 * (strand_count records).  Root validity is the caller's contract on device
 * (the host-side check of asset.cpp:468-483 cannot read device memory). */
uint32_t nxg_generate_packed_from_roots(const nxg_device_asset *asset,
                                        const nxg_generation_params *params,
                                        const nxg_noise_octaves *octaves,
                                        const nxg_root_sample *roots,
                                        const nxg_packed_output *output,
                                        const nxg_launch_config *config,
                                        void *workspace, uint64_t workspace_bytes,
                                        void *stream);

/* ---- float expression runtime (xgen_expression.h:12-100) --------------- */

/* XgenScalarOp (xgen_expression.h:12-40) values 0..26 verbatim; 27.. are
 * extensions with no reference counterpart:
 *   POW(a, b)        powf(a, b)
 *   MIX(a, b, t)     fit(t, 0, 1, a, b), operation for operation
 *   NOISE3(x, y, z)  gradient_noise (generate.h:370-402), value in [0, 1]
 *   TEXEL(f, u, v)   point sample of texel map `auxiliary` (nxg_texel_map):
 *                    face clamp(floor(f), 0, faces-1), column
 *                    clamp(floor(u*R), 0, R-1), row clamp(floor(v*R), 0, R-1) */
enum {
    NXG_OP_CONSTANT = 0, NXG_OP_INPUT, NXG_OP_NEGATE, NXG_OP_LOGICAL_NOT, NXG_OP_ADD,
    NXG_OP_SUBTRACT, NXG_OP_MULTIPLY, NXG_OP_DIVIDE, NXG_OP_LESS, NXG_OP_LESS_EQUAL,
    NXG_OP_GREATER, NXG_OP_GREATER_EQUAL, NXG_OP_EQUAL, NXG_OP_NOT_EQUAL, NXG_OP_LOGICAL_AND,
    NXG_OP_LOGICAL_OR, NXG_OP_SELECT, NXG_OP_HASH, NXG_OP_RANDOM, NXG_OP_MINIMUM, NXG_OP_MAXIMUM,
    NXG_OP_CLAMP, NXG_OP_FIT, NXG_OP_GAMMA, NXG_OP_CONTRAST, NXG_OP_SMOOTHSTEP, NXG_OP_RAMP,
    NXG_OP_POW = 27, NXG_OP_MIX = 28, NXG_OP_NOISE3 = 29, NXG_OP_TEXEL = 30,
    NXG_OP_COUNT = 31
};

/* XgenFloatScalarInstruction (xgen_expression.h:83-89): 16 bytes. */
typedef struct nxg_expr_instruction {
    uint8_t op;
    uint8_t reserved;
    uint16_t operand_count;
    uint32_t operand_offset;
    uint32_t auxiliary;
    float immediate;
} nxg_expr_instruction;

/* XgenFloatRampPoint / XgenRamp (xgen_expression.h:48-57). */
typedef struct nxg_expr_ramp_point { float position; float value; uint32_t interpolation; } nxg_expr_ramp_point;
typedef struct nxg_expr_ramp { uint32_t point_offset; uint32_t point_count; } nxg_expr_ramp;

/* Per-root variables a program input can bind to (the generator's analogue
 * of the Classic runtime's input binding, xgen_classic_runtime.cpp:504-553). */
enum {
    NXG_ROOT_VAR_POSITION_X = 0, NXG_ROOT_VAR_POSITION_Y, NXG_ROOT_VAR_POSITION_Z, /* rest pose */
    NXG_ROOT_VAR_NORMAL_X, NXG_ROOT_VAR_NORMAL_Y, NXG_ROOT_VAR_NORMAL_Z,
    NXG_ROOT_VAR_U, NXG_ROOT_VAR_V,      /* root texcoord */
    NXG_ROOT_VAR_FACE,                   /* triangle index */
    NXG_ROOT_VAR_BARY1, NXG_ROOT_VAR_BARY2,
    NXG_ROOT_VAR_ID,                     /* strand index */
    NXG_ROOT_VAR_COUNT
};

/* XgenFloatExpressionProgram (xgen_expression.h:91-100), flattened, host
 * memory.  Program input i reads root variable input_bindings[i]. */
typedef struct nxg_expr_program {
    const nxg_expr_instruction *instructions;
    const uint32_t *operands;
    const nxg_expr_ramp_point *ramp_points;
    const nxg_expr_ramp *ramps;
    const uint32_t *input_bindings;
    uint32_t instruction_count;
    uint32_t operand_count;
    uint32_t ramp_point_count;
    uint32_t ramp_count;
    uint32_t input_count;
    uint32_t result;
    uint32_t expression_seed;
    uint32_t random_call_count;
} nxg_expr_program;

/* A face-major R x R float texel map (device memory for the GPU entry
 * points, host memory for the CPU ones). */
typedef struct nxg_texel_map {
    const float *texels;
    uint32_t face_count;
    uint32_t resolution;
} nxg_texel_map;

/* Per-root attribute slots (4 floats per strand, strand-major).  Program k
 * of a set writes slot k; slots without a program hold 1.0.  Generation
 * multiplies length_scale by slot 0, the width profile by slot 1 and every
 * octave's noise amplitude by slot 2; slot 3 is carried for consumers. */
enum { NXG_ATTR_LENGTH = 0, NXG_ATTR_WIDTH = 1, NXG_ATTR_NOISE = 2, NXG_ATTR_CLUMP = 3, NXG_ATTR_COUNT = 4 };
#define NXG_EXPR_MAX_VALUES 256u
#define NXG_EXPR_MAX_PROGRAMS 4u
#define NXG_EXPR_MAX_TEXEL_MAPS 4u

/* The expression compiler (compile_xgen_scalar_expression followed by
 * make_xgen_float_expression_program, src/xgen_expression.cpp:50-557,
 * 771-818; XgenExpressionCompileOptions, xgen_expression.h:102-110): XGen /
 * SeExpr scalar expression text to this float IR, instruction for
 * instruction the reference's program, with its diagnostics.  Limits of 0
 * take the reference's defaults.  extensions = 1 also accepts pow(a, b),
 * mix(a, b, t), noise(x, y, z) and texel0..texel3(f, u, v) (the POW, MIX,
 * NOISE3 and TEXEL ops). */
typedef struct nxg_expr_compile_options {
    const char *expression_name;  /* with object_type, seeds rand() call sites */
    const char *object_type;
    uint64_t max_source_bytes;
    uint64_t max_tokens;
    uint64_t max_instructions;
    uint64_t max_locals;
    uint64_t max_ramp_points;
    uint32_t extensions;
    uint32_t reserved;
} nxg_expr_compile_options;

/* Caller buffers with capacities in elements (input_names: bytes, the
 * program's $-variable names NUL-terminated back to back, in input order);
 * every count is written even when a buffer is too small
 * (NXG_ERROR_BUFFER_TOO_SMALL), so a first call with zero capacities sizes
 * the second. */
typedef struct nxg_expr_compiled {
    nxg_expr_instruction *instructions;
    uint32_t *operands;
    nxg_expr_ramp_point *ramp_points;
    nxg_expr_ramp *ramps;
    char *input_names;
    uint32_t instruction_capacity;
    uint32_t operand_capacity;
    uint32_t ramp_point_capacity;
    uint32_t ramp_capacity;
    uint32_t input_name_capacity;
    uint32_t instruction_count;
    uint32_t operand_count;
    uint32_t ramp_point_count;
    uint32_t ramp_count;
    uint32_t input_count;
    uint32_t input_name_bytes;
    uint32_t result;
    uint32_t expression_seed;
    uint32_t random_call_count;
} nxg_expr_compiled;

/* NXG_OK, NXG_ERROR_INVALID_PROGRAM (the reference's message in err) or
 * NXG_ERROR_BUFFER_TOO_SMALL. */
uint32_t nxg_expr_compile(const char *source, uint64_t source_bytes, const nxg_expr_compile_options *options,
                          nxg_expr_compiled *out, char *err, uint64_t err_capacity);

/* xgen_seexpr_hash, xgen_string_seed, xgen_face_seed
 * (src/xgen_expression.cpp:659-661, 738-769). */
double nxg_seexpr_hash(const double *arguments, uint32_t count);
uint32_t nxg_string_seed(const char *text);
double nxg_face_seed(uint32_t description_id, const char *patch_name, uint32_t face_id);

/* Checks one program the way the reference interpreter does while it runs
 * (operand ranges, SSA order, arity, input/ramp indices, result index), plus
 * the runtime limit NXG_EXPR_MAX_VALUES.  NXG_OK or NXG_ERROR_INVALID_PROGRAM. */
uint32_t nxg_expr_validate(const nxg_expr_program *program);

/* Host descriptor of a packed program set: nxg_expr_pack fills everything
 * but `image`, which the caller points at its device copy of the packed
 * bytes (16-byte aligned). */
typedef struct nxg_expr_set {
    const void *image;
    uint64_t byte_size;
    uint32_t program_count;
    uint32_t max_values;
    uint32_t program_offset[NXG_EXPR_MAX_PROGRAMS];
    uint32_t texel_maps_required; /* 1 + the largest texel-map index used */
} nxg_expr_set;

/* Packs up to NXG_EXPR_MAX_PROGRAMS validated programs into one position-
 * independent byte image for device upload.  With out == NULL only
 * set->byte_size is written (size query). */
uint32_t nxg_expr_pack(const nxg_expr_program *programs, uint32_t program_count, void *out,
                       uint64_t capacity, nxg_expr_set *set);

/* Evaluates the set's programs for every strand's root (sample_root,
 * generate.h:101-135) on the GPU: attributes[4*s + k] = program k at root s.
 * Context per root: u, v = root texcoord, face_seed = runtime hash of
 * (0, 0, face), t = 0, no random prefix.  A non-finite value (an error in
 * the reference interpreter) writes NaN for that root and sets bit 0 of
 * *error_word (optional, device memory, never cleared by the library). */
uint32_t nxg_evaluate_root_attributes(const nxg_device_asset *asset,
                                      const nxg_generation_params *params,
                                      const nxg_expr_set *set,
                                      const nxg_texel_map *maps, uint32_t map_count,
                                      float *attributes, uint64_t attribute_capacity,
                                      uint32_t *error_word, void *stream);

/* nxg_generate_packed with per-root attribute multipliers (device memory,
 * 4 floats per strand; see NXG_ATTR_*).  All-ones attributes reproduce
 * nxg_generate_packed bit for bit. */
uint32_t nxg_generate_packed_attributed(const nxg_device_asset *asset,
                                        const nxg_deformed_geometry *deformed,
                                        const nxg_generation_params *params,
                                        const nxg_noise_octaves *octaves,
                                        const float *attributes, uint64_t attribute_capacity,
                                        const nxg_packed_output *output,
                                        const nxg_stencil_output *stencil,
                                        const nxg_launch_config *config,
                                        void *workspace, uint64_t workspace_bytes,
                                        void *stream);

/* The strand half of nxg_generate_packed_attributed, from a workspace that
 * nxg_prepare_packed_generation filled (the same rule as
 * nxg_generate_packed_prepared). */
uint32_t nxg_generate_packed_attributed_prepared(const nxg_device_asset *asset,
                                                 const nxg_deformed_geometry *deformed,
                                                 const nxg_generation_params *params,
                                                 const nxg_noise_octaves *octaves,
                                                 const float *attributes, uint64_t attribute_capacity,
                                                 const nxg_packed_output *output,
                                                 const nxg_stencil_output *stencil,
                                                 const nxg_launch_config *config,
                                                 void *workspace, uint64_t workspace_bytes,
                                                 void *stream);

/* ---- attribute surrogate (extension, no reference counterpart) --------- */

/* A 10 -> 64 -> 64 -> 64 -> 64 -> 4 MLP with SiLU hidden activations, BF16
 * weights, BF16 activations carried as two terms (hi = bf16(h),
 * lo = bf16(h - hi), two MFMAs per product), FP32 accumulation (gfx950
 * v_mfma_f32_16x16x32_bf16).  Inputs are the root features written by
 * nxg_root_features; outputs are the NXG_ATTR_* attributes.  All state lives
 * in one caller-owned device buffer laid out by nxg_mlp_get_layout: FP32
 * master parameters (input-major padded weights 16x64, 64x64 x3, 64x16, each
 * followed by its bias), Adam moments, the BF16 parameter copy, the BF16
 * weights in MFMA operand order, normalisation (input mean[10], input
 * 1/std[10], output scale[4], output offset[4]), the last loss and per-block
 * gradient partials.  The caller uploads parameters and normalisation, then
 * calls nxg_mlp_sync; every training step refreshes both BF16 copies. */
#define NXG_MLP_INPUTS 10u
#define NXG_MLP_HIDDEN 64u
#define NXG_MLP_HIDDEN_LAYERS 4u
#define NXG_MLP_OUTPUTS 4u
#define NXG_MLP_MAX_BATCH 8192u

typedef struct nxg_mlp_layout {
    uint64_t param_count;
    uint64_t params_offset;
    uint64_t adam_m_offset;
    uint64_t adam_v_offset;
    uint64_t params_bf16_offset;
    uint64_t fragments_offset; /* BF16 weights in MFMA operand order (inference) */
    uint64_t norm_offset;
    uint64_t loss_offset;
    uint64_t partials_offset; /* [batch/64][param_count + 1] FP32 */
    uint64_t max_batch;
    uint64_t total_bytes;
} nxg_mlp_layout;

void nxg_mlp_get_layout(nxg_mlp_layout *layout);

/* Refresh both BF16 copies (flat and MFMA operand order) from the FP32
 * master parameters. */
uint32_t nxg_mlp_sync(void *state, void *stream);

/* Forward + backward on samples [batch_first, batch_first + batch_size)
 * (modulo sample_count) of features [n][10] / targets [n][4] (attributes),
 * mean-squared loss on normalised targets; leaves each 64-sample tile's
 * gradient (and loss in the last slot) in the partials region. */
uint32_t nxg_mlp_gradients(void *state, const float *features, const float *targets,
                           uint64_t sample_count, uint64_t batch_first, uint32_t batch_size,
                           void *stream);

/* nxg_mlp_gradients, then the partials reduced in tile order and one Adam
 * step (beta1 0.9, beta2 0.999, eps 1e-8; step counts from 1). */
uint32_t nxg_mlp_train_step(void *state, const float *features, const float *targets,
                            uint64_t sample_count, uint64_t batch_first, uint32_t batch_size,
                            float learning_rate, uint32_t step, void *stream);

/* outputs[n][4] = attributes predicted for features[n][10]. */
uint32_t nxg_mlp_forward(const void *state, const float *features, float *outputs, uint64_t count,
                         void *stream);

/* Surrogate features of every strand's root (sample_root, generate.h:101-135):
 * position xyz, normal xyz, texcoord uv, texel(face, b1, b2) and
 * texel(face, b2, b1) of `map`; features[n][10]. */
uint32_t nxg_root_features(const nxg_device_asset *asset, const nxg_generation_params *params,
                           const nxg_texel_map *map, float *features, uint64_t capacity,
                           void *stream);

/* Fused root features + MLP: attributes[4*s + k] predicted for strand s.
 * prepared_workspace (optional, with its byte size): a workspace that
 * nxg_prepare_packed_generation filled for the same asset, seed and
 * strand_count; the roots are then processed in its order (grouped by root
 * triangle, so a wavefront's feature gathers share cache lines), which
 * changes no value. */
uint32_t nxg_surrogate_attributes(const nxg_device_asset *asset, const nxg_generation_params *params,
                                  const void *state, const nxg_texel_map *map, float *attributes,
                                  uint64_t attribute_capacity, const void *prepared_workspace,
                                  uint64_t workspace_bytes, void *stream);

/* The surrogate's check against the exact runtime: evaluates the set's
 * programs exactly for every sample_stride-th root (strands s with s % stride == stride / 2); a 64-root
 * tile any of whose samples differs from `attributes` by more than
 * relative_threshold * max(|exact|, 1e-6) in any attribute is re-evaluated
 * exactly for all its roots.  tile_flags: ceil(N/64) words of device
 * scratch; stats (device, optional): [0] tiles re-evaluated, [1] samples
 * checked.  Both are cleared by this call. */
uint32_t nxg_verify_attributes(const nxg_device_asset *asset, const nxg_generation_params *params,
                               const nxg_expr_set *set, const nxg_texel_map *maps, uint32_t map_count,
                               float *attributes, uint64_t attribute_capacity,
                               float relative_threshold, uint32_t sample_stride,
                               uint32_t *tile_flags, uint64_t tile_flag_capacity, uint32_t *stats,
                               void *stream);

/* ---- motion samples ------------------------------------------------------ */

/* DeviceMotionSampleDescriptor (device_contract.h:42-45): one deformation of
 * the asset at one time. */
typedef struct nxg_motion_sample {
    nxg_deformed_geometry deformation;
    float time;
} nxg_motion_sample;

/* DeviceMotionOutputDescriptor (device_contract.h:47-52): sample-major
 * absolute positions, strand_count * cvs_per_strand per sample; capacity in
 * elements. */
typedef struct nxg_motion_output {
    nxg_vec3 *points;
    uint64_t point_capacity;
} nxg_motion_output;

/* validate_device_motion_generation_request (device_contract.h:256-288):
 * samples present, times finite and strictly increasing, each sample's
 * deformation valid for the asset, output large enough. */
uint32_t nxg_validate_motion_generation(const nxg_device_asset *asset, const nxg_motion_sample *samples,
                                        uint32_t sample_count, const nxg_generation_params *params,
                                        const nxg_motion_output *output, const nxg_launch_config *config);

/* Workspace of nxg_generate_motion: sample_count packed-generation workspaces. */
uint64_t nxg_motion_generation_workspace_size(const nxg_asset_header *header, const nxg_generation_params *params,
                                              uint32_t sample_count);

/* The positions of generate_packed under each sample's deformation
 * (generate.h:612-642 with the deformed frame, generate.h:137-187): every
 * sample draws the same roots (triangle and barycentrics); positions are
 * those nxg_generate_packed writes for that deformation.  Two kernels per
 * sample on `stream`; allocates nothing. */
uint32_t nxg_generate_motion(const nxg_device_asset *asset, const nxg_motion_sample *samples,
                             uint32_t sample_count, const nxg_generation_params *params,
                             const nxg_noise_octaves *octaves, const nxg_motion_output *output,
                             const nxg_launch_config *config, void *workspace, uint64_t workspace_bytes,
                             void *stream);

/* ---- batched descriptions ------------------------------------------------
 * Many descriptions generated by one persistent launch.  The reference runs
 * a collection as one chain of dispatches per description
 * (src/luisa/xgen_classic_collection.cpp:108-221, ClassicCollectionPipeline::
 * encode); here every wavefront claims 64-strand tiles of the whole batch
 * from one device counter, so small and large descriptions share the GPU.
 * Each item's results are identical to nxg_generate_packed (or
 * nxg_generate_packed_attributed) called for that item alone. */
#define NXG_BATCH_MAX_ITEMS 12u

typedef struct nxg_batch_item {
    const nxg_device_asset *asset;
    const nxg_deformed_geometry *deformed;  /* optional */
    const nxg_generation_params *params;
    const nxg_noise_octaves *octaves;       /* optional */
    float *attributes;                      /* optional float4 per strand (see exprs) */
    uint64_t attribute_capacity;            /* floats */
    const nxg_expr_set *exprs;              /* optional: attributes are first evaluated by these programs */
    const nxg_texel_map *maps;              /* texel maps of exprs */
    uint32_t map_count;
    uint32_t reserved;
    const nxg_packed_output *output;
    const nxg_stencil_output *stencil;      /* optional */
    void *workspace;                        /* nxg_packed_generation_workspace_size bytes */
    uint64_t workspace_bytes;
    /* optional learned surrogate for an item with exprs (extension, no reference counterpart): its
     * attributes are predicted by this MLP state (nxg_surrogate_attributes, with surrogate_map as the
     * features' texel map, roots in the item's strand order), then checked and, where a sampled root
     * misses surrogate_threshold, re-evaluated by exprs (nxg_verify_attributes with surrogate_stride,
     * surrogate_flags and surrogate_stats) */
    const void *surrogate;
    const nxg_texel_map *surrogate_map;
    float surrogate_threshold;
    uint32_t surrogate_stride;
    uint32_t *surrogate_flags;              /* device scratch, one word per 64-strand tile */
    uint64_t surrogate_flag_capacity;
    uint32_t *surrogate_stats;              /* optional device [2]: tiles re-evaluated, samples checked */
} nxg_batch_item;

#define NXG_BATCH_ORDER_GIVEN 0u     /* descriptions in item order */
#define NXG_BATCH_ORDER_HEAVY_FIRST 1u /* costliest strands first: CVs x noise passes x octaves */

typedef struct nxg_batch_tuning {
    uint32_t waves_per_block;  /* 1, 2 or 4 */
    uint32_t chunk_cvs;        /* CVs staged in LDS per store burst: 4, 8 or 16 (16 needs waves <= 2) */
    uint32_t tiles_per_claim;  /* 64-strand tiles a wavefront takes per counter claim: 1..16 */
    uint32_t blocks_per_cu;    /* persistent grid = blocks_per_cu x CUs; 0 = resident maximum */
    uint32_t order;            /* NXG_BATCH_ORDER_* */
} nxg_batch_tuning;

/* Defaults used when a NULL tuning is passed. */
void nxg_batch_default_tuning(nxg_batch_tuning *tuning);

/* Validates every item as nxg_generate_packed would, plus the tuning. */
uint32_t nxg_batch_validate(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning);

/* On `stream`: the strand order of every item (four launches over all
 * items), the attributes of items that carry exprs (nxg_evaluate_root_
 * attributes, one launch each; or, with a surrogate, the surrogate's
 * prediction in the item's strand order and the exact check and fallback),
 * one guide preparation over all items (which also zeroes *counter, a
 * caller-owned device word), then the persistent strand kernel.  Allocates nothing, never synchronises, records no memset: the
 * call may be captured into a graph.  Calls sharing one counter must not
 * overlap. */
uint32_t nxg_generate_batch(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning,
                            uint32_t *counter, void *stream);

/* nxg_generate_batch captured once into an instantiated graph (host handle
 * in *graph); replay with nxg_batch_graph_launch on any stream; the items'
 * device buffers must outlive the graph.  In the graph the items' attribute
 * work (exact programs; a surrogate's network, in root index order, and its
 * check) runs on three branches beside the order kernels and joins before
 * the strand kernel; the outputs are those of nxg_generate_batch. */
uint32_t nxg_batch_graph_create(const nxg_batch_item *items, uint32_t count, const nxg_batch_tuning *tuning,
                                uint32_t *counter, void **graph);
uint32_t nxg_batch_graph_launch(void *graph, void *stream);
uint32_t nxg_batch_graph_destroy(void *graph);

/* Built-in candidate tunings (returns how many; writes at most capacity). */
uint32_t nxg_batch_tuning_candidates(nxg_batch_tuning *candidates, uint32_t capacity);

/* Autotuner: runs each candidate `repeats` times after one warm-up on
 * `stream`, synchronising, and writes the median milliseconds of each
 * (times_ms[candidate_count]) and the fastest tuning to *best.  Not
 * capturable; the items' outputs hold the last run's results. */
uint32_t nxg_batch_autotune(const nxg_batch_item *items, uint32_t count, uint32_t *counter, void *stream,
                            const nxg_batch_tuning *candidates, uint32_t candidate_count, uint32_t repeats,
                            nxg_batch_tuning *best, float *times_ms);

/* ---- Classic CutFX on generated curves -------------------------------- */

/* The reference's Classic float runtime (apply_xgen_classic_float_runtime_
 * plan_cpu, include/nanoxgen/xgen_classic_runtime.h:185-190) run with one
 * CutFX module whose amount is amount_scale * $cLength + amount_offset
 * (times an optional per-strand multiplier), no length/width/taper
 * expressions, on packed curves such as nxg_generate_packed writes.  Per
 * strand (src/xgen_classic_runtime.cpp:1648-1678, 1806-1821):
 *   cLength = curve_spline_length (2n+4 chords of the uniform cubic
 *             B-spline with reflected ends, src/xgen_classic_runtime.cpp:
 *             587-663); width = 2 * radius[0] / radius_scale (0 when
 *             radius_scale <= 0);
 *   culled when cLength < 1e-4 or width < 1e-4;
 *   cut_from_tip_and_rebuild (src/xgen_classic_runtime.cpp:692-743):
 *             walk back from the tip in 1/(2n+4) parameter steps until
 *             max(amount, 0) of chord length is removed, then resample n
 *             CVs uniformly on [0, cut]; culled when cut < 1e-4;
 * then the kept strands are compacted in order, with their point counts,
 * roots and root uvs (src/xgen_classic_runtime.cpp:1893-1918).  The
 * device counterpart in the reference is make_classic_runtime_cut_kernel
 * (src/luisa/xgen_classic_runtime.cpp:409-468), which flags dropped strands
 * instead of compacting them. */
#define NXG_CLASSIC_MAX_CVS 256u
#define NXG_CLASSIC_MAX_STRANDS (1u << 30)
#define NXG_CLASSIC_NONFINITE_AMOUNT 1u /* "Classic CutFX amount produced a non-finite value" */
#define NXG_CLASSIC_NONFINITE_WIDTH 2u  /* "Classic width expression produced a non-finite value" */

typedef struct nxg_classic_curves {
    const nxg_packed_point *points;  /* strand_count x cvs_per_strand, strand-major, 16-byte aligned */
    const uint32_t *point_counts;    /* optional: copied for kept strands (else cvs_per_strand) */
    const nxg_root_sample *roots;    /* optional */
    const nxg_vec2 *root_uvs;        /* optional */
    uint32_t strand_count;           /* 0..NXG_CLASSIC_MAX_STRANDS */
    uint32_t cvs_per_strand;         /* 2..NXG_CLASSIC_MAX_CVS */
} nxg_classic_curves;

typedef struct nxg_classic_cut_params {
    float amount_scale;          /* amount = (amount_scale * cLength + amount_offset) * multiplier[s] */
    float amount_offset;
    float radius_scale;          /* the radius_scale the curves were written with */
    uint32_t multiplier_stride;  /* floats between consecutive strands' multipliers (4: an attribute column) */
    const float *multiplier;     /* optional, device memory */
} nxg_classic_cut_params;

/* Outputs must not overlap the inputs.  Capacities are in elements and must
 * hold every input strand (the kept count is known only on the device). */
typedef struct nxg_classic_output {
    nxg_packed_point *points;
    uint32_t *point_counts;      /* optional */
    nxg_root_sample *roots;      /* optional; needs curves.roots */
    nxg_vec2 *root_uvs;          /* optional; needs curves.root_uvs */
    uint64_t point_capacity;
    uint64_t point_count_capacity;
    uint64_t root_capacity;
    uint64_t root_uv_capacity;
    uint32_t *kept_count;        /* device word: strands written (0 when status != 0) */
    uint32_t *status;            /* optional device word: 0 or NXG_CLASSIC_NONFINITE_* bits */
} nxg_classic_output;

/* Host-side checks of a request; NXG_OK or the first failing code. */
uint32_t nxg_validate_classic_cut(const nxg_classic_curves *curves, const nxg_classic_cut_params *cut,
                                  const nxg_classic_output *output);

/* Device workspace for nxg_classic_cut_apply: per-strand cut parameters and
 * the per-tile kept counts. */
uint64_t nxg_classic_cut_workspace_size(uint32_t strand_count);

/* On `stream`, without synchronising or allocating: the cut chain above
 * over every strand, then the stable compaction.  Where the reference would
 * throw (a non-finite width or amount on any strand) nothing is written,
 * *kept_count is 0 and *status carries the NXG_CLASSIC_NONFINITE_* bits.
 * Positions follow the reference's CPU operation order bit for bit. */
uint32_t nxg_classic_cut_apply(const nxg_classic_curves *curves, const nxg_classic_cut_params *cut,
                               const nxg_classic_output *output, void *workspace, uint64_t workspace_bytes,
                               void *stream);

/* Renderer endpoints (add_xgen_classic_renderer_endpoints,
 * src/xgen_classic_runtime.cpp:1921-1956): every strand of cvs_per_strand
 * points becomes cvs_per_strand + 2, a phantom point 2 p[0] - p[1] before the
 * root and 2 p[n-1] - p[n-2] after the tip, each with its neighbour's radius;
 * point_counts (optional) receive cvs_per_strand + 2.  With
 * device_strand_count non-NULL the strand count is read on the device (for
 * instance the kept count nxg_classic_cut_apply wrote) and strand_count is
 * only its upper bound.  Output must not overlap input.  On `stream`. */
uint32_t nxg_classic_add_endpoints(const nxg_packed_point *points, uint32_t strand_count, uint32_t cvs_per_strand,
                                   const uint32_t *device_strand_count, nxg_packed_point *out_points,
                                   uint64_t out_point_capacity, uint32_t *out_point_counts,
                                   uint64_t out_point_count_capacity, void *stream);

/* ClumpingFX (one ClassicFloatClumpModule of the reference's float runtime,
 * src/xgen_classic_runtime.cpp:1680-1793) with world-space clump guides
 * (ClassicClumpRuntimeData, include/nanoxgen/xgen_classic_runtime.h:59-88,
 * prepared as prepare_xgen_classic_clump_runtime_data does,
 * src/xgen_classic_runtime.cpp:839-890) and constant module expressions.
 * Per strand whose guide index is not NXG_INVALID_INDEX and whose mask is
 * positive: cut_to_length_and_rebuild to the guide's spline length
 * (src/xgen_classic_runtime.cpp:750-799), blend CVs 1.. toward the guide's
 * render axis by mask * amount * (1 - 2 clump_scale), then add the guide's
 * clump-noise axis (build_clump_noise_axis, src/xgen_classic_runtime.cpp:
 * 1041-1186: two gradient-noise samples along a transported frame, scaled by
 * mask * noise * noise_scale).  Other strands are copied unchanged.  The
 * device counterpart in the reference is make_classic_runtime_clump_kernel
 * (src/luisa/xgen_classic_runtime.cpp:481-764).  The noise frame uses
 * acosf/cosf/sinf, so noisy positions agree with the reference to a
 * tolerance; with noise off they are bit-identical. */
#define NXG_CLUMP_MAX_CVS 64u
#define NXG_CLASSIC_INVALID_GUIDE 4u  /* "Classic ClumpingFX guide index is invalid" */

typedef struct nxg_clump_guides {
    const nxg_vec3 *axes;                /* guide_count x cvs_per_strand CVs (guide_axes), device */
    const nxg_vec3 *normals;             /* guide_count, device */
    const nxg_vec3 *tangents;            /* guide_count, device */
    const nxg_vec3 *reference_positions; /* optional noise-domain points (guide_reference_positions),
                                            else each guide's first CV */
    uint32_t guide_count;
    uint32_t reserved;
} nxg_clump_guides;

typedef struct nxg_clump_params {
    float mask;               /* mask, clamped to [0, 1]; <= 0 leaves strands alone */
    float amount;             /* clump, clamped to [0, 1] (after the multiplier) */
    float clump_scale;        /* clumpScale */
    float noise;              /* noise */
    float noise_scale;        /* noiseScale */
    float noise_frequency;    /* noiseFrequency */
    float noise_correlation;  /* noiseCorrelation, in percent */
    uint32_t amount_stride;   /* floats between consecutive strands' amount multipliers */
    const float *amount_multiplier; /* optional, device: clump = amount * multiplier[s * stride] */
} nxg_clump_params;

/* Device workspace of nxg_classic_clump_apply. */
uint64_t nxg_classic_clump_workspace_size(uint32_t guide_count, uint32_t cvs_per_strand);

/* On `stream`: out_points may equal points (in place) or must not overlap
 * them; strand_guides is device memory (one index per strand); *status
 * (optional device word) receives 0 or NXG_CLASSIC_INVALID_GUIDE / _NONFINITE_AMOUNT
 * bits (such strands are copied unchanged). */
uint32_t nxg_classic_clump_apply(const nxg_packed_point *points, uint32_t strand_count, uint32_t cvs_per_strand,
                                 const uint32_t *strand_guides, const nxg_clump_guides *guides,
                                 const nxg_clump_params *params, nxg_packed_point *out_points,
                                 uint64_t out_point_capacity, uint32_t *status, void *workspace,
                                 uint64_t workspace_bytes, void *stream);

/* ---- Classic primitive expressions --------------------------------------- */

/* apply_xgen_classic_float_runtime_plan_cpu (include/nanoxgen/
 * xgen_classic_runtime.h:185-203, src/xgen_classic_runtime.cpp:1420-1918)
 * for a ClassicFloatRuntimePlan (xgen_classic_runtime.h:130-165) of
 * SplinePrimitive expressions (length, width, taper, taperStart, widthRamp)
 * and an ordered list of FX effects (the reference's plan.effects,
 * xgen_classic_runtime.h:98-106): ClumpingFX, NoiseFX (mode 0) and CutFX
 * (rebuildType 1) modules in any order, up to NXG_CLASSIC_MAX_CLUMPS,
 * _MAX_NOISES and _MAX_CUTS of them.  Per strand, in the reference's order
 * (src/xgen_classic_runtime.cpp:1619-1858):
 *   cLength; length expression -> scale_curve about the root, cLength *=
 *   scale (an error when not finite or negative); raw width from the width
 *   expression, else 2 radius[0] / radius_scale; culled (skipped) when
 *   cLength < 1e-4 or width < 1e-4; cWidth = width;
 * then each effect in list order:
 *   ClumpingFX (:1680-1793) for a strand with a guide: mask (clamped to
 *   [0, 1]; <= 0 leaves the strand alone); noise, noiseFrequency,
 *   noiseCorrelation and noiseScale(t) in the guide's context (its uv and
 *   SeExpr prefix, :1041-1186) build the noise axis, of magnitude
 *   mask noise noiseScale; cut_to_length_and_rebuild to the guide's spline
 *   length (cLength of the rebuilt curve); clump amount (clamped);
 *   CVs 1.. blended toward the guide's render axis by
 *   mask amount (1 - 2 clumpScale(t)), plus the noise axis; cLength again
 *   (with root_relative the goal is taken about the strand's root,
 *   :1627-1628, 1774-1776, and a module with local guide axes clumps
 *   toward their render axis moved to the guide's root);
 *   NoiseFX (apply_noise, :897-1039): mask (> 1e-6), frequency, correlation
 *   and preserveLength at the context's t; the transported frame of the
 *   strand from its surface normal and tangent; per CV magnitude(t) *
 *   magnitudeScale(t) * mask times the classic noise of the domain moved
 *   along each axis; the preserve-length rescale; cLength again;
 *   CutFX: amount, cut_from_tip_and_rebuild, culled when the cut parameter
 *   < 1e-4 (a culled strand still runs the later effects, whose errors
 *   count), cLength of the rebuilt curve;
 * taper, taperStart; per CV t =
 *   cv / (n - 1), scale = 1 - taper (t - start) / (1 - start) past the
 *   start, times widthRamp(t), radius = 0.5 max(cWidth scale, 0)
 *   radius_scale;
 * then the kept strands are compacted in order.  Program inputs bind as the
 * reference's input_value does (src/xgen_classic_runtime.cpp:504-553):
 * NXG_CLASSIC_VAR_ID ($id: the strand's primitive id, else its index),
 * _CLENGTH, _CWIDTH, and NXG_CLASSIC_VAR_COLUMN0 + k for column k of the
 * per-strand input row (the reference's [PTEX, custom, $Prefg noise]
 * runtime_inputs, xgen_classic_runtime.h:193-195).  rand() reads the
 * strand's SeExpr prefix (required when any program calls it), else u, v of
 * its root and the face seed xgen_runtime_face_seed(description_id, name,
 * face) (src/xgen_expression.cpp:724-736). */
#define NXG_CLASSIC_VAR_ID 0u
#define NXG_CLASSIC_VAR_CLENGTH 1u
#define NXG_CLASSIC_VAR_CWIDTH 2u
#define NXG_CLASSIC_VAR_COLUMN0 3u
#define NXG_CLASSIC_MAX_COLUMNS 183u  /* 61 PTEX + 61 custom + 61 $Prefg-noise inputs */
#define NXG_CLASSIC_MAX_PROGRAM_INPUTS 32u
#define NXG_CLASSIC_MAX_PROGRAM_SLOTS 32u
/* status bits beyond NXG_CLASSIC_NONFINITE_AMOUNT / _WIDTH */
#define NXG_CLASSIC_NEGATIVE_LENGTH 8u      /* "Classic length expression produced a negative value" */
#define NXG_CLASSIC_NONFINITE_TAPER 16u     /* "Classic taper expression produced a non-finite value" */
#define NXG_CLASSIC_NONFINITE_PROFILE 32u   /* "Classic width profile produced a non-finite value" */
#define NXG_CLASSIC_EXPRESSION_ERROR 64u    /* "XGen float expression produced a non-finite value"
                                               (or a non-finite input or context) */
#define NXG_CLASSIC_VARIABLE_COUNTS 128u    /* "Classic runtime requires fixed point counts" */
#define NXG_CLASSIC_NONFINITE_CLUMP 256u    /* "Classic ClumpingFX noise magnitude is non-finite" */
#define NXG_CLASSIC_NONFINITE_NOISE_PARAMETER 512u /* "Classic NoiseFX produced a non-finite parameter" */
#define NXG_CLASSIC_NONFINITE_NOISE 1024u   /* "Classic NoiseFX magnitude is non-finite" */

/* Modules of one plan: ClumpingFX modules (each with its own
 * ClassicClumpRuntimeData), NoiseFX modules and CutFX modules, applied in the
 * order of the plan's effect list (the reference's plan.effects,
 * src/xgen_classic_runtime.cpp:1679-1826).  A CutFX that other effects follow
 * rebuilds the strand at once; a strand it culls still runs them. */
#define NXG_CLASSIC_MAX_CLUMPS 3u
#define NXG_CLASSIC_MAX_NOISES 2u
#define NXG_CLASSIC_MAX_CUTS 2u
#define NXG_CLASSIC_MAX_EFFECTS 8u
#define NXG_CLASSIC_CLUMP_EXPRS 7u
#define NXG_CLASSIC_NOISE_EXPRS 6u

/* ClassicFloatEffectType (xgen_classic_runtime.h:98-102), same numbering. */
enum { NXG_CLASSIC_EFFECT_CLUMP = 0, NXG_CLASSIC_EFFECT_NOISE = 1, NXG_CLASSIC_EFFECT_CUT = 2 };

/* ClassicFloatEffect (xgen_classic_runtime.h:104-107): an effect type and
 * the index of its module among the plan's modules of that type. */
typedef struct nxg_classic_effect {
    uint32_t type;
    uint32_t module;
} nxg_classic_effect;

enum {
    NXG_CLASSIC_EXPR_LENGTH = 0,
    NXG_CLASSIC_EXPR_WIDTH,
    NXG_CLASSIC_EXPR_CUT_AMOUNT,
    NXG_CLASSIC_EXPR_TAPER,
    NXG_CLASSIC_EXPR_TAPER_START,
    NXG_CLASSIC_EXPR_WIDTH_RAMP,
    /* the first ClumpingFX module (ClassicFloatClumpModule, xgen_classic_runtime.h:43-56); module m's
     * programs follow at NXG_CLASSIC_EXPR_CLUMP_OF(m, k) */
    NXG_CLASSIC_EXPR_CLUMP_MASK,
    NXG_CLASSIC_EXPR_CLUMP,
    NXG_CLASSIC_EXPR_CLUMP_SCALE,
    NXG_CLASSIC_EXPR_CLUMP_NOISE,
    NXG_CLASSIC_EXPR_CLUMP_NOISE_SCALE,
    NXG_CLASSIC_EXPR_CLUMP_NOISE_FREQUENCY,
    NXG_CLASSIC_EXPR_CLUMP_NOISE_CORRELATION,
    /* the first NoiseFX module (ClassicFloatNoiseModule, xgen_classic_runtime.h:32-40); module n's
     * programs follow at NXG_CLASSIC_EXPR_NOISE_OF(n, k) */
    NXG_CLASSIC_EXPR_NOISE_MASK = NXG_CLASSIC_EXPR_CLUMP_MASK + NXG_CLASSIC_MAX_CLUMPS * NXG_CLASSIC_CLUMP_EXPRS,
    NXG_CLASSIC_EXPR_NOISE_MAGNITUDE,
    NXG_CLASSIC_EXPR_NOISE_MAGNITUDE_SCALE,
    NXG_CLASSIC_EXPR_NOISE_FREQUENCY,
    NXG_CLASSIC_EXPR_NOISE_CORRELATION,
    NXG_CLASSIC_EXPR_NOISE_PRESERVE_LENGTH,
    /* the amounts of CutFX modules 1 .. NXG_CLASSIC_MAX_CUTS - 1 (module 0's is NXG_CLASSIC_EXPR_CUT_AMOUNT) */
    NXG_CLASSIC_EXPR_CUT_AMOUNT_1 = NXG_CLASSIC_EXPR_NOISE_MASK + NXG_CLASSIC_MAX_NOISES * NXG_CLASSIC_NOISE_EXPRS,
    NXG_CLASSIC_EXPR_COUNT = NXG_CLASSIC_EXPR_CUT_AMOUNT_1 + NXG_CLASSIC_MAX_CUTS - 1
};
/* Program k (one of NXG_CLASSIC_EXPR_CLUMP_MASK .. _NOISE_CORRELATION) of ClumpingFX module m. */
#define NXG_CLASSIC_EXPR_CLUMP_OF(m, k) ((k) + NXG_CLASSIC_CLUMP_EXPRS * (m))
/* Program k (one of NXG_CLASSIC_EXPR_NOISE_MASK .. _NOISE_PRESERVE_LENGTH) of NoiseFX module n. */
#define NXG_CLASSIC_EXPR_NOISE_OF(n, k) ((k) + NXG_CLASSIC_NOISE_EXPRS * (n))
/* The amount program of CutFX module c. */
#define NXG_CLASSIC_EXPR_CUT_AMOUNT_OF(c) ((c) == 0u ? (uint32_t)NXG_CLASSIC_EXPR_CUT_AMOUNT \
                                                    : (uint32_t)NXG_CLASSIC_EXPR_CUT_AMOUNT_1 + (c) - 1u)

/* Host memory.  expressions[k] NULL: absent.  Bindings are NXG_CLASSIC_VAR_*;
 * programs hold the reference's ops only (no pow/mix/noise3/texel), at most
 * NXG_CLASSIC_MAX_PROGRAM_INPUTS inputs and NXG_CLASSIC_MAX_PROGRAM_SLOTS live
 * values.  Each clump needs all seven of its expressions; its noiseFrequency
 * and noiseCorrelation may read no input (they are evaluated once per
 * guide).  Each NoiseFX needs all six of its expressions, each CutFX its
 * amount.
 *
 * The effects: with effect_count > 0, effects[0 .. effect_count) in order,
 * naming every module of the plan (clump_count ClumpingFX, noise_count
 * NoiseFX, cut_count CutFX modules) exactly once.  With effect_count 0 the
 * short form applies: clump_count clumps in order, the NoiseFX (has_noise)
 * and then the CutFX (has_cut) after the first clump_count -
 * clumps_after_cut of them (noise_count and cut_count must be 0 then). */
typedef struct nxg_classic_plan {
    const nxg_expr_program *expressions[NXG_CLASSIC_EXPR_COUNT];
    uint32_t has_cut;               /* short form: one CutFX; needs expressions[NXG_CLASSIC_EXPR_CUT_AMOUNT] */
    uint32_t clump_count;           /* ClumpingFX modules (0..NXG_CLASSIC_MAX_CLUMPS) */
    uint32_t description_id;        /* descriptionId */
    uint32_t description_name_seed; /* nxg_string_seed of the description name after its last ':' */
    uint32_t clumps_after_cut;      /* short form: the last this many clumps follow the CutFX */
    uint32_t has_noise;             /* short form: one NoiseFX after the clumps, before the CutFX (needs
                                       clumps_after_cut 0) */
    uint32_t noise_count;           /* NoiseFX modules of the effect list (0..NXG_CLASSIC_MAX_NOISES) */
    uint32_t cut_count;             /* CutFX modules of the effect list (0..NXG_CLASSIC_MAX_CUTS) */
    uint32_t effect_count;          /* 0: the short form */
    nxg_classic_effect effects[NXG_CLASSIC_MAX_EFFECTS];
} nxg_classic_plan;

/* What nxg_classic_plan_pack reports about a plan. */
typedef struct nxg_classic_plan_info {
    uint64_t byte_size;           /* of the device image */
    uint64_t present;             /* bit k: expressions[k] given */
    uint32_t has_cut;
    uint32_t clump_count;
    uint32_t columns_required;    /* 1 + the largest input column read (0: none) */
    uint32_t needs_random_prefix; /* some program calls rand() */
    uint32_t final_length;        /* taper, taperStart or widthRamp reads cLength */
    uint32_t clump_length;        /* bit m: module m's clump amount or clumpScale reads cLength */
    uint32_t description_id;
    uint32_t description_name_seed;
    uint32_t program_offset[NXG_CLASSIC_EXPR_COUNT];
    uint32_t uniform;             /* bit k: expressions[k] reads no input and calls no rand(), so its
                                     value depends on t alone (evaluated once per call) */
    uint32_t clumps_after_cut;    /* the short form's (0 for an effect list) */
    uint32_t has_noise;           /* some NoiseFX module */
    uint32_t noise_count;
    uint32_t cut_count;
    uint32_t effect_count;        /* the effects, short form expanded */
    nxg_classic_effect effects[NXG_CLASSIC_MAX_EFFECTS];
    uint32_t trailing_cut;        /* the last effect is a CutFX (the write resamples the cut strands) */
    uint32_t rewrites;            /* some effect other than a trailing CutFX changes the CVs (they need
                                     cvs_per_strand <= NXG_CLUMP_MAX_CVS and the workspace copy) */
    uint64_t length_readers;      /* bit k: expressions[k] reads $cLength */
} nxg_classic_plan_info;

/* Validates and packs a plan into a position-independent image (out NULL:
 * size only).  NXG_OK, NXG_ERROR_INVALID_PROGRAM, NXG_ERROR_BUFFER_TOO_SMALL
 * or NXG_ERROR_NULL_ARGUMENT. */
uint32_t nxg_classic_plan_pack(const nxg_classic_plan *plan, void *out, uint64_t capacity,
                               nxg_classic_plan_info *info);

/* One ClumpingFX module's geometry binding (ClassicClumpRuntimeData,
 * xgen_classic_runtime.h:59-94); every field is required. */
typedef struct nxg_classic_clump_binding {
    const uint32_t *strand_guides;    /* device: guide per strand, NXG_INVALID_INDEX for none */
    const nxg_clump_guides *guides;   /* host struct of device arrays (world-space guide CVs) */
    const nxg_vec2 *guide_uvs;        /* device: the guides' uv (the noise expressions' context) */
    const uint32_t *guide_random_prefixes; /* device: the guides' SeExpr prefixes */
    const nxg_vec3 *local_axes;       /* optional, device: translation-free copies of the guide CVs
                                         (guide_local_axes, xgen_classic_runtime.h:65-70), guide-major;
                                         they set the guides' spline lengths and, with root_relative,
                                         their render and noise axes */
} nxg_classic_clump_binding;

typedef struct nxg_classic_runtime_inputs {
    const void *plan_image;           /* device copy of the packed plan, 16-byte aligned */
    const nxg_classic_plan_info *plan;/* host: the packer's report for that image */
    const uint32_t *primitive_ids;    /* optional, device: $id per strand */
    const uint32_t *random_prefixes;  /* device; required when plan->needs_random_prefix */
    const float *columns;             /* device, row-major [strand][column_count]; NULL iff column_count 0 */
    uint32_t column_count;            /* >= plan->columns_required */
    float radius_scale;               /* finite, >= 0 */
    const nxg_classic_clump_binding *clumps; /* host array, one per ClumpingFX module (plan->clump_count);
                                                with a clump cvs_per_strand <= NXG_CLUMP_MAX_CVS */
    /* the NoiseFX binding: one surface tangent per strand (required with a NoiseFX, which also needs
     * cvs_per_strand <= NXG_CLUMP_MAX_CVS) and optional per-strand noise domain positions (else the
     * strand's current root) */
    const nxg_vec3 *surface_tangents;
    const nxg_vec3 *noise_domain_positions;
    uint32_t root_relative;           /* the reference's root_relative (xgen_classic_runtime.h:203): clump
                                         goals about each strand's root position */
    uint32_t reserved;
} nxg_classic_runtime_inputs;

/* Host-side checks (the reference's argument checks,
 * src/xgen_classic_runtime.cpp:1432-1512, plus capacities). */
uint32_t nxg_validate_classic_runtime(const nxg_classic_curves *curves, const nxg_classic_runtime_inputs *inputs,
                                      const nxg_classic_output *output);

/* Device workspace for nxg_classic_runtime_apply of a plan (the packer's
 * report; NULL: a plan without ClumpingFX or NoiseFX); guide_count is the
 * sum of the clump modules' guide counts. */
uint64_t nxg_classic_runtime_workspace_size(const nxg_classic_plan_info *plan, uint32_t strand_count,
                                            uint32_t cvs_per_strand, uint32_t guide_count);

/* On `stream`, without synchronising or allocating: the chain above, then
 * the stable compaction into `output` (points, point counts, roots, root
 * uvs).  Where the reference would throw, nothing is written, *kept_count is
 * 0 and *status carries the NXG_CLASSIC_* bits of every failing strand.
 * Positions and radii follow the reference's CPU operation order bit for
 * bit, except where a clump's noise axis or a NoiseFX frame evaluates
 * acos, cos and sin (the device's acosf, cosf and sinf): there positions
 * agree within 1e-5 of the strands' bounding-box diagonal. */
uint32_t nxg_classic_runtime_apply(const nxg_classic_curves *curves, const nxg_classic_runtime_inputs *inputs,
                                   const nxg_classic_output *output, void *workspace, uint64_t workspace_bytes,
                                   void *stream);

/* ---- renderer curve payload (include/nanoxgen/curve_payload.h) -------- */

/* build_curve_batches (src/curve_payload.cpp:150-301) on the GPU: curves
 * split into batches of at most max_strands_per_batch strands, each point
 * transformed by the row-major affine transform (skipped when it is the
 * identity), widths turned into radii (0.5 * width * radius_scale), curves
 * optionally resampled to resample_cv_count CVs in normalised index space
 * (linear between CVs), motion samples transformed (and resampled) alike.
 * The batches are written back to back, so batch b is the strands
 * [b * max, min((b + 1) * max, N)) and its points start at
 * batch_points[b] (batch_points[batch_count] = the total).  Motion points
 * are batch-major and, within a batch, sample-major, as CurveBatch holds them
 * (curve_payload.h:62-72).  Uniform primvars need no copy: batch b's slice
 * of a per-strand array starts at strand b * max. */
#define NXG_PAYLOAD_MAX_MOTION_SAMPLES 16u
#define NXG_PAYLOAD_INVALID_COUNTS 8u  /* a curve with < 2 CVs, or counts not summing to point_count */

typedef struct nxg_curve_set {   /* CurveSetView (curve_payload.h:27-38), device arrays */
    const uint32_t *point_counts; /* optional: every curve then has fixed_cvs_per_strand CVs */
    const nxg_vec3 *points;       /* point_count */
    const float *widths;          /* optional, per point (full diameters); else constant_width */
    const nxg_root_sample *roots; /* optional (root uvs when root_uvs is NULL) */
    const nxg_vec2 *root_uvs;     /* optional */
    uint64_t point_count;         /* < 2^32 */
    uint32_t strand_count;
    uint32_t fixed_cvs_per_strand;
    float constant_width;
    float time;                   /* motion samples: the sample time */
} nxg_curve_set;

typedef struct nxg_curve_payload_options {  /* CurvePayloadOptions (curve_payload.h:78-85) */
    uint32_t max_strands_per_batch;         /* 64K in the reference's defaults */
    uint32_t resample_cv_count;             /* 0: keep each curve's CVs; else >= 2 */
    float radius_scale;
    float transform[12];                    /* row-major 3x4 */
} nxg_curve_payload_options;

typedef struct nxg_curve_payload_output {
    nxg_packed_point *points;      /* all batches' points */
    uint32_t *point_counts;        /* one per strand */
    nxg_vec2 *root_uvs;            /* one per strand, when the base has uvs or roots; else may be NULL */
    nxg_vec3 *motion_points;       /* motion_sample_count x all points */
    uint64_t *batch_points;        /* batch_count + 1 prefix offsets (device) */
    uint32_t *status;              /* optional device word: 0 or NXG_PAYLOAD_INVALID_COUNTS */
    uint64_t point_capacity;
    uint64_t point_count_capacity;
    uint64_t root_uv_capacity;
    uint64_t motion_point_capacity;
    uint64_t batch_capacity;
} nxg_curve_payload_output;

/* Batch count and total output points the request makes; host-side checks of
 * src/curve_payload.cpp:150-196 (NXG_ERROR_INVALID_BUILD_INPUT otherwise;
 * the message in err).  With variable counts and no resampling the total is
 * base->point_count. */
uint32_t nxg_curve_payload_plan(const nxg_curve_set *base, const nxg_curve_set *motion, uint32_t motion_count,
                                const nxg_curve_payload_options *options, uint32_t *batch_count,
                                uint64_t *output_points, char *err, uint64_t err_capacity);

uint64_t nxg_curve_payload_workspace_size(uint32_t strand_count, uint32_t motion_count);

/* On `stream`, without synchronising. */
uint32_t nxg_curve_payload_build(const nxg_curve_set *base, const nxg_curve_set *motion, uint32_t motion_count,
                                 const nxg_curve_payload_options *options, const nxg_curve_payload_output *output,
                                 void *workspace, uint64_t workspace_bytes, void *stream);

/* ---- .nxc curve cache (include/nanoxgen/curve_cache.h) ---------------- */

/* A pointer-free, directly uploadable blob of finished curves: the 128-byte
 * CurveCacheHeader (curve_cache.h:29-47), then the sections in this order,
 * each 16-byte aligned: point counts, packed points, motion times, motion
 * points (sample-major), texcoords, patch uvs, face uvs, face ids; absent
 * sections have offset 0.  content_hash is the word hash of every byte after
 * the header (src/curve_cache.cpp:35-53). */
#define NXG_CURVE_CACHE_MAGIC 0x454843414347584eull /* "NXGCACHE" */
#define NXG_CURVE_CACHE_VERSION_MAJOR 1u
#define NXG_CURVE_CACHE_HAS_TEXCOORDS 1u
#define NXG_CURVE_CACHE_HAS_PATCH_UVS 2u
#define NXG_CURVE_CACHE_HAS_FACE_UVS 4u
#define NXG_CURVE_CACHE_HAS_FACE_IDS 8u
#define NXG_CURVE_CACHE_HAS_MOTION 16u

typedef struct nxg_curve_cache_header {
    uint64_t magic;
    uint16_t version_major;
    uint16_t version_minor;
    uint32_t flags;
    uint64_t byte_size;
    uint64_t content_hash;
    uint32_t strand_count;
    uint32_t point_count;
    uint32_t motion_sample_count;
    uint32_t reserved;
    uint64_t point_counts_offset;
    uint64_t points_offset;
    uint64_t motion_times_offset;
    uint64_t motion_points_offset;
    uint64_t texcoords_offset;
    uint64_t patch_uvs_offset;
    uint64_t face_uvs_offset;
    uint64_t face_ids_offset;
    uint8_t padding[16]; /* the reference header is 64-byte aligned: 128 bytes */
} nxg_curve_cache_header;

/* CurveCacheBuildInput (curve_cache.h:56-67).  Optional arrays may be NULL
 * (counts then ignored); motion_times has motion_sample_count entries and
 * motion_points motion_sample_count * point_count, sample-major. */
typedef struct nxg_curve_cache_input {
    const uint32_t *point_counts;
    const nxg_packed_point *points;
    const nxg_vec2 *texcoords;    /* point_count */
    const nxg_vec2 *patch_uvs;    /* strand_count */
    const nxg_vec2 *face_uvs;     /* strand_count */
    const uint32_t *face_ids;     /* strand_count */
    const float *motion_times;    /* always host memory */
    const nxg_vec3 *motion_points;
    uint32_t strand_count;
    uint32_t point_count;
    uint32_t motion_sample_count;
    uint32_t reserved;
} nxg_curve_cache_input;

/* The header of the cache the input would make (sizes, flags, offsets; no
 * hash), from counts alone: NXG_OK or NXG_ERROR_INVALID_BUILD_INPUT. */
uint32_t nxg_curve_cache_layout(const nxg_curve_cache_input *input, nxg_curve_cache_header *header);

/* build_curve_cache (src/curve_cache.cpp:104-164) from host arrays into a
 * 64-byte aligned caller buffer; *byte_size always receives the size.  The
 * reference's message goes to err on NXG_ERROR_INVALID_BUILD_INPUT. */
uint32_t nxg_curve_cache_build(const nxg_curve_cache_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                               char *err, uint64_t err_capacity);

/* validate_curve_cache (src/curve_cache.cpp:166-239): "" or the reference's
 * message; NXG_OK or NXG_ERROR_INVALID_ASSET. */
uint32_t nxg_curve_cache_validate(const void *blob, uint64_t byte_size, char *msg, uint64_t msg_capacity);

/* The same cache from device arrays (points, counts, uvs, ids and motion
 * points on the device; motion_times on the host): one kernel lays every
 * section out in device_scratch (>= the layout's byte_size, 64-byte aligned),
 * one download fills `out`, and the host adds the content hash and runs the
 * validator, as build_curve_cache does.  Synchronises `stream`. */
uint32_t nxg_curve_cache_build_gpu(const nxg_curve_cache_input *input, void *device_scratch,
                                   uint64_t scratch_capacity, void *out, uint64_t capacity, uint64_t *byte_size,
                                   char *err, uint64_t err_capacity, void *stream);

/* Library build identity: gfx target and compile flags that affect parity. */
const char *nxg_build_info(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* NANOXGEN_AMD_H */
