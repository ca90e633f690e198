// Internal declarations shared by the host translation units of
// libnanoxgen_amd.so.  Not part of the public ABI (include/nanoxgen_amd.h).
#pragma once

#include <cstddef>
#include <cstdint>

#include "../../include/nanoxgen_amd.h"

namespace nxg {

// The asset view's checks of validate_common (pointer, alignment, header
// metadata, capacity).
uint32_t validate_asset_view(const nxg_device_asset *asset);
// Shared checks of device_contract.h:118-181 (validate_common_device_input).
uint32_t validate_common(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                         const nxg_generation_params *params, const nxg_launch_config *config);

// device_contract.h:120-135 plus the octave extension.
bool valid_generation_parameters(const nxg_generation_params &p);
bool valid_octaves(const nxg_noise_octaves *octaves);

// Output checks of device_contract.h:232-253.
uint32_t validate_packed_output(const nxg_generation_params &p, const nxg_packed_output *out);

// Workspace layout, each region 256-byte aligned:
//   guide state       guide_count x 2 float4 (root + support radius, normal)
//   guide offsets     guide_count x cvs_per_strand float4
//   strand order      strand_count u32: strands grouped by root triangle
//   bin counts        kOrderBins x order tiles u32: per-tile counting-sort
//                     histograms, tile-major, scanned in place in bin-major
//                     order into starts
//   scan partials     one u32 per kScanBlock counts
//   strand keys       strand_count u32: each strand's bin and its rank among
//                     its tile's strands of that bin (written by the count,
//                     read by the scatter)
struct WorkspaceLayout {
    uint64_t guide_state_offset;
    uint64_t guide_offsets_offset;
    uint64_t order_offset;
    uint64_t counts_offset;
    uint64_t partials_offset;
    uint64_t keys_offset;
    uint64_t total;
};
constexpr uint32_t kScanBlock = 1024u;   // values per scan block
constexpr uint32_t kOrderBins = 8192u;   // counting-sort bins (LDS histogram per tile)
constexpr uint32_t kOrderTile = 8192u;   // strands per counting-sort tile
// Triangles per bin: the smallest power of two that fits the triangles in
// kOrderBins bins (bins hold consecutive triangle indices).
inline uint32_t order_shift(uint32_t triangle_count) {
    uint32_t shift = 0u;
    while ((uint64_t{triangle_count} + (uint64_t{1} << shift) - 1u) >> shift > kOrderBins) ++shift;
    return shift;
}
inline uint32_t order_tiles(uint32_t strand_count) { return (strand_count + kOrderTile - 1u) / kOrderTile; }
WorkspaceLayout workspace_layout(uint32_t guide_count, uint32_t cvs_per_strand, uint32_t strand_count,
                                 uint32_t triangle_count);

// Classic cut workspace, each region 256-byte aligned:
//   cut parameters   strand_count float: the rebuild parameter of a kept
//                    strand, -1 for a culled one
//   tile counts      one u32 per 64-strand tile: kept strands (bits 0..29)
//                    and the NXG_CLASSIC_NONFINITE_* bits seen (30..31),
//                    scanned in place into the tile's first output slot
//                    within its group of kScanBlock tiles
//   group totals     one u32 per kScanBlock tiles, scanned likewise
//   status           one u32: the error bits of the whole call
//   sample tables    the spline samples every strand of the call shares
//                    (cLength's 2n+3 and the cut walk's), 32 B each
struct ClassicLayout {
    uint64_t cut_offset;
    uint64_t tiles_offset;
    uint64_t groups_offset;
    uint64_t status_offset;
    uint64_t table_offset;
    uint64_t total;
};
constexpr uint32_t kClassicTile = 64u;
constexpr uint32_t kClassicTableEntries = 2u * NXG_CLASSIC_MAX_CVS + 8u;  // per table
constexpr uint64_t kClassicTableBytes = 2u * kClassicTableEntries * 32u;
inline uint32_t classic_tiles(uint32_t strand_count) { return (strand_count + kClassicTile - 1u) / kClassicTile; }
inline uint32_t classic_groups(uint32_t strand_count) { return (classic_tiles(strand_count) + kScanBlock - 1u) / kScanBlock; }
inline ClassicLayout classic_layout(uint32_t strand_count) {
    auto a = [](uint64_t v) { return (v + 255u) & ~uint64_t{255u}; };
    ClassicLayout l{};
    l.cut_offset = 0u;
    l.tiles_offset = a(uint64_t{strand_count} * 4u);
    l.groups_offset = l.tiles_offset + a(uint64_t{classic_tiles(strand_count)} * 4u);
    l.status_offset = l.groups_offset + a(uint64_t{classic_groups(strand_count)} * 4u);
    l.table_offset = l.status_offset + 256u;
    l.total = l.table_offset + a(kClassicTableBytes);
    return l;
}

// Classic runtime workspace, after the cut layout (each region 256-byte
// aligned):
//   uniform values  {value, ok} pairs of the plan's programs of t alone:
//                   [k] program k at t = 0, then widthRamp, clumpScale and
//                   noiseScale at every CV (kClassicCvTables tables)
//   strand state    32 B per strand: the search's values for the write
//   clump only:     the clumped strands (strand_count x cvs float4), and per
//                   guide its render axis and unit noise axis (cvs float4
//                   each) and {spline length, noise frequency-correlation ok}
constexpr uint32_t kClassicScalarSlots = 48u;  // every program at t = 0, three spares at t = 1
constexpr uint32_t kClassicCvTables = 1u + 2u * NXG_CLASSIC_MAX_CLUMPS;  // widthRamp; clumpScale, noiseScale per module
constexpr uint64_t kClassicUniformBytes = (kClassicScalarSlots + kClassicCvTables * NXG_CLASSIC_MAX_CVS) * 8u;
struct ClassicRuntimeLayout {
    uint64_t uniform_offset, state_offset, clumped_offset, render_offset, noise_offset, guide_offset, total;
};
inline ClassicRuntimeLayout classic_runtime_layout(uint32_t strand_count, uint32_t cvs, uint32_t guides,
                                                   bool rewrites = false) {
    auto a = [](uint64_t v) { return (v + 255u) & ~uint64_t{255u}; };
    ClassicRuntimeLayout l{};
    l.uniform_offset = classic_layout(strand_count).total;
    l.state_offset = l.uniform_offset + a(kClassicUniformBytes);
    l.clumped_offset = l.state_offset + a(uint64_t{strand_count} * 32u);
    // the strands a clump or a NoiseFX rewrites
    const uint64_t clumped = guides || rewrites ? uint64_t{strand_count} * cvs * 16u : 0u;
    l.render_offset = l.clumped_offset + a(clumped);
    l.noise_offset = l.render_offset + a(uint64_t{guides} * cvs * 16u);
    l.guide_offset = l.noise_offset + a(uint64_t{guides} * cvs * 16u);
    l.total = l.guide_offset + a(uint64_t{guides} * 8u);
    return l;
}

// SeExpr hashing and XGen string seeds in double, as the reference computes
// them (expr_compile.cpp).
uint32_t seexpr_component(double x);
uint32_t seexpr_finish(uint32_t state);
double seexpr_hash(const double *args, std::size_t n);
uint32_t string_seed(const char *text, std::size_t size);

// Curve cache (curve_cache.cpp): layout from counts, the builder's data
// checks, the content hash, and the finishing step (hash, header, validation).
uint32_t curve_cache_layout(const nxg_curve_cache_input &in, nxg_curve_cache_header &h, char *err, uint64_t cap);
uint32_t curve_cache_check_data(const nxg_curve_cache_header &h, const uint32_t *counts, bool motion_present,
                                const float *motion_times, char *err, uint64_t cap);
uint64_t curve_cache_hash(const uint8_t *bytes, uint64_t size);
uint32_t curve_cache_finish(uint8_t *blob, nxg_curve_cache_header h, char *err, uint64_t cap);

// Asset builder (asset_builder.cpp).  build_blob with host_stencil = false
// leaves the triangle stencil zeroed and the content hash unset; finish_blob
// sets the hash once the stencil is in place.
uint32_t build_blob(const nxg_asset_build_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                    char *err, uint64_t err_capacity, bool host_stencil);
void finish_blob(void *blob);

}  // namespace nxg
