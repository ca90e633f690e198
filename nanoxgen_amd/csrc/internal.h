// Internal declarations shared by the host translation units of
// libnanoxgen_amd.so.  Not part of the public ABI (include/nanoxgen_amd.h).
#pragma once

#include <cstdint>

#include "../../include/nanoxgen_amd.h"

namespace nxg {

// Shared checks of device_contract.h:118-181 (validate_common_device_input).
uint32_t validate_common(const nxg_device_asset *asset, const nxg_deformed_geometry *deformed,
                         const nxg_generation_params *params, const nxg_launch_config *config);

// device_contract.h:120-135 plus the octave extension.
bool valid_generation_parameters(const nxg_generation_params &p);
bool valid_octaves(const nxg_noise_octaves *octaves);

// Output checks of device_contract.h:232-253.
uint32_t validate_packed_output(const nxg_generation_params &p, const nxg_packed_output *out);

// Workspace layout: per-guide state (2 float4) then the guide offset table
// (guide_count * cvs_per_strand float4), each 256-byte aligned.
struct WorkspaceLayout {
    uint64_t guide_state_offset;
    uint64_t guide_offsets_offset;
    uint64_t total;
};
WorkspaceLayout workspace_layout(uint32_t guide_count, uint32_t cvs_per_strand);

// Asset builder (asset_builder.cpp).  build_blob with host_stencil = false
// leaves the triangle stencil zeroed and the content hash unset; finish_blob
// sets the hash once the stencil is in place.
uint32_t build_blob(const nxg_asset_build_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                    char *err, uint64_t err_capacity, bool host_stencil);
void finish_blob(void *blob);

}  // namespace nxg
