// The .nxc curve cache assembled on the GPU from device-resident curves.
//
//   cache_assemble_kernel   every section of the cache's layout copied into
//                           one device buffer in a single launch: blockIdx.y
//                           picks the section, blocks stride its 16-byte
//                           chunks (4-byte words for the tails), so the whole
//                           cache leaves the device in one contiguous download.
// The content hash (src/curve_cache.cpp:35-53) is a serial chain over 64-bit
// words by definition, so it, the header and the reference's checks run on
// the host copy (curve_cache.cpp).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/nanoxgen_amd.h"
#include "internal.h"

namespace nxg {
namespace {

constexpr uint32_t kCacheSections = 7u;  // motion times come from the host

struct CacheCopy {
    const uint8_t *src[kCacheSections];
    uint64_t dst[kCacheSections];
    uint64_t bytes[kCacheSections];  // multiples of 4; 0: absent
    uint8_t *blob;
};

__global__ void __launch_bounds__(256) cache_assemble_kernel(CacheCopy c) {
    const uint32_t k = blockIdx.y;
    const uint64_t bytes = c.bytes[k];
    if (bytes == 0u) return;
    const uint8_t *src = c.src[k];
    uint8_t *dst = c.blob + c.dst[k];
    const uint64_t stride = uint64_t{gridDim.x} * blockDim.x;
    const uint64_t first = uint64_t{blockIdx.x} * blockDim.x + threadIdx.x;
    if ((reinterpret_cast<uintptr_t>(src) & 15u) == 0u) {  // dst offsets are 16-byte aligned
        const uint64_t chunks = bytes / 16u;
        for (uint64_t i = first; i < chunks; i += stride)
            reinterpret_cast<uint4 *>(dst)[i] = reinterpret_cast<const uint4 *>(src)[i];
        for (uint64_t i = chunks * 4u + first; i < bytes / 4u; i += stride)
            reinterpret_cast<uint32_t *>(dst)[i] = reinterpret_cast<const uint32_t *>(src)[i];
    } else {
        for (uint64_t i = first; i < bytes / 4u; i += stride)
            reinterpret_cast<uint32_t *>(dst)[i] = reinterpret_cast<const uint32_t *>(src)[i];
    }
}

}  // namespace
}  // namespace nxg

extern "C" uint32_t nxg_curve_cache_build_gpu(const nxg_curve_cache_input *input, void *device_scratch,
                                              uint64_t scratch_capacity, void *out, uint64_t capacity,
                                              uint64_t *byte_size, char *err, uint64_t err_capacity,
                                              void *stream_ptr) {
    using namespace nxg;
    if (!input) return NXG_ERROR_NULL_ARGUMENT;
    nxg_curve_cache_header h;
    if (uint32_t s = curve_cache_layout(*input, h, err, err_capacity); s != NXG_OK) return s;
    if (byte_size) *byte_size = h.byte_size;
    if (!out || !device_scratch) return NXG_ERROR_NULL_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(out) % 64u != 0u || reinterpret_cast<uintptr_t>(device_scratch) % 64u != 0u)
        return NXG_ERROR_MISALIGNED_ASSET;
    if (capacity < h.byte_size || scratch_capacity < h.byte_size) return NXG_ERROR_BUFFER_TOO_SMALL;
    const bool motion_present = input->motion_times && input->motion_points;
    if (h.motion_sample_count != 0u && !motion_present) {
        // the reference checks counts first; they are on the device here
        if (err && err_capacity) std::snprintf(err, err_capacity, "%s",
                                               "curve cache motion times and positions must both be present");
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }
    const uint64_t n = h.strand_count, p = h.point_count, m = h.motion_sample_count;
    CacheCopy c{};
    c.blob = static_cast<uint8_t *>(device_scratch);
    const void *srcs[kCacheSections] = {input->point_counts, input->points, input->motion_points, input->texcoords,
                                       input->patch_uvs, input->face_uvs, input->face_ids};
    const uint64_t offs[kCacheSections] = {h.point_counts_offset, h.points_offset, h.motion_points_offset,
                                           h.texcoords_offset, h.patch_uvs_offset, h.face_uvs_offset,
                                           h.face_ids_offset};
    const uint64_t sizes[kCacheSections] = {n * 4u, p * sizeof(nxg_packed_point), m * p * sizeof(nxg_vec3),
                                            p * sizeof(nxg_vec2), n * sizeof(nxg_vec2), n * sizeof(nxg_vec2), n * 4u};
    for (uint32_t k = 0; k < kCacheSections; ++k) {
        const bool present = offs[k] != 0u && srcs[k];
        if (present && reinterpret_cast<uintptr_t>(srcs[k]) % 4u != 0u) return NXG_ERROR_MISALIGNED_ASSET;
        c.src[k] = static_cast<const uint8_t *>(srcs[k]);
        c.dst[k] = offs[k];
        c.bytes[k] = present ? sizes[k] : 0u;
    }
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    hipLaunchKernelGGL(cache_assemble_kernel, dim3(1024, kCacheSections), dim3(256), 0, stream, c);
    if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    uint8_t *blob = static_cast<uint8_t *>(out);
    if (hipMemcpyAsync(blob, device_scratch, h.byte_size, hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return NXG_ERROR_HIP;
    // the header, the padding between sections and the host-side motion times
    std::memset(blob, 0, sizeof(h));
    uint64_t ends[kCacheSections + 1][2];
    uint32_t count = 0u;
    for (uint32_t k = 0; k < kCacheSections; ++k)
        if (c.bytes[k]) ends[count][0] = c.dst[k], ends[count++][1] = c.dst[k] + c.bytes[k];
    if (m) ends[count][0] = h.motion_times_offset, ends[count++][1] = h.motion_times_offset + m * 4u;
    for (uint32_t a = 0; a < count; ++a) {  // zero from each section's end to the next section's start
        uint64_t next = h.byte_size;
        for (uint32_t b = 0; b < count; ++b)
            if (ends[b][0] >= ends[a][1] && ends[b][0] < next) next = ends[b][0];
        std::memset(blob + ends[a][1], 0, next - ends[a][1]);
    }
    std::memset(blob + sizeof(h), 0, (count ? h.point_counts_offset : h.byte_size) - sizeof(h));
    if (m) std::memcpy(blob + h.motion_times_offset, input->motion_times, m * 4u);
    if (uint32_t s = curve_cache_check_data(h, reinterpret_cast<const uint32_t *>(blob + h.point_counts_offset),
                                            motion_present, input->motion_times, err, err_capacity);
        s != NXG_OK)
        return s;
    return curve_cache_finish(blob, h, err, err_capacity);
}
