// Asset build with the triangle guide stencil computed on the GPU.
//
// The stencil (src/asset.cpp:252-269) is the O(triangles x guides) part of
// build_asset: for every triangle, the k guides with the smallest key
//   |centroid - guide root|^2 + (1e20 if dot(face normal, guide normal) < -0.25)
// in ascending key order, ties to the lower guide index.  Everything else
// (normals, the FP64 LIFO alias table, guide records, the FNV-1a content
// hash) stays on the host, where its operation order is sequential.
//
//   stencil_kernel   one lane per triangle; guide roots and normals staged in
//                    LDS tiles that every lane reads in the same order
//                    (broadcast reads); each lane keeps its sorted top-k in
//                    LDS and inserts exactly as the host builder does, so the
//                    rows are identical to it bit for bit.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "../../include/nanoxgen_amd.h"
#include "internal.h"

namespace nxg {
namespace {

constexpr uint32_t kGuideTile = 512;
constexpr uint32_t kBlock = 256;

struct P3 {
    float x, y, z;
};
__device__ __forceinline__ P3 ld3(const nxg_vec3 &v) { return {v.x, v.y, v.z}; }
__device__ __forceinline__ P3 operator+(P3 a, P3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ P3 operator-(P3 a, P3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ P3 operator/(P3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
__device__ __forceinline__ float dot3(P3 a, P3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ P3 cross3(P3 a, P3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ P3 normalize3(P3 v) {  // generate.h:67-70
    const float l2 = dot3(v, v);
    return l2 > 1.0e-20f ? v / sqrtf(l2) : P3{0.0f, 1.0f, 0.0f};
}

__global__ void __launch_bounds__(kBlock) stencil_kernel(const uint8_t *blob, nxg_asset_header h, uint32_t k,
                                                         uint32_t *stencil) {
    __shared__ float4 groot[kGuideTile];
    __shared__ float4 gnorm[kGuideTile];
    __shared__ float keys[NXG_GUIDE_STENCIL_SIZE][kBlock];
    __shared__ uint32_t ids[NXG_GUIDE_STENCIL_SIZE][kBlock];
    const nxg_vec3 *positions = reinterpret_cast<const nxg_vec3 *>(blob + h.positions_offset);
    const nxg_uint3 *triangles = reinterpret_cast<const nxg_uint3 *>(blob + h.triangles_offset);
    const nxg_guide_record *guides = reinterpret_cast<const nxg_guide_record *>(blob + h.guides_offset);
    const uint32_t lane = threadIdx.x;
    const uint32_t ti = blockIdx.x * kBlock + lane;
    const bool live = ti < h.triangle_count;
    P3 centroid{0.0f, 0.0f, 0.0f}, fn{0.0f, 1.0f, 0.0f};
    if (live) {
        const nxg_uint3 t = triangles[ti];
        const P3 p0 = ld3(positions[t.x]), p1 = ld3(positions[t.y]), p2 = ld3(positions[t.z]);
        centroid = (p0 + p1 + p2) / 3.0f;
        fn = normalize3(cross3(p1 - p0, p2 - p0));
    }
    const uint32_t G = h.guide_count;
    const uint32_t count = k < G ? k : G;
    uint32_t filled = 0u;
    float kth = 0.0f;  // keys[count - 1] once filled == count
    for (uint32_t base = 0; base < G; base += kGuideTile) {
        __syncthreads();
        const uint32_t n = G - base < kGuideTile ? G - base : kGuideTile;
        for (uint32_t i = lane; i < n; i += kBlock) {
            const nxg_guide_record g = guides[base + i];
            groot[i] = make_float4(g.root_position.x, g.root_position.y, g.root_position.z, 0.0f);
            gnorm[i] = make_float4(g.root_normal.x, g.root_normal.y, g.root_normal.z, 0.0f);
        }
        __syncthreads();
        if (!live) continue;
        for (uint32_t j = 0; j < n; ++j) {
            const float4 gr = groot[j], gn = gnorm[j];
            const float penalty = dot3(fn, P3{gn.x, gn.y, gn.z}) < -0.25f ? 1.0e20f : 0.0f;
            const P3 d = centroid - P3{gr.x, gr.y, gr.z};
            const float key = dot3(d, d) + penalty;
            if (filled == count && !(key < kth)) continue;
            uint32_t pos = filled < count ? filled++ : count - 1u;
            while (pos > 0u && key < keys[pos - 1u][lane]) {
                keys[pos][lane] = keys[pos - 1u][lane];
                ids[pos][lane] = ids[pos - 1u][lane];
                --pos;
            }
            keys[pos][lane] = key;
            ids[pos][lane] = base + j;
            if (filled == count) kth = keys[count - 1u][lane];
        }
    }
    if (!live) return;
    uint32_t *row = stencil + uint64_t{ti} * NXG_GUIDE_STENCIL_SIZE;
#pragma unroll
    for (uint32_t j = 0; j < NXG_GUIDE_STENCIL_SIZE; ++j) row[j] = j < filled ? ids[j][lane] : NXG_INVALID_INDEX;
}

}  // namespace
}  // namespace nxg

extern "C" {

uint32_t nxg_asset_build_gpu(const nxg_asset_build_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                             char *err, uint64_t err_capacity, void *device_out, uint64_t device_capacity,
                             void *stream_ptr) {
    uint64_t size = 0u;
    if (uint32_t s = nxg::build_blob(input, out, capacity, &size, err, err_capacity, false); s != NXG_OK) {
        if (byte_size) *byte_size = size;
        return s;
    }
    if (byte_size) *byte_size = size;
    if (!device_out) return NXG_ERROR_NULL_ASSET;
    if (reinterpret_cast<uintptr_t>(device_out) % NXG_BLOB_ALIGNMENT != 0u) return NXG_ERROR_MISALIGNED_ASSET;
    if (device_capacity < size) return NXG_ERROR_ASSET_CAPACITY_TOO_SMALL;
    nxg_asset_header h;
    std::memcpy(&h, out, sizeof(h));
    const uint32_t k = input->stencil_k == 0u ? NXG_GUIDE_STENCIL_SIZE : input->stencil_k;
    hipStream_t stream = static_cast<hipStream_t>(stream_ptr);
    uint8_t *dev = static_cast<uint8_t *>(device_out);
    uint8_t *host = static_cast<uint8_t *>(out);
    const uint64_t stencil_bytes = uint64_t{h.triangle_count} * NXG_GUIDE_STENCIL_SIZE * 4u;
    if (hipMemcpyAsync(dev, host, size, hipMemcpyHostToDevice, stream) != hipSuccess) return NXG_ERROR_HIP;
    hipLaunchKernelGGL(nxg::stencil_kernel, dim3((h.triangle_count + nxg::kBlock - 1u) / nxg::kBlock),
                       dim3(nxg::kBlock), 0, stream, dev, h, k,
                       reinterpret_cast<uint32_t *>(dev + h.triangle_guides_offset));
    if (hipGetLastError() != hipSuccess) return NXG_ERROR_HIP;
    if (hipMemcpyAsync(host + h.triangle_guides_offset, dev + h.triangle_guides_offset, stencil_bytes,
                       hipMemcpyDeviceToHost, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return NXG_ERROR_HIP;
    nxg::finish_blob(out);
    if (hipMemcpyAsync(dev, host, sizeof(nxg_asset_header), hipMemcpyHostToDevice, stream) != hipSuccess ||
        hipStreamSynchronize(stream) != hipSuccess)
        return NXG_ERROR_HIP;
    return NXG_OK;
}

}  // extern "C"
