// The .nxc curve cache (include/nanoxgen/curve_cache.h of the reference):
// layout, builder, content hash and validator in C-ABI terms, with the
// reference's checks in its order and its messages
// (src/curve_cache.cpp:12-239).  The GPU assembly path (cache_gpu.hip) shares
// the layout and the host finishing step defined here.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "internal.h"

static_assert(sizeof(nxg_curve_cache_header) == 128, "CurveCacheHeader is 64-byte aligned: 128 bytes");

namespace nxg {
namespace {

constexpr uint64_t kSection = 16u;  // section alignment (src/curve_cache.cpp:12)
constexpr uint32_t kKnownFlags = NXG_CURVE_CACHE_HAS_TEXCOORDS | NXG_CURVE_CACHE_HAS_PATCH_UVS |
                                 NXG_CURVE_CACHE_HAS_FACE_UVS | NXG_CURVE_CACHE_HAS_FACE_IDS |
                                 NXG_CURVE_CACHE_HAS_MOTION;

uint64_t align_up(uint64_t v, uint64_t a) { return (v + a - 1u) & ~(a - 1u); }

void put(char *msg, uint64_t cap, const char *text) {
    if (msg && cap) std::snprintf(msg, cap, "%s", text);
}

bool finite2(const nxg_vec2 &v) { return std::isfinite(v.x) && std::isfinite(v.y); }

template <class T>
bool section_fits(const nxg_curve_cache_header &h, uint64_t offset, uint64_t count) {
    if (count == 0u) return offset == 0u;
    if (offset < sizeof(nxg_curve_cache_header) || offset > h.byte_size) return false;
    if (count > std::numeric_limits<uint64_t>::max() / sizeof(T)) return false;
    const uint64_t size = count * sizeof(T);
    return size <= h.byte_size - offset && offset % alignof(T) == 0u;
}

}  // namespace

// content_hash (src/curve_cache.cpp:35-53): per 64-bit word, then the tail.
uint64_t curve_cache_hash(const uint8_t *bytes, uint64_t size) {
    constexpr uint64_t prime = 0x9e3779b185ebca87ull;
    uint64_t hash = 0x243f6a8885a308d3ull ^ size;
    uint64_t offset = 0u;
    for (; offset + 8u <= size; offset += 8u) {
        uint64_t word;
        std::memcpy(&word, bytes + offset, 8u);
        hash ^= word;
        hash *= prime;
        hash ^= hash >> 29u;
    }
    uint64_t tail = 0u;
    if (size - offset) std::memcpy(&tail, bytes + offset, size - offset);
    hash ^= tail + prime;
    hash *= 0xc2b2ae3d27d4eb4full;
    hash ^= hash >> 32u;
    return hash;
}

// Sizes, flags and offsets of build_curve_cache's append order
// (src/curve_cache.cpp:138-156): counts, points, motion times, motion
// points, texcoords, patch uvs, face uvs, face ids.
uint32_t curve_cache_layout(const nxg_curve_cache_input &in, nxg_curve_cache_header &h, char *err, uint64_t cap) {
    h = nxg_curve_cache_header{};
    if (in.strand_count == 0u || in.point_count == 0u || !in.point_counts || !in.points) {
        put(err, cap, "curve cache needs curves and points");
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }
    h.magic = NXG_CURVE_CACHE_MAGIC;
    h.version_major = NXG_CURVE_CACHE_VERSION_MAJOR;
    h.strand_count = in.strand_count;
    h.point_count = in.point_count;
    h.motion_sample_count = in.motion_sample_count;
    if (in.texcoords) h.flags |= NXG_CURVE_CACHE_HAS_TEXCOORDS;
    if (in.patch_uvs) h.flags |= NXG_CURVE_CACHE_HAS_PATCH_UVS;
    if (in.face_uvs) h.flags |= NXG_CURVE_CACHE_HAS_FACE_UVS;
    if (in.face_ids) h.flags |= NXG_CURVE_CACHE_HAS_FACE_IDS;
    if (in.motion_sample_count) h.flags |= NXG_CURVE_CACHE_HAS_MOTION;
    uint64_t end = sizeof(nxg_curve_cache_header);
    auto append = [&](bool present, uint64_t bytes) -> uint64_t {
        if (!present || bytes == 0u) return 0u;
        const uint64_t offset = align_up(end, kSection);
        end = offset + bytes;
        return offset;
    };
    const uint64_t n = in.strand_count, p = in.point_count, m = in.motion_sample_count;
    h.point_counts_offset = append(true, n * 4u);
    h.points_offset = append(true, p * sizeof(nxg_packed_point));
    h.motion_times_offset = append(m != 0u, m * 4u);
    h.motion_points_offset = append(m != 0u, m * p * sizeof(nxg_vec3));
    h.texcoords_offset = append(in.texcoords, p * sizeof(nxg_vec2));
    h.patch_uvs_offset = append(in.patch_uvs, n * sizeof(nxg_vec2));
    h.face_uvs_offset = append(in.face_uvs, n * sizeof(nxg_vec2));
    h.face_ids_offset = append(in.face_ids, n * 4u);
    h.byte_size = end;
    return NXG_OK;
}

// The builder's own checks of the data, in its order (src/curve_cache.cpp:
// 112-136): counts, their sum, motion presence, motion times.
uint32_t curve_cache_check_data(const nxg_curve_cache_header &h, const uint32_t *counts, bool motion_present,
                                const float *motion_times, char *err, uint64_t cap) {
    uint64_t sum = 0u;
    for (uint32_t s = 0; s < h.strand_count; ++s) {
        if (counts[s] < 2u) {
            put(err, cap, "curve cache curve has fewer than two CVs");
            return NXG_ERROR_INVALID_BUILD_INPUT;
        }
        sum += counts[s];
    }
    if (sum != h.point_count) {
        put(err, cap, "curve cache point counts do not match point array");
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }
    if (h.motion_sample_count != 0u && !motion_present) {
        put(err, cap, "curve cache motion times and positions must both be present");
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }
    float previous = -std::numeric_limits<float>::infinity();
    for (uint32_t k = 0; k < h.motion_sample_count; ++k) {
        if (!std::isfinite(motion_times[k]) || motion_times[k] <= previous) {
            put(err, cap, "curve cache motion times must be finite and increasing");
            return NXG_ERROR_INVALID_BUILD_INPUT;
        }
        previous = motion_times[k];
    }
    return NXG_OK;
}

// Hash and header into a laid-out blob, then the reference's post-build
// validation (src/curve_cache.cpp:157-163).
uint32_t curve_cache_finish(uint8_t *blob, nxg_curve_cache_header h, char *err, uint64_t cap) {
    std::memcpy(blob, &h, sizeof(h));
    h.content_hash = curve_cache_hash(blob + sizeof(h), h.byte_size - sizeof(h));
    std::memcpy(blob, &h, sizeof(h));
    char msg[256] = "";
    if (nxg_curve_cache_validate(blob, h.byte_size, msg, sizeof(msg)) != NXG_OK) {
        char full[320];
        std::snprintf(full, sizeof(full), "built an invalid curve cache: %s", msg);
        put(err, cap, full);
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }
    return NXG_OK;
}

}  // namespace nxg

extern "C" {

uint32_t nxg_curve_cache_layout(const nxg_curve_cache_input *input, nxg_curve_cache_header *header) {
    if (!input || !header) return NXG_ERROR_NULL_ARGUMENT;
    return nxg::curve_cache_layout(*input, *header, nullptr, 0u);
}

uint32_t nxg_curve_cache_build(const nxg_curve_cache_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                               char *err, uint64_t err_capacity) {
    if (!input) return NXG_ERROR_NULL_ARGUMENT;
    nxg_curve_cache_header h;
    if (uint32_t s = nxg::curve_cache_layout(*input, h, err, err_capacity); s != NXG_OK) return s;
    if (byte_size) *byte_size = h.byte_size;
    if (!out) return NXG_ERROR_NULL_ARGUMENT;
    if (reinterpret_cast<uintptr_t>(out) % 64u != 0u) return NXG_ERROR_MISALIGNED_ASSET;
    if (capacity < h.byte_size) return NXG_ERROR_BUFFER_TOO_SMALL;
    if (uint32_t s = nxg::curve_cache_check_data(h, input->point_counts,
                                                 input->motion_times && input->motion_points,
                                                 input->motion_times, err, err_capacity);
        s != NXG_OK)
        return s;
    uint8_t *blob = static_cast<uint8_t *>(out);
    std::memset(blob, 0, h.byte_size);
    const uint64_t n = input->strand_count, p = input->point_count, m = input->motion_sample_count;
    auto copy = [&](uint64_t offset, const void *src, uint64_t bytes) {
        if (offset) std::memcpy(blob + offset, src, bytes);
    };
    copy(h.point_counts_offset, input->point_counts, n * 4u);
    copy(h.points_offset, input->points, p * sizeof(nxg_packed_point));
    copy(h.motion_times_offset, input->motion_times, m * 4u);
    copy(h.motion_points_offset, input->motion_points, m * p * sizeof(nxg_vec3));
    copy(h.texcoords_offset, input->texcoords, p * sizeof(nxg_vec2));
    copy(h.patch_uvs_offset, input->patch_uvs, n * sizeof(nxg_vec2));
    copy(h.face_uvs_offset, input->face_uvs, n * sizeof(nxg_vec2));
    copy(h.face_ids_offset, input->face_ids, n * 4u);
    return nxg::curve_cache_finish(blob, h, err, err_capacity);
}

uint32_t nxg_curve_cache_validate(const void *blob_ptr, uint64_t size, char *msg, uint64_t cap) {
    using nxg::put;
    auto fail = [&](const char *text) {
        put(msg, cap, text);
        return static_cast<uint32_t>(NXG_ERROR_INVALID_ASSET);
    };
    put(msg, cap, "");
    const uint8_t *bytes = static_cast<const uint8_t *>(blob_ptr);
    if (!bytes || size < sizeof(nxg_curve_cache_header)) return fail("blob is smaller than CurveCacheHeader");
    if (reinterpret_cast<uintptr_t>(bytes) % 64u != 0u) return fail("curve-cache base is not 64-byte aligned");
    nxg_curve_cache_header h;
    std::memcpy(&h, bytes, sizeof(h));
    if (h.magic != NXG_CURVE_CACHE_MAGIC) return fail("bad curve-cache magic");
    if (h.version_major != NXG_CURVE_CACHE_VERSION_MAJOR) return fail("unsupported curve-cache major version");
    if ((h.flags & ~nxg::kKnownFlags) != 0u) return fail("curve cache has unknown flags");
    if (h.byte_size != size) return fail("curve-cache byte_size mismatch");
    if (h.strand_count == 0u || h.point_count == 0u) return fail("curve cache is empty");
    if (((h.flags & NXG_CURVE_CACHE_HAS_MOTION) != 0u) != (h.motion_sample_count != 0u))
        return fail("curve-cache motion flag/count mismatch");
    const uint64_t motion_points = uint64_t{h.motion_sample_count} * h.point_count;
    auto optional = [&](uint32_t flag, uint64_t count) { return (h.flags & flag) ? count : 0u; };
    using nxg::section_fits;
    if (!section_fits<uint32_t>(h, h.point_counts_offset, h.strand_count) ||
        !section_fits<nxg_packed_point>(h, h.points_offset, h.point_count) ||
        !section_fits<float>(h, h.motion_times_offset, h.motion_sample_count) ||
        !section_fits<nxg_vec3>(h, h.motion_points_offset, motion_points) ||
        !section_fits<nxg_vec2>(h, h.texcoords_offset, optional(NXG_CURVE_CACHE_HAS_TEXCOORDS, h.point_count)) ||
        !section_fits<nxg_vec2>(h, h.patch_uvs_offset, optional(NXG_CURVE_CACHE_HAS_PATCH_UVS, h.strand_count)) ||
        !section_fits<nxg_vec2>(h, h.face_uvs_offset, optional(NXG_CURVE_CACHE_HAS_FACE_UVS, h.strand_count)) ||
        !section_fits<uint32_t>(h, h.face_ids_offset, optional(NXG_CURVE_CACHE_HAS_FACE_IDS, h.strand_count)))
        return fail("one or more curve-cache sections are out of bounds");
    if (nxg::curve_cache_hash(bytes + sizeof(h), size - sizeof(h)) != h.content_hash)
        return fail("curve-cache content hash mismatch");
    auto at = [&](uint64_t offset) { return offset ? bytes + offset : nullptr; };
    const uint32_t *counts = reinterpret_cast<const uint32_t *>(at(h.point_counts_offset));
    uint64_t sum = 0u;
    for (uint32_t s = 0; s < h.strand_count; ++s) {
        if (counts[s] < 2u) return fail("curve cache contains a curve with fewer than two CVs");
        sum += counts[s];
    }
    if (sum != h.point_count) return fail("curve-cache point counts do not sum to point_count");
    const nxg_packed_point *points = reinterpret_cast<const nxg_packed_point *>(at(h.points_offset));
    const nxg_vec2 *tex = reinterpret_cast<const nxg_vec2 *>(at(h.texcoords_offset));
    for (uint32_t i = 0; i < h.point_count; ++i) {
        const nxg_packed_point v = points[i];
        if (!std::isfinite(v.x) || !std::isfinite(v.y) || !std::isfinite(v.z) || !std::isfinite(v.radius) ||
            v.radius < 0.0f)
            return fail("curve cache contains an invalid point or radius");
        if (tex && !nxg::finite2(tex[i])) return fail("curve cache contains an invalid texcoord");
    }
    const float *times = reinterpret_cast<const float *>(at(h.motion_times_offset));
    float previous = -std::numeric_limits<float>::infinity();
    for (uint32_t k = 0; k < h.motion_sample_count; ++k) {
        if (!std::isfinite(times[k]) || times[k] <= previous) return fail("curve cache contains invalid motion times");
        previous = times[k];
    }
    const nxg_vec3 *mp = reinterpret_cast<const nxg_vec3 *>(at(h.motion_points_offset));
    for (uint64_t i = 0; i < motion_points; ++i)
        if (!std::isfinite(mp[i].x) || !std::isfinite(mp[i].y) || !std::isfinite(mp[i].z))
            return fail("curve cache contains an invalid motion position");
    const nxg_vec2 *patch = reinterpret_cast<const nxg_vec2 *>(at(h.patch_uvs_offset));
    const nxg_vec2 *face = reinterpret_cast<const nxg_vec2 *>(at(h.face_uvs_offset));
    for (uint32_t s = 0; s < h.strand_count; ++s) {
        if (patch && !nxg::finite2(patch[s])) return fail("curve cache contains an invalid patch UV");
        if (face && !nxg::finite2(face[s])) return fail("curve cache contains an invalid face UV");
    }
    return NXG_OK;
}

}  // extern "C"
