// Host side of the .nxg asset format: builder, validator and device
// descriptor, behind the C ABI.  Restates the reference's build_asset /
// validate_asset / make_device_asset_descriptor (src/asset.cpp:16-378) with
// the same byte layout, the same float/double operation order and the same
// error messages.  The O(triangles x guides) stencil search runs on all host
// cores; its result does not depend on the thread count.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "internal.h"

namespace {

struct F3 {
    float x, y, z;
};
inline F3 f3(const nxg_vec3 &v) { return {v.x, v.y, v.z}; }
inline F3 operator+(F3 a, F3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline F3 operator-(F3 a, F3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline F3 operator/(F3 a, float s) { return {a.x / s, a.y / s, a.z / s}; }
inline float dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 cross(F3 a, F3 b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
inline F3 normalize(F3 v) {  // generate.h:67-70
    const float l2 = dot(v, v);
    return l2 > 1.0e-20f ? v / std::sqrt(l2) : F3{0.0f, 1.0f, 0.0f};
}
inline nxg_vec3 vec(F3 a) { return {a.x, a.y, a.z}; }

constexpr uint64_t kHeader = sizeof(nxg_asset_header);
uint64_t align16(uint64_t v) { return (v + 15u) & ~uint64_t{15u}; }

const float kGradients[NXG_NOISE_GRADIENT_COUNT][3] = {
#define NXG_GRADIENT(x, y, z) {x, y, z},
#include "seexpr_gradients.inc"
#undef NXG_GRADIENT
};

struct Layout {
    nxg_asset_header header{};
    uint64_t total_cvs{};
};

// Section placement of asset.cpp:19-32 and 271-288: every section starts
// 16-byte aligned after the 128-byte header; an absent section has offset 0.
Layout plan(const nxg_asset_build_input &in) {
    Layout L;
    for (uint32_t g = 0; g < in.guide_count; ++g) L.total_cvs += in.guide_cv_counts[g];
    nxg_asset_header &h = L.header;
    h.magic = NXG_MAGIC;
    h.version_major = NXG_VERSION_MAJOR;
    h.version_minor = NXG_VERSION_MINOR;
    h.flags = NXG_ASSET_HAS_NORMALS | (in.texcoords ? NXG_ASSET_HAS_TEXCOORDS : 0u);
    h.vertex_count = in.vertex_count;
    h.triangle_count = in.triangle_count;
    h.guide_count = in.guide_count;
    h.guide_cv_count = static_cast<uint32_t>(L.total_cvs);
    h.guide_stencil_size = NXG_GUIDE_STENCIL_SIZE;
    h.noise_gradient_count = NXG_NOISE_GRADIENT_COUNT;
    uint64_t end = kHeader;
    auto place = [&end](uint64_t bytes) {
        const uint64_t off = align16(end);
        end = off + bytes;
        return off;
    };
    h.positions_offset = place(uint64_t{in.vertex_count} * 12u);
    h.normals_offset = place(uint64_t{in.vertex_count} * 12u);
    h.texcoords_offset = in.texcoords ? place(uint64_t{in.vertex_count} * 8u) : 0u;
    h.triangles_offset = place(uint64_t{in.triangle_count} * 12u);
    h.alias_table_offset = place(uint64_t{in.triangle_count} * 8u);
    h.guides_offset = place(uint64_t{in.guide_count} * sizeof(nxg_guide_record));
    h.guide_cvs_offset = place(L.total_cvs * 12u);
    h.triangle_guides_offset = place(uint64_t{in.triangle_count} * NXG_GUIDE_STENCIL_SIZE * 4u);
    h.noise_gradients_offset = place(uint64_t{NXG_NOISE_GRADIENT_COUNT} * 12u);
    h.byte_size = end;
    return L;
}

// check_input (asset.cpp:99-150) for the flattened input, plus the two
// extension fields.  Returns nullptr when the input is acceptable.
const char *check_input(const nxg_asset_build_input &in) {
    if (!in.positions || in.vertex_count == 0u) return "asset needs vertices";
    if (!in.triangles || in.triangle_count == 0u) return "asset needs triangles";
    if (in.guide_count == 0u || !in.guide_cv_counts || !in.guide_cvs)
        return "asset needs at least one guide";
    for (uint32_t i = 0; i < in.triangle_count; ++i) {
        const nxg_uint3 t = in.triangles[i];
        if (t.x >= in.vertex_count || t.y >= in.vertex_count || t.z >= in.vertex_count)
            return "triangle index out of bounds";
    }
    for (uint32_t g = 0; g < in.guide_count; ++g) {
        if (in.guide_cv_counts[g] < 2u || in.guide_cv_counts[g] > 65535u)
            return "each guide needs 2..65535 CVs";
    }
    if (in.stencil_k > NXG_GUIDE_STENCIL_SIZE) return "stencil_k must be in 0..8";
    if (in.triangle_weights) {
        for (uint32_t i = 0; i < in.triangle_count; ++i) {
            if (!std::isfinite(in.triangle_weights[i]) || in.triangle_weights[i] < 0.0f)
                return "triangle weights must be finite and non-negative";
        }
    }
    return nullptr;
}

void set_err(char *err, uint64_t cap, const char *msg) {
    if (err && cap) std::snprintf(err, cap, "%s", msg);
}

// Walker alias table (asset.cpp:43-73): areas summed in double, LIFO
// small/large worklists, probability stored as float.
bool build_alias(const nxg_asset_build_input &in, nxg_alias_entry *table) {
    const std::size_t n = in.triangle_count;
    std::vector<double> scaled(n);
    double area_sum = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
        const nxg_uint3 t = in.triangles[i];
        const F3 a = f3(in.positions[t.x]);
        const F3 c = cross(f3(in.positions[t.y]) - a, f3(in.positions[t.z]) - a);
        scaled[i] = static_cast<double>(0.5f * std::sqrt(dot(c, c)));
        if (in.triangle_weights) scaled[i] = scaled[i] * static_cast<double>(in.triangle_weights[i]);
        area_sum += scaled[i];
    }
    if (!(area_sum > 0.0)) return false;
    for (double &p : scaled) p = p * static_cast<double>(n) / area_sum;
    std::vector<uint32_t> small, large;
    small.reserve(n);
    large.reserve(n);
    for (std::size_t i = 0; i < n; ++i) (scaled[i] < 1.0 ? small : large).push_back(uint32_t(i));
    while (!small.empty() && !large.empty()) {
        const uint32_t s = small.back();
        small.pop_back();
        const uint32_t l = large.back();
        large.pop_back();
        table[s] = {static_cast<float>(scaled[s]), l};
        scaled[l] = scaled[l] + scaled[s] - 1.0;
        (scaled[l] < 1.0 ? small : large).push_back(l);
    }
    for (const uint32_t i : large) table[i] = {1.0f, i};
    for (const uint32_t i : small) table[i] = {1.0f, i};
    return true;
}

unsigned worker_count(uint64_t work_items) {
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const uint64_t useful = std::max<uint64_t>(1u, work_items / 64u);
    return static_cast<unsigned>(std::min<uint64_t>(hw, useful));
}

template <typename Fn>
void parallel_ranges(uint64_t count, uint64_t chunk, Fn &&fn) {
    const uint64_t chunks = (count + chunk - 1u) / chunk;
    const unsigned workers = worker_count(chunks * 64u);
    std::atomic<uint64_t> next{0};
    auto run = [&] {
        for (;;) {
            const uint64_t c = next.fetch_add(1u, std::memory_order_relaxed);
            if (c >= chunks) return;
            fn(c * chunk, std::min(count, (c + 1u) * chunk));
        }
    };
    if (workers <= 1u) {
        run();
        return;
    }
    std::vector<std::thread> pool;
    for (unsigned i = 0; i < workers; ++i) pool.emplace_back(run);
    for (auto &t : pool) t.join();
}

// Triangle guide stencil (asset.cpp:252-269): the k smallest keys
// d2(centroid, guide root) + (1e20 if dot(face normal, guide normal) < -0.25),
// ties broken by the lower guide index, in ascending order.  Each worker keeps
// a sorted top-k by insertion; guides are visited in index order, so an equal
// key never displaces an earlier index, which is the (key, index) order the
// reference's std::partial_sort over pairs produces.
void build_stencil(const nxg_asset_build_input &in, const nxg_guide_record *guides,
                   uint32_t k, uint32_t *stencil) {
    const uint32_t G = in.guide_count;
    const uint32_t count = std::min(k, G);
    parallel_ranges(in.triangle_count, 256u, [&](uint64_t begin, uint64_t end) {
        float keys[NXG_GUIDE_STENCIL_SIZE];
        uint32_t ids[NXG_GUIDE_STENCIL_SIZE];
        for (uint64_t ti = begin; ti < end; ++ti) {
            const nxg_uint3 t = in.triangles[ti];
            const F3 p0 = f3(in.positions[t.x]), p1 = f3(in.positions[t.y]), p2 = f3(in.positions[t.z]);
            const F3 centroid = (p0 + p1 + p2) / 3.0f;
            const F3 fn = normalize(cross(p1 - p0, p2 - p0));
            uint32_t filled = 0;
            for (uint32_t g = 0; g < G; ++g) {
                const float penalty = dot(fn, f3(guides[g].root_normal)) < -0.25f ? 1.0e20f : 0.0f;
                const F3 d = centroid - f3(guides[g].root_position);
                const float key = dot(d, d) + penalty;
                if (filled == count && !(key < keys[count - 1u])) continue;
                uint32_t pos = filled < count ? filled++ : count - 1u;
                while (pos > 0u && key < keys[pos - 1u]) {
                    keys[pos] = keys[pos - 1u];
                    ids[pos] = ids[pos - 1u];
                    --pos;
                }
                keys[pos] = key;
                ids[pos] = g;
            }
            uint32_t *row = stencil + ti * NXG_GUIDE_STENCIL_SIZE;
            for (uint32_t j = 0; j < NXG_GUIDE_STENCIL_SIZE; ++j)
                row[j] = j < filled ? ids[j] : NXG_INVALID_INDEX;
        }
    });
}

uint64_t fnv1a(const uint8_t *b, uint64_t n) {  // asset.cpp:34-41
    uint64_t hash = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; ++i) {
        hash ^= b[i];
        hash *= 1099511628211ull;
    }
    return hash;
}

bool section_fits(const nxg_asset_header &h, uint64_t offset, uint64_t count, uint64_t elem) {
    if (count == 0u) return offset == 0u;  // asset.cpp:152-158
    if (offset < kHeader || offset > h.byte_size) return false;
    return count * elem <= h.byte_size - offset;
}

const char *validate(const void *blob, uint64_t size) {  // asset.cpp:295-338
    if (size < kHeader) return "blob is smaller than AssetHeader";
    if (reinterpret_cast<uintptr_t>(blob) % NXG_BLOB_ALIGNMENT != 0u)
        return "blob base is not 64-byte aligned";
    nxg_asset_header h;
    std::memcpy(&h, blob, sizeof(h));
    if (h.magic != NXG_MAGIC) return "bad NanoXGen magic";
    if (h.version_major != NXG_VERSION_MAJOR || h.version_minor != NXG_VERSION_MINOR)
        return "unsupported NanoXGen asset version";
    if (h.byte_size != size) return "header byte_size does not match file size";
    if (h.triangle_count == 0u || h.vertex_count == 0u || h.guide_count == 0u)
        return "required section is empty";
    if (h.guide_stencil_size != NXG_GUIDE_STENCIL_SIZE) return "unsupported guide stencil size";
    if (h.noise_gradient_count != NXG_NOISE_GRADIENT_COUNT) return "unsupported noise gradient table";
    if (!section_fits(h, h.positions_offset, h.vertex_count, 12u) ||
        !section_fits(h, h.normals_offset, h.vertex_count, 12u) ||
        !section_fits(h, h.texcoords_offset, (h.flags & NXG_ASSET_HAS_TEXCOORDS) ? h.vertex_count : 0u, 8u) ||
        !section_fits(h, h.triangles_offset, h.triangle_count, 12u) ||
        !section_fits(h, h.alias_table_offset, h.triangle_count, 8u) ||
        !section_fits(h, h.guides_offset, h.guide_count, sizeof(nxg_guide_record)) ||
        !section_fits(h, h.guide_cvs_offset, h.guide_cv_count, 12u) ||
        !section_fits(h, h.triangle_guides_offset, uint64_t{h.triangle_count} * h.guide_stencil_size, 4u) ||
        !section_fits(h, h.noise_gradients_offset, h.noise_gradient_count, 12u))
        return "one or more sections are out of bounds";
    const uint8_t *b = static_cast<const uint8_t *>(blob);
    if (fnv1a(b + kHeader, size - kHeader) != h.content_hash) return "content hash mismatch";
    const nxg_uint3 *tris = reinterpret_cast<const nxg_uint3 *>(b + h.triangles_offset);
    for (uint32_t i = 0; i < h.triangle_count; ++i) {
        if (tris[i].x >= h.vertex_count || tris[i].y >= h.vertex_count || tris[i].z >= h.vertex_count)
            return "triangle index out of bounds";
    }
    const nxg_guide_record *g = reinterpret_cast<const nxg_guide_record *>(b + h.guides_offset);
    for (uint32_t i = 0; i < h.guide_count; ++i) {
        if (g[i].cv_count < 2u || g[i].first_cv > h.guide_cv_count ||
            g[i].cv_count > h.guide_cv_count - g[i].first_cv)
            return "guide CV range out of bounds";
    }
    return nullptr;
}

}  // namespace

extern "C" {

uint32_t nxg_asset_build_size(const nxg_asset_build_input *input, uint64_t *byte_size) {
    if (!input || !byte_size) return NXG_ERROR_NULL_ARGUMENT;
    if (check_input(*input)) return NXG_ERROR_INVALID_BUILD_INPUT;
    *byte_size = plan(*input).header.byte_size;
    return NXG_OK;
}

uint32_t nxg_asset_build(const nxg_asset_build_input *input, void *out, uint64_t capacity,
                         uint64_t *byte_size, char *err, uint64_t err_capacity) {
    return nxg::build_blob(input, out, capacity, byte_size, err, err_capacity, true);
}

}  // extern "C"

// The whole build with host_stencil; without it the stencil section stays
// zero and the content hash is left for the caller to set (finish_blob).
uint32_t nxg::build_blob(const nxg_asset_build_input *input, void *out, uint64_t capacity, uint64_t *byte_size,
                         char *err, uint64_t err_capacity, bool host_stencil) {
    set_err(err, err_capacity, "");
    if (!input || !out) return NXG_ERROR_NULL_ARGUMENT;
    const nxg_asset_build_input &in = *input;
    if (const char *msg = check_input(in)) {
        set_err(err, err_capacity, msg);
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }
    if (reinterpret_cast<uintptr_t>(out) % NXG_BLOB_ALIGNMENT != 0u) {
        set_err(err, err_capacity, "blob base is not 64-byte aligned");
        return NXG_ERROR_MISALIGNED_ASSET;
    }
    Layout L = plan(in);
    nxg_asset_header &h = L.header;
    if (byte_size) *byte_size = h.byte_size;
    if (capacity < h.byte_size) {
        set_err(err, err_capacity, "output buffer is too small");
        return NXG_ERROR_BUFFER_TOO_SMALL;
    }
    uint8_t *blob = static_cast<uint8_t *>(out);
    std::memset(blob, 0, h.byte_size);
    const uint32_t V = in.vertex_count, T = in.triangle_count, G = in.guide_count;

    // vertex normals (asset.cpp:75-86): given normals are kept as they are
    nxg_vec3 *normals = reinterpret_cast<nxg_vec3 *>(blob + h.normals_offset);
    if (in.normals) {
        std::memcpy(normals, in.normals, std::size_t{V} * 12u);
    } else {
        for (uint32_t i = 0; i < T; ++i) {
            const nxg_uint3 t = in.triangles[i];
            const F3 p0 = f3(in.positions[t.x]);
            const F3 n = cross(f3(in.positions[t.y]) - p0, f3(in.positions[t.z]) - p0);
            normals[t.x] = vec(f3(normals[t.x]) + n);
            normals[t.y] = vec(f3(normals[t.y]) + n);
            normals[t.z] = vec(f3(normals[t.z]) + n);
        }
        for (uint32_t i = 0; i < V; ++i) normals[i] = vec(normalize(f3(normals[i])));
    }
    if (!build_alias(in, reinterpret_cast<nxg_alias_entry *>(blob + h.alias_table_offset))) {
        set_err(err, err_capacity, "mesh has zero total area");
        return NXG_ERROR_INVALID_BUILD_INPUT;
    }

    // guide records (asset.cpp:234-250) with the automatic support radius
    // of asset.cpp:88-97 (1.75 x nearest other guide root, at least 1e-4)
    nxg_guide_record *guides = reinterpret_cast<nxg_guide_record *>(blob + h.guides_offset);
    std::vector<uint64_t> first(G);
    for (uint64_t g = 0, acc = 0; g < G; ++g) {
        first[g] = acc;
        acc += in.guide_cv_counts[g];
    }
    parallel_ranges(G, 64u, [&](uint64_t begin, uint64_t end) {
        for (uint64_t g = begin; g < end; ++g) {
            nxg_guide_record r{};
            r.first_cv = static_cast<uint32_t>(first[g]);
            r.cv_count = static_cast<uint16_t>(in.guide_cv_counts[g]);
            r.triangle_index = in.guide_triangles ? in.guide_triangles[g] : NXG_INVALID_INDEX;
            float radius = in.guide_support_radii ? in.guide_support_radii[g] : 0.0f;
            if (!(radius > 0.0f)) {
                const F3 root = f3(in.guide_cvs[first[g]]);
                float nearest = std::numeric_limits<float>::max();
                for (uint32_t j = 0; j < G; ++j) {
                    if (j == g) continue;
                    const F3 d = root - f3(in.guide_cvs[first[j]]);
                    nearest = std::min(nearest, std::sqrt(dot(d, d)));
                }
                radius = std::isfinite(nearest) ? std::max(nearest * 1.75f, 1.0e-4f) : 1.0f;
            }
            r.support_radius = radius;
            if (in.guide_barycentrics) r.barycentric = in.guide_barycentrics[g];
            if (in.guide_root_uvs) r.root_uv = in.guide_root_uvs[g];
            r.root_position = in.guide_cvs[first[g]];
            const nxg_vec3 up{0.0f, 1.0f, 0.0f};
            r.root_normal = vec(normalize(f3(in.guide_root_normals ? in.guide_root_normals[g] : up)));
            guides[g] = r;
        }
    });

    std::memcpy(blob + h.positions_offset, in.positions, std::size_t{V} * 12u);
    if (in.texcoords) std::memcpy(blob + h.texcoords_offset, in.texcoords, std::size_t{V} * 8u);
    std::memcpy(blob + h.triangles_offset, in.triangles, std::size_t{T} * 12u);
    std::memcpy(blob + h.guide_cvs_offset, in.guide_cvs, std::size_t(L.total_cvs) * 12u);
    if (host_stencil)
        build_stencil(in, guides, in.stencil_k == 0u ? NXG_GUIDE_STENCIL_SIZE : in.stencil_k,
                      reinterpret_cast<uint32_t *>(blob + h.triangle_guides_offset));
    std::memcpy(blob + h.noise_gradients_offset, kGradients, sizeof(kGradients));
    std::memcpy(blob, &h, sizeof(h));
    if (host_stencil) finish_blob(blob);
    return NXG_OK;
}

// content_hash (asset.cpp:34-41, 288): FNV-1a over every byte after the header
void nxg::finish_blob(void *out) {
    uint8_t *blob = static_cast<uint8_t *>(out);
    nxg_asset_header h;
    std::memcpy(&h, blob, sizeof(h));
    h.content_hash = fnv1a(blob + kHeader, h.byte_size - kHeader);
    std::memcpy(blob, &h, sizeof(h));
}

extern "C" {

uint32_t nxg_asset_validate(const void *blob, uint64_t byte_size, char *msg, uint64_t msg_capacity) {
    set_err(msg, msg_capacity, "");
    if (!blob) {
        set_err(msg, msg_capacity, "blob is smaller than AssetHeader");
        return NXG_ERROR_INVALID_ASSET;
    }
    if (const char *m = validate(blob, byte_size)) {
        set_err(msg, msg_capacity, m);
        return NXG_ERROR_INVALID_ASSET;
    }
    return NXG_OK;
}

uint32_t nxg_make_device_asset(const void *host_blob, uint64_t host_byte_size,
                               const void *device_data, uint64_t device_byte_capacity,
                               nxg_device_asset *out) {
    if (!out) return NXG_ERROR_NULL_ARGUMENT;
    if (!host_blob || validate(host_blob, host_byte_size)) return NXG_ERROR_INVALID_ASSET;
    if (!device_data) return NXG_ERROR_NULL_ASSET;
    if (reinterpret_cast<uintptr_t>(device_data) % NXG_BLOB_ALIGNMENT != 0u)
        return NXG_ERROR_MISALIGNED_ASSET;
    if (device_byte_capacity < host_byte_size) return NXG_ERROR_ASSET_CAPACITY_TOO_SMALL;
    out->data = device_data;
    std::memcpy(&out->header, host_blob, sizeof(nxg_asset_header));
    out->byte_capacity = device_byte_capacity;
    return NXG_OK;
}

}  // extern "C"
