/*
 * ORACLE (test infrastructure only): restatement of the library's opt-in
 * inverse-CDF root sampler, which has no reference counterpart (it replaces
 * the alias draw of sample_root, generate.h:107-112, when a caller asks for
 * it).  Stated from its definition, sequentially:
 *   mass of bucket j      keep_j = ceil(prob_j * 2^24) * 2^8 (prob_j >= 1:
 *                         2^32; not > 0, NaN included: 0) stays with j; the
 *                         rest of its 2^32 goes to alias_j, or to j itself
 *                         when alias_j is out of range
 *   prefix sums           C[t] = sum of the masses of triangles 0..t (u64)
 *   target of a strand    floor((h0 * 2^32 + h1) * T / 2^32), h0 and h1 the
 *                         hashes behind random01's dimensions 0 and 1
 *   triangle              the smallest t with C[t] > target
 *   root                  sample_root's barycentrics and interpolation on
 *                         that triangle (oracle_strand.inc)
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "oracle.h"

static uint64_t keep_of(float p) {
    uint64_t k = 0u;
    if (p >= 1.0f) k = (uint64_t)1u << 24;
    else if (p > 0.0f) k = (uint64_t)ceil((double)p * 16777216.0);
    return k << 8;
}

void orc_cdf_prefix(const void *blob, uint64_t *prefix) {
    const nxg_asset_header *h = (const nxg_asset_header *)blob;
    const nxg_alias_entry *alias = (const nxg_alias_entry *)((const uint8_t *)blob + h->alias_table_offset);
    const uint32_t T = h->triangle_count;
    memset(prefix, 0, sizeof(uint64_t) * T);
    for (uint32_t j = 0; j < T; ++j) {
        const uint64_t keep = keep_of(alias[j].probability);
        prefix[j] += keep;
        prefix[alias[j].alias < T ? alias[j].alias : j] += ((uint64_t)1u << 32) - keep;
    }
    for (uint32_t t = 1; t < T; ++t) prefix[t] += prefix[t - 1];
}

uint32_t orc_cdf_triangle(const uint64_t *prefix, uint32_t T, uint32_t seed, uint32_t strand) {
    const uint64_t h0 = orc_hash32(seed ^ orc_hash32(strand + 0x9e3779b9u * 1u));
    const uint64_t h1 = orc_hash32(seed ^ orc_hash32(strand + 0x9e3779b9u * 2u));
    const uint64_t target = h0 * T + ((h1 * T) >> 32);
    uint32_t lo = 0u, hi = T - 1u;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        if (prefix[mid] > target) hi = mid;
        else lo = mid + 1u;
    }
    return lo;
}

void orc_cdf_roots(const void *blob, const uint64_t *prefix, uint32_t seed, uint32_t strand_count,
                   nxg_root_sample *roots) {
    const nxg_asset_header *h = (const nxg_asset_header *)blob;
    for (uint32_t s = 0; s < strand_count; ++s)
        orc_root_on_triangle_f32(blob, seed, s, orc_cdf_triangle(prefix, h->triangle_count, seed, s), &roots[s]);
}
