// Device half of the float expression runtime, shared by the generator's
// root-attribute kernel and the Classic primitive kernels: the SeExpr hash
// helpers, ramp evaluation and instruction fetch from a packed program image
// in the constant address space.  One instruction's evaluation is the
// fragment expr_ops.inc, expanded inside each interpreter loop.  Each
// includer gets its own copy (file-local definitions).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/nanoxgen_amd.h"
#include "expr_format.h"

namespace nxg {
namespace {

__constant__ uint8_t kPermutation[256] = {
#include "seexpr_permutation.inc"
};

// x ^ (y & m) in one v_bitop3 (truth table 0x78, indexed x*4 + y*2 + m).
__device__ __forceinline__ uint32_t xor_and(uint32_t x, uint32_t y, uint32_t m) {
    return __builtin_amdgcn_bitop3_b32(x, y, m, 0x78);
}

// xgen_seexpr_hash_finish (xgen_expression.cpp:637-651)
__device__ uint32_t seexpr_hash_finish(uint32_t state) {
    state ^= state >> 11u;
    state = xor_and(state, state << 7u, 0x9d2c5680u);
    state = xor_and(state, state << 15u, 0xefc60000u);
    state ^= state >> 18u;
    const uint32_t b3 = kPermutation[state & 0xffu];
    const uint32_t b2 = kPermutation[((state >> 8u) + b3) & 0xffu];
    const uint32_t b1 = kPermutation[((state >> 16u) + b2) & 0xffu];
    const uint32_t b0 = kPermutation[((state >> 24u) + b1) & 0xffu];
    return b0 | (b1 << 8u) | (b2 << 16u) | (b3 << 24u);
}

// xgen_runtime_hash_component (xgen_expression.cpp:663-709): the top 32
// fraction bits of frexp(double(x) * 8.539734222673566) xor the exponent,
// from a 24 x 53-bit integer product (no FP64).
__device__ uint32_t runtime_hash_component(float argument) {
    constexpr uint64_t cm = 0x001114580b45d474ull;
    const uint32_t bits = __float_as_uint(argument);
    const uint32_t absolute = bits & 0x7fffffffu;
    if (absolute == 0u) return 0u;
    const uint32_t fexp = absolute >> 23u;
    const uint32_t mant = fexp == 0u ? (absolute & 0x7fffffu) : ((absolute & 0x7fffffu) | 0x800000u);
    const uint64_t p0 = uint64_t{mant} * static_cast<uint32_t>(cm);
    const uint64_t p1 = uint64_t{mant} * static_cast<uint32_t>(cm >> 32u);
    const uint64_t low = p0 + (p1 << 32u);
    const uint64_t high = (p1 >> 32u) + (low < p0 ? 1u : 0u);
    const uint32_t length = high != 0u ? 128u - static_cast<uint32_t>(__clzll(high))
                                       : 64u - static_cast<uint32_t>(__clzll(low));
    const uint32_t shift = length - 32u;
    const uint32_t top = static_cast<uint32_t>((low >> shift) | (high << (64u - shift)));
    const uint64_t remainder = low & ((uint64_t{1} << shift) - 1u);
    const uint64_t ceil_div = ((low >> 32u) | (high << 32u)) + ((low & 0xffffffffu) != 0u ? 1u : 0u);
    const uint32_t fraction = top - (remainder < ceil_div ? 1u : 0u);
    const int32_t value_exponent = fexp == 0u ? -149 : static_cast<int32_t>(fexp) - 150;
    const uint32_t exponent = static_cast<uint32_t>(static_cast<int32_t>(length) + value_exponent - 49);
    return ((bits >> 31u) != 0u ? 0u - fraction : fraction) ^ exponent;
}

__device__ __forceinline__ uint32_t hash_step(uint32_t state, float x) {
    return state * 1664525u + runtime_hash_component(x) + 1013904223u;
}
__device__ __forceinline__ float hash_to_unit(uint32_t state) {
    return static_cast<float>(seexpr_hash_finish(state)) * 0x1p-32f;
}

__device__ __forceinline__ float clamp_std(float v, float lo, float hi) {  // std::clamp
    return v < lo ? lo : (hi < v ? hi : v);
}

// Pointers into a packed program image, in the constant address space (the
// image is read-only for a whole launch, so wave-uniform reads of it are
// scalar loads).
#define ImageOf(T) const __attribute__((address_space(4))) T *
typedef const __attribute__((address_space(4))) uint8_t *Image;

// A whole instruction as eight scalar-loaded words (structs cannot be
// copied out of another address space directly).
__device__ __forceinline__ PackedInstruction load_instruction(ImageOf(PackedInstruction) p) {
    ImageOf(uint32_t) w = (ImageOf(uint32_t))p;
    uint32_t v[sizeof(PackedInstruction) / 4u];
#pragma unroll
    for (uint32_t j = 0; j < sizeof(PackedInstruction) / 4u; ++j) v[j] = w[j];
    PackedInstruction r;
    __builtin_memcpy(&r, v, sizeof(r));
    return r;
}

// evaluate_ramp_points<float> (xgen_expression.cpp:559-615)
__device__ float ramp_eval(ImageOf(nxg_expr_ramp_point) pts, uint32_t n, float x) {
    if (x <= pts[0].position) return clamp_std(pts[0].value, 0.0f, 1.0f);
    if (x >= pts[n - 1u].position) return clamp_std(pts[n - 1u].value, 0.0f, 1.0f);
    uint32_t upper = 1u;
    while (upper < n && x >= pts[upper].position) ++upper;
    const uint32_t lower = upper - 1u;
    const float p0 = pts[lower].position, p1 = pts[upper].position;
    const float k0 = pts[lower].value, k1 = pts[upper].value;
    const float u = (x - p0) / (p1 - p0);
    float value = k0;
    const uint32_t mode = pts[lower].interpolation;
    if (mode == 1u) {
        value = k0 + u * (k1 - k0);
    } else if (mode >= 2u) {
        value = k0 * (u - 1.0f) * (u - 1.0f) * (2.0f * u + 1.0f) + k1 * u * u * (3.0f - 2.0f * u);
        if (mode == 3u) {
            const float scale = 1.0f / (p1 - p0);
            const float d0 = lower == 0u ? (k1 - k0) / (scale * (2.0f * p1 - 2.0f * p0))
                                         : (k1 - pts[lower - 1u].value) / (scale * (p1 - pts[lower - 1u].position));
            const float d1 = upper + 1u == n ? (k1 - k0) / (scale * (2.0f * p1 - 2.0f * p0))
                                             : (pts[upper + 1u].value - k0) / (scale * (pts[upper + 1u].position - p0));
            value = k0 * (u - 1.0f) * (u - 1.0f) * (2.0f * u + 1.0f) + k1 * u * u * (3.0f - 2.0f * u) +
                    d0 * u * (u - 1.0f) * (u - 1.0f) + d1 * u * u * (u - 1.0f);
        }
    }
    return clamp_std(value, 0.0f, 1.0f);
}

// XgenExpressionFloatContext (xgen_expression.h:124-132) less its inputs.
struct ExprContext {
    float u, v, face_seed, t;
    uint32_t random_prefix;
    bool has_random_prefix;
};

}  // namespace
}  // namespace nxg
