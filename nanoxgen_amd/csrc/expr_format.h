// Device image of a packed expression-program set (host packer in expr.cpp,
// device interpreter in generate.hip).  All offsets are bytes from the start
// of the image, which is 16-byte aligned in device memory.
#pragma once

#include <cstdint>

#include "../../include/nanoxgen_amd.h"

namespace nxg {

constexpr uint32_t kPackedMagic = 0x5045584eu;  // "NXEP"
constexpr uint32_t kMaxInputs = 32u;
constexpr uint32_t kMaxOperands = 16u;

struct PackedSet {
    uint32_t magic;
    uint32_t program_count;
    uint32_t max_values;  // largest instruction count of the set
    uint32_t byte_size;
    uint32_t program_offset[NXG_EXPR_MAX_PROGRAMS];
};

struct PackedProgram {
    uint32_t instruction_count;
    uint32_t result;
    uint32_t input_count;
    uint32_t instructions_offset;
    uint32_t operands_offset;
    uint32_t ramps_offset;
    uint32_t ramp_points_offset;
    uint32_t reserved;
    uint32_t input_bindings[kMaxInputs];
};

// One instruction; for NXG_OP_INPUT the auxiliary already holds the bound
// root variable (NXG_ROOT_VAR_*).
struct PackedInstruction {
    uint8_t op;
    uint8_t reserved;
    uint16_t operand_count;
    uint32_t operand_offset;  // index into the program's operand array
    uint32_t auxiliary;
    float immediate;
};
static_assert(sizeof(PackedInstruction) == 16, "packed instruction is 16 bytes");

bool validate_program(const nxg_expr_program &p);

}  // namespace nxg
