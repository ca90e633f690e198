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

// LDS value slots: the root variables (NXG_ROOT_VAR_*) occupy slots
// 0..NXG_ROOT_VAR_COUNT-1 for the whole set; every other SSA value gets a
// slot from a liveness-based allocation made by the packer, so a slot is
// reused once the value it held has no later reader.
constexpr uint32_t kVarSlots = NXG_ROOT_VAR_COUNT;

struct PackedSet {
    uint32_t magic;
    uint32_t program_count;
    uint32_t max_values;  // value slots per root the largest program needs (vars included)
    uint32_t byte_size;
    uint32_t program_offset[NXG_EXPR_MAX_PROGRAMS];
};

struct PackedProgram {
    uint32_t instruction_count;    // dispatched instructions
    uint32_t result;               // slot holding the result value
    uint32_t input_count;
    uint32_t instructions_offset;  // instruction_count + 1 entries (a zeroed one ends the array)
    uint32_t operands_offset;      // operand slots of instructions with more than 4 operands
    uint32_t ramps_offset;
    uint32_t ramp_points_offset;
    uint32_t slot_count;
    uint32_t var_mask;             // root variables the program reads (checked finite once per root)
    uint32_t reserved[3];
    uint32_t input_bindings[kMaxInputs];
};

// Operand slot encoding: below kImmediate a value slot; kImmediate | k the
// instruction's k-th immediate (an inlined finite constant).
constexpr uint16_t kImmediate = 0x8000u;

// One dispatched instruction, 32 bytes (one scalar load).  Inputs are not
// dispatched (their uses read the root variable's slot) and finite constants
// used only by instructions of at most 4 operands are inlined as immediates.
struct PackedInstruction {
    uint8_t op;
    uint8_t reserved;
    uint16_t operand_count;
    uint16_t dest;            // slot the value is written to
    uint16_t reserved2;
    uint32_t operand_offset;  // operand_count > 4: index into the program's operand slots
    uint32_t auxiliary;
    float immediate;          // constant value, random call-site seed, or immediate 0
    uint16_t slot[4];         // operand_count <= 4: operand slots (see kImmediate)
    float immediate1;         // immediate 1
};
static_assert(sizeof(PackedInstruction) == 32, "packed instruction is 32 bytes");

bool validate_program(const nxg_expr_program &p);

}  // namespace nxg
