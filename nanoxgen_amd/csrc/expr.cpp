// Host half of the float expression runtime: program validation and packing
// into a position-independent device image.  Validation restates the checks
// the reference interpreter makes while it runs
// (src/xgen_expression.cpp:1007-1148: operand ranges, SSA order, arity,
// input and ramp indices, result index), so the device interpreter can run
// without them; only non-finite values remain a run-time condition.
#include <cstring>
#include <functional>
#include <queue>
#include <vector>

#include "expr_format.h"
#include "internal.h"

namespace nxg {
namespace {

// Required operand count per op; -1: any count up to kMaxOperands; -2: 0..3.
int arity(uint32_t op) {
    switch (op) {
        case NXG_OP_CONSTANT: case NXG_OP_INPUT: case NXG_OP_RAMP: return 0;
        case NXG_OP_NEGATE: case NXG_OP_LOGICAL_NOT: return 1;
        case NXG_OP_ADD: case NXG_OP_SUBTRACT: case NXG_OP_MULTIPLY: case NXG_OP_DIVIDE:
        case NXG_OP_LESS: case NXG_OP_LESS_EQUAL: case NXG_OP_GREATER: case NXG_OP_GREATER_EQUAL:
        case NXG_OP_EQUAL: case NXG_OP_NOT_EQUAL: case NXG_OP_LOGICAL_AND: case NXG_OP_LOGICAL_OR:
        case NXG_OP_MINIMUM: case NXG_OP_MAXIMUM: case NXG_OP_GAMMA: case NXG_OP_CONTRAST:
        case NXG_OP_POW:
            return 2;
        case NXG_OP_SELECT: case NXG_OP_CLAMP: case NXG_OP_SMOOTHSTEP: case NXG_OP_MIX:
        case NXG_OP_NOISE3: case NXG_OP_TEXEL:
            return 3;
        case NXG_OP_FIT: return 5;
        case NXG_OP_HASH: return -1;
        case NXG_OP_RANDOM: return -2;
        default: return -3;
    }
}

}  // namespace

bool validate_program(const nxg_expr_program &p) {
    if (p.instruction_count == 0u || p.instruction_count > NXG_EXPR_MAX_VALUES) return false;
    if (!p.instructions || (p.operand_count && !p.operands)) return false;
    if (p.input_count > kMaxInputs || (p.input_count && !p.input_bindings)) return false;
    for (uint32_t i = 0; i < p.input_count; ++i)
        if (p.input_bindings[i] >= NXG_ROOT_VAR_COUNT) return false;
    if ((p.ramp_count && !p.ramps) || (p.ramp_point_count && !p.ramp_points)) return false;
    for (uint32_t r = 0; r < p.ramp_count; ++r) {
        const nxg_expr_ramp ramp = p.ramps[r];
        if (ramp.point_count == 0u || uint64_t{ramp.point_offset} + ramp.point_count > p.ramp_point_count)
            return false;
        for (uint32_t k = 0; k < ramp.point_count; ++k)
            if (p.ramp_points[ramp.point_offset + k].interpolation > 3u) return false;
    }
    for (uint32_t i = 0; i < p.instruction_count; ++i) {
        const nxg_expr_instruction &in = p.instructions[i];
        const int need = arity(in.op);
        if (need == -3) return false;
        if (need >= 0 && in.operand_count != static_cast<uint32_t>(need)) return false;
        if (need == -1 && (in.operand_count == 0u || in.operand_count > kMaxOperands)) return false;
        if (need == -2 && in.operand_count > 3u) return false;
        if (uint64_t{in.operand_offset} + in.operand_count > p.operand_count) return false;
        for (uint32_t k = 0; k < in.operand_count; ++k)
            if (p.operands[in.operand_offset + k] >= i) return false;
        if (in.op == NXG_OP_INPUT && in.auxiliary >= p.input_count) return false;
        if (in.op == NXG_OP_RAMP && in.auxiliary >= p.ramp_count) return false;
        if (in.op == NXG_OP_TEXEL && in.auxiliary >= NXG_EXPR_MAX_TEXEL_MAPS) return false;
    }
    return p.result < p.instruction_count;
}

namespace {

uint64_t align16(uint64_t v) { return (v + 15u) & ~uint64_t{15u}; }

uint64_t program_bytes(const nxg_expr_program &p) {
    return align16(sizeof(PackedProgram)) + align16(uint64_t{p.instruction_count + 1u} * sizeof(PackedInstruction)) +
           align16(uint64_t{p.operand_count} * 4u) + align16(uint64_t{p.ramp_count} * sizeof(nxg_expr_ramp)) +
           align16(uint64_t{p.ramp_point_count} * sizeof(nxg_expr_ramp_point));
}


// Liveness-based slot allocation over the SSA order of a validated program:
// a value's slot returns to the pool after its last reader (operands are
// read before the result is written, so a result may take the slot of one of
// its own operands); the result value is never released; inputs use their
// root variable's reserved slot.  Returns the slot count, vars included.
uint32_t allocate_slots(const nxg_expr_program &p, std::vector<uint16_t> &slot_of) {
    const uint32_t n = p.instruction_count;
    std::vector<int64_t> last_use(n, -1);
    for (uint32_t i = 0; i < n; ++i) {
        const nxg_expr_instruction &in = p.instructions[i];
        for (uint32_t k = 0; k < in.operand_count; ++k) last_use[p.operands[in.operand_offset + k]] = i;
    }
    last_use[p.result] = n;
    slot_of.assign(n, 0u);
    std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> free_slots;
    std::vector<bool> released(n, false);
    uint32_t next = kVarSlots;
    for (uint32_t i = 0; i < n; ++i) {
        const nxg_expr_instruction &in = p.instructions[i];
        for (uint32_t k = 0; k < in.operand_count; ++k) {
            const uint32_t v = p.operands[in.operand_offset + k];
            if (last_use[v] == i && !released[v] && p.instructions[v].op != NXG_OP_INPUT) {
                released[v] = true;
                free_slots.push(slot_of[v]);
            }
        }
        if (in.op == NXG_OP_INPUT) {
            slot_of[i] = static_cast<uint16_t>(p.input_bindings[in.auxiliary]);
            continue;
        }
        uint32_t slot;
        if (free_slots.empty()) {
            slot = next++;
        } else {
            slot = free_slots.top();
            free_slots.pop();
        }
        slot_of[i] = static_cast<uint16_t>(slot);
        if (last_use[i] < 0) free_slots.push(slot);  // never read: free after the write
    }
    return next;
}

}  // namespace
}  // namespace nxg

extern "C" {

uint32_t nxg_expr_validate(const nxg_expr_program *program) {
    if (!program) return NXG_ERROR_NULL_ARGUMENT;
    return nxg::validate_program(*program) ? NXG_OK : NXG_ERROR_INVALID_PROGRAM;
}

uint32_t nxg_expr_pack(const nxg_expr_program *programs, uint32_t count, void *out, uint64_t capacity,
                       nxg_expr_set *info) {
    using namespace nxg;
    if (!programs || count == 0u || count > NXG_EXPR_MAX_PROGRAMS) return NXG_ERROR_INVALID_PROGRAM;
    uint64_t total = align16(sizeof(PackedSet));
    uint32_t max_values = 0, maps_required = 0;
    std::vector<std::vector<uint16_t>> slots(count);
    std::vector<uint32_t> slot_counts(count);
    for (uint32_t k = 0; k < count; ++k) {
        if (!validate_program(programs[k])) return NXG_ERROR_INVALID_PROGRAM;
        for (uint32_t i = 0; i < programs[k].instruction_count; ++i) {
            const nxg_expr_instruction &in = programs[k].instructions[i];
            if (in.op == NXG_OP_TEXEL && in.auxiliary + 1u > maps_required) maps_required = in.auxiliary + 1u;
        }
        total += program_bytes(programs[k]);
        slot_counts[k] = allocate_slots(programs[k], slots[k]);
        if (slot_counts[k] > max_values) max_values = slot_counts[k];
    }
    if (!info) return NXG_ERROR_NULL_ARGUMENT;
    info->byte_size = total;
    if (!out) return NXG_OK;
    if (capacity < total) return NXG_ERROR_BUFFER_TOO_SMALL;
    uint8_t *base = static_cast<uint8_t *>(out);
    std::memset(base, 0, total);
    PackedSet set{};
    set.magic = kPackedMagic;
    set.program_count = count;
    set.max_values = max_values;
    set.byte_size = static_cast<uint32_t>(total);
    uint64_t at = align16(sizeof(PackedSet));
    for (uint32_t k = 0; k < count; ++k) {
        const nxg_expr_program &p = programs[k];
        const std::vector<uint16_t> &slot_of = slots[k];
        set.program_offset[k] = static_cast<uint32_t>(at);
        PackedProgram pp{};
        pp.instruction_count = p.instruction_count;
        pp.result = slot_of[p.result];
        pp.input_count = p.input_count;
        pp.slot_count = slot_counts[k];
        for (uint32_t i = 0; i < p.input_count; ++i) pp.input_bindings[i] = p.input_bindings[i];
        uint64_t cursor = at + align16(sizeof(PackedProgram));
        pp.instructions_offset = static_cast<uint32_t>(cursor);
        cursor += align16(uint64_t{p.instruction_count + 1u} * sizeof(PackedInstruction));
        pp.operands_offset = static_cast<uint32_t>(cursor);
        cursor += align16(uint64_t{p.operand_count} * 4u);
        pp.ramps_offset = static_cast<uint32_t>(cursor);
        cursor += align16(uint64_t{p.ramp_count} * sizeof(nxg_expr_ramp));
        pp.ramp_points_offset = static_cast<uint32_t>(cursor);
        cursor += align16(uint64_t{p.ramp_point_count} * sizeof(nxg_expr_ramp_point));
        std::memcpy(base + at, &pp, sizeof(pp));
        PackedInstruction *ins = reinterpret_cast<PackedInstruction *>(base + pp.instructions_offset);
        uint32_t *operand_slots = reinterpret_cast<uint32_t *>(base + pp.operands_offset);
        for (uint32_t i = 0; i < p.instruction_count; ++i) {
            const nxg_expr_instruction &in = p.instructions[i];
            PackedInstruction &o = ins[i];
            o.op = in.op;
            o.operand_count = in.operand_count;
            o.dest = slot_of[i];
            o.operand_offset = in.operand_offset;
            o.auxiliary = in.op == NXG_OP_INPUT ? p.input_bindings[in.auxiliary] : in.auxiliary;
            o.immediate = in.immediate;
            for (uint32_t j = 0; j < in.operand_count; ++j) {
                const uint16_t s = slot_of[p.operands[in.operand_offset + j]];
                if (in.operand_count <= 4u) o.slot[j] = s;
                operand_slots[in.operand_offset + j] = s;
            }
        }
        if (p.ramp_count)
            std::memcpy(base + pp.ramps_offset, p.ramps, uint64_t{p.ramp_count} * sizeof(nxg_expr_ramp));
        if (p.ramp_point_count)
            std::memcpy(base + pp.ramp_points_offset, p.ramp_points,
                        uint64_t{p.ramp_point_count} * sizeof(nxg_expr_ramp_point));
        at = cursor;
    }
    std::memcpy(base, &set, sizeof(set));
    info->program_count = set.program_count;
    info->max_values = set.max_values;
    info->texel_maps_required = maps_required;
    for (uint32_t k = 0; k < NXG_EXPR_MAX_PROGRAMS; ++k) info->program_offset[k] = set.program_offset[k];
    return NXG_OK;
}

}  // extern "C"
