// Host half of the float expression runtime: program validation and packing
// into a position-independent device image.  Validation restates the checks
// the reference interpreter makes while it runs
// (src/xgen_expression.cpp:1007-1148: operand ranges, SSA order, arity,
// input and ramp indices, result index), so the device interpreter can run
// without them; only non-finite values remain a run-time condition.
#include <cmath>
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

namespace {

// The checks shared by both binding spaces: input bindings below
// binding_limit.
bool validate_bounded(const nxg_expr_program &p, uint32_t binding_limit) {
    if (p.instruction_count == 0u || p.instruction_count > NXG_EXPR_MAX_VALUES) return false;
    if (!p.instructions || (p.operand_count && !p.operands)) return false;
    if (p.input_count > kMaxInputs || (p.input_count && !p.input_bindings)) return false;
    for (uint32_t i = 0; i < p.input_count; ++i)
        if (p.input_bindings[i] >= binding_limit) return false;
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

}  // namespace

bool validate_program(const nxg_expr_program &p) { return validate_bounded(p, NXG_ROOT_VAR_COUNT); }

namespace {

// A Classic primitive program: NXG_CLASSIC_VAR_* bindings and the
// reference's ops only (the Classic runtime has no extension ops).
bool validate_classic_program(const nxg_expr_program &p) {
    if (!validate_bounded(p, NXG_CLASSIC_VAR_COLUMN0 + NXG_CLASSIC_MAX_COLUMNS)) return false;
    if (p.input_count > NXG_CLASSIC_MAX_PROGRAM_INPUTS) return false;
    for (uint32_t i = 0; i < p.instruction_count; ++i) {
        const uint32_t op = p.instructions[i].op;
        if (op == NXG_OP_POW || op == NXG_OP_MIX || op == NXG_OP_NOISE3 || op == NXG_OP_TEXEL) return false;
    }
    return true;
}

uint64_t align16(uint64_t v) { return (v + 15u) & ~uint64_t{15u}; }

uint64_t program_bytes(const nxg_expr_program &p) {
    return align16(sizeof(PackedProgram)) + align16(uint64_t{p.instruction_count + 1u} * sizeof(PackedInstruction)) +
           align16(uint64_t{p.operand_count} * 4u) + align16(uint64_t{p.ramp_count} * sizeof(nxg_expr_ramp)) +
           align16(uint64_t{p.ramp_point_count} * sizeof(nxg_expr_ramp_point));
}


// How one validated program is packed.
//   dispatched[i]  instruction i runs on the device (inputs never do: their
//                  uses read the root variable's slot; finite constants whose
//                  every reader has at most 4 operands and a free immediate
//                  are inlined into those readers)
//   operand[i][j]  the encoded operand slot (a value slot, or kImmediate | k)
//   imm[i][k]      the immediates of instruction i
// Slots: a liveness-based allocation over the SSA order: a value's slot
// returns to the pool after its last reader (operands are read before the
// result is written, so a result may take the slot of one of its own
// operands); the result value is never released.
struct PackPlan {
    std::vector<bool> dispatched;
    std::vector<uint16_t> slot_of;
    std::vector<std::vector<uint16_t>> operand;
    std::vector<float> imm0, imm1;
    uint32_t slot_count = kVarSlots;
    uint32_t var_mask = 0u;
    uint32_t dispatched_count = 0u;
};

bool is_finite(float v) { return std::isfinite(v); }

// classic: input i lives in slot i (the interpreter loads the program's
// inputs before it runs), so the first input_count slots are reserved;
// otherwise inputs read the root variables' reserved slots.
PackPlan plan_program(const nxg_expr_program &p, bool classic) {
    const uint32_t n = p.instruction_count;
    PackPlan plan;
    plan.dispatched.assign(n, true);
    plan.slot_of.assign(n, 0u);
    plan.operand.assign(n, {});
    plan.imm0.assign(n, 0.0f);
    plan.imm1.assign(n, 0.0f);
    auto ins = [&](uint32_t i) -> const nxg_expr_instruction & { return p.instructions[i]; };
    auto opnd = [&](uint32_t i, uint32_t j) { return p.operands[ins(i).operand_offset + j]; };
    // immediates each instruction can host: RANDOM keeps `immediate` for its seed
    std::vector<uint32_t> free_imm(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t op = ins(i).op;
        free_imm[i] = (op == NXG_OP_CONSTANT || op == NXG_OP_INPUT || ins(i).operand_count > 4u) ? 0u
                      : op == NXG_OP_RANDOM                                                      ? 1u
                                                                                                 : 2u;
    }
    // which constants can be inlined everywhere they are read
    std::vector<bool> inline_const(n, false);
    std::vector<std::vector<uint32_t>> readers(n);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < ins(i).operand_count; ++j) {
            const uint32_t v = opnd(i, j);
            if (readers[v].empty() || readers[v].back() != i) readers[v].push_back(i);
        }
    std::vector<uint32_t> used_imm(n, 0u);
    for (uint32_t c = 0; c < n; ++c) {
        if (ins(c).op != NXG_OP_CONSTANT || c == p.result || !is_finite(ins(c).immediate)) continue;
        bool ok = true;
        for (uint32_t r : readers[c]) ok = ok && used_imm[r] < free_imm[r];
        if (!ok) continue;
        inline_const[c] = true;
        for (uint32_t r : readers[c]) ++used_imm[r];
    }
    for (uint32_t i = 0; i < n; ++i) {
        if (ins(i).op == NXG_OP_INPUT) {
            plan.dispatched[i] = false;
            if (!classic) plan.var_mask |= 1u << p.input_bindings[ins(i).auxiliary];
        } else if (inline_const[i]) {
            plan.dispatched[i] = false;
        }
    }
    // slots
    std::vector<int64_t> last_use(n, -1);
    for (uint32_t i = 0; i < n; ++i)
        for (uint32_t j = 0; j < ins(i).operand_count; ++j) last_use[opnd(i, j)] = i;
    last_use[p.result] = n;
    std::priority_queue<uint32_t, std::vector<uint32_t>, std::greater<uint32_t>> free_slots;
    std::vector<bool> released(n, false);
    uint32_t next = classic ? p.input_count : kVarSlots;
    for (uint32_t i = 0; i < n; ++i) {
        const nxg_expr_instruction &in = ins(i);
        if (in.op == NXG_OP_INPUT) {
            plan.slot_of[i] = static_cast<uint16_t>(classic ? in.auxiliary : p.input_bindings[in.auxiliary]);
            continue;
        }
        if (!plan.dispatched[i]) continue;
        for (uint32_t j = 0; j < in.operand_count; ++j) {
            const uint32_t v = opnd(i, j);
            if (last_use[v] == i && !released[v] && plan.dispatched[v]) {
                released[v] = true;
                free_slots.push(plan.slot_of[v]);
            }
        }
        uint32_t slot;
        if (free_slots.empty()) {
            slot = next++;
        } else {
            slot = free_slots.top();
            free_slots.pop();
        }
        plan.slot_of[i] = static_cast<uint16_t>(slot);
        if (last_use[i] < 0) free_slots.push(slot);  // never read: free after the write
    }
    plan.slot_count = next;
    // operand encodings and immediates
    for (uint32_t i = 0; i < n; ++i) {
        if (!plan.dispatched[i]) continue;
        ++plan.dispatched_count;
        plan.imm0[i] = ins(i).immediate;
        uint32_t k = ins(i).op == NXG_OP_RANDOM || ins(i).op == NXG_OP_CONSTANT ? 1u : 0u;
        std::vector<std::pair<uint32_t, uint16_t>> seen;  // constant -> encoding within this reader
        for (uint32_t j = 0; j < ins(i).operand_count; ++j) {
            const uint32_t v = opnd(i, j);
            uint16_t enc = plan.slot_of[v];
            if (inline_const[v]) {
                enc = 0xffffu;
                for (const auto &e : seen)
                    if (e.first == v) enc = e.second;
                if (enc == 0xffffu) {
                    enc = static_cast<uint16_t>(kImmediate | k);
                    (k == 0u ? plan.imm0[i] : plan.imm1[i]) = ins(v).immediate;
                    seen.push_back({v, enc});
                    ++k;
                }
            }
            plan.operand[i].push_back(enc);
        }
    }
    return plan;
}

// Writes one planned program at byte `at` of the image (its header, the
// dispatched instructions, operand slots and ramps); returns the offset
// after it.
uint64_t write_program(uint8_t *base, uint64_t at, const nxg_expr_program &p, const PackPlan &plan) {
    PackedProgram pp{};
    pp.instruction_count = plan.dispatched_count;
    pp.result = plan.slot_of[p.result];
    pp.input_count = p.input_count;
    pp.slot_count = plan.slot_count;
    pp.var_mask = plan.var_mask;
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
    uint32_t d = 0;
    for (uint32_t i = 0; i < p.instruction_count; ++i) {
        if (!plan.dispatched[i]) continue;
        const nxg_expr_instruction &in = p.instructions[i];
        PackedInstruction &o = ins[d++];
        o.op = in.op;
        o.operand_count = in.operand_count;
        o.dest = plan.slot_of[i];
        o.operand_offset = in.operand_offset;
        o.auxiliary = in.auxiliary;
        o.immediate = plan.imm0[i];
        o.immediate1 = plan.imm1[i];
        for (uint32_t j = 0; j < in.operand_count; ++j) {
            const uint16_t e = plan.operand[i][j];
            if (in.operand_count <= 4u) o.slot[j] = e;
            operand_slots[in.operand_offset + j] = e;
        }
    }
    if (p.ramp_count)
        std::memcpy(base + pp.ramps_offset, p.ramps, uint64_t{p.ramp_count} * sizeof(nxg_expr_ramp));
    if (p.ramp_point_count)
        std::memcpy(base + pp.ramp_points_offset, p.ramp_points,
                    uint64_t{p.ramp_point_count} * sizeof(nxg_expr_ramp_point));
    return cursor;
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
    std::vector<PackPlan> plans;
    plans.reserve(count);
    for (uint32_t k = 0; k < count; ++k) {
        if (!validate_program(programs[k])) return NXG_ERROR_INVALID_PROGRAM;
        for (uint32_t i = 0; i < programs[k].instruction_count; ++i) {
            const nxg_expr_instruction &in = programs[k].instructions[i];
            if (in.op == NXG_OP_TEXEL && in.auxiliary + 1u > maps_required) maps_required = in.auxiliary + 1u;
        }
        total += program_bytes(programs[k]);
        plans.push_back(plan_program(programs[k], false));
        if (plans.back().slot_count > max_values) max_values = plans.back().slot_count;
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
        const PackPlan &plan = plans[k];
        set.program_offset[k] = static_cast<uint32_t>(at);
        at = write_program(base, at, p, plan);
    }
    std::memcpy(base, &set, sizeof(set));
    info->program_count = set.program_count;
    info->max_values = set.max_values;
    info->texel_maps_required = maps_required;
    for (uint32_t k = 0; k < NXG_EXPR_MAX_PROGRAMS; ++k) info->program_offset[k] = set.program_offset[k];
    return NXG_OK;
}

uint32_t nxg_classic_plan_pack(const nxg_classic_plan *plan, void *out, uint64_t capacity,
                               nxg_classic_plan_info *info) {
    using namespace nxg;
    if (!plan || !info) return NXG_ERROR_NULL_ARGUMENT;
    if (plan->clump_count > NXG_CLASSIC_MAX_CLUMPS) return NXG_ERROR_INVALID_PROGRAM;
    nxg_classic_plan_info r{};
    uint32_t noises = 0u, cuts = 0u;
    auto push = [&](uint32_t type, uint32_t module) { r.effects[r.effect_count++] = nxg_classic_effect{type, module}; };
    if (plan->effect_count == 0u) {
        // the short form: clumps, then the NoiseFX and the CutFX before the last clumps_after_cut clumps
        if (plan->noise_count || plan->cut_count) return NXG_ERROR_INVALID_PROGRAM;
        if (plan->clumps_after_cut > plan->clump_count || (plan->clumps_after_cut && !plan->has_cut))
            return NXG_ERROR_INVALID_PROGRAM;
        if (plan->has_noise && plan->clumps_after_cut) return NXG_ERROR_INVALID_PROGRAM;
        const uint32_t before = plan->clump_count - plan->clumps_after_cut;
        for (uint32_t m = 0; m < before; ++m) push(NXG_CLASSIC_EFFECT_CLUMP, m);
        if (plan->has_noise) push(NXG_CLASSIC_EFFECT_NOISE, 0u);
        if (plan->has_cut) push(NXG_CLASSIC_EFFECT_CUT, 0u);
        for (uint32_t m = before; m < plan->clump_count; ++m) push(NXG_CLASSIC_EFFECT_CLUMP, m);
        noises = plan->has_noise ? 1u : 0u;
        cuts = plan->has_cut ? 1u : 0u;
        r.clumps_after_cut = plan->clumps_after_cut;
    } else {
        // an effect list naming every module exactly once
        if (plan->effect_count > NXG_CLASSIC_MAX_EFFECTS || plan->noise_count > NXG_CLASSIC_MAX_NOISES ||
            plan->cut_count > NXG_CLASSIC_MAX_CUTS || plan->has_cut || plan->has_noise || plan->clumps_after_cut)
            return NXG_ERROR_INVALID_PROGRAM;
        const uint32_t counts[3] = {plan->clump_count, plan->noise_count, plan->cut_count};
        uint32_t seen[3] = {0u, 0u, 0u};
        for (uint32_t e = 0; e < plan->effect_count; ++e) {
            const nxg_classic_effect &x = plan->effects[e];
            if (x.type > NXG_CLASSIC_EFFECT_CUT || x.module >= counts[x.type] || ((seen[x.type] >> x.module) & 1u))
                return NXG_ERROR_INVALID_PROGRAM;
            seen[x.type] |= 1u << x.module;
            push(x.type, x.module);
        }
        for (uint32_t t = 0; t < 3u; ++t)
            if (seen[t] != (1u << counts[t]) - 1u) return NXG_ERROR_INVALID_PROGRAM;
        noises = plan->noise_count;
        cuts = plan->cut_count;
    }
    // every module's programs
    for (uint32_t m = 0; m < plan->clump_count; ++m)
        for (uint32_t k = 0; k < NXG_CLASSIC_CLUMP_EXPRS; ++k)
            if (!plan->expressions[NXG_CLASSIC_EXPR_CLUMP_OF(m, NXG_CLASSIC_EXPR_CLUMP_MASK + k)])
                return NXG_ERROR_INVALID_PROGRAM;
    for (uint32_t n = 0; n < noises; ++n)
        for (uint32_t k = 0; k < NXG_CLASSIC_NOISE_EXPRS; ++k)
            if (!plan->expressions[NXG_CLASSIC_EXPR_NOISE_OF(n, NXG_CLASSIC_EXPR_NOISE_MASK + k)])
                return NXG_ERROR_INVALID_PROGRAM;
    for (uint32_t c = 0; c < cuts; ++c)
        if (!plan->expressions[NXG_CLASSIC_EXPR_CUT_AMOUNT_OF(c)]) return NXG_ERROR_INVALID_PROGRAM;
    r.has_cut = cuts ? 1u : 0u;
    r.cut_count = cuts;
    r.clump_count = plan->clump_count;
    r.has_noise = noises ? 1u : 0u;
    r.noise_count = noises;
    r.trailing_cut = r.effect_count && r.effects[r.effect_count - 1u].type == NXG_CLASSIC_EFFECT_CUT ? 1u : 0u;
    r.rewrites = r.effect_count > r.trailing_cut ? 1u : 0u;
    r.description_id = plan->description_id;
    r.description_name_seed = plan->description_name_seed;
    // the module and role of program k (role: its index in module 0's block), and whether the plan uses it
    auto role_of = [](uint32_t k, uint32_t &module) -> uint32_t {
        module = 0u;
        if (k >= NXG_CLASSIC_EXPR_CUT_AMOUNT_1) {
            module = k - NXG_CLASSIC_EXPR_CUT_AMOUNT_1 + 1u;
            return NXG_CLASSIC_EXPR_CUT_AMOUNT;
        }
        if (k >= NXG_CLASSIC_EXPR_NOISE_MASK) {
            module = (k - NXG_CLASSIC_EXPR_NOISE_MASK) / NXG_CLASSIC_NOISE_EXPRS;
            return k - NXG_CLASSIC_NOISE_EXPRS * module;
        }
        if (k >= NXG_CLASSIC_EXPR_CLUMP_MASK) {
            module = (k - NXG_CLASSIC_EXPR_CLUMP_MASK) / NXG_CLASSIC_CLUMP_EXPRS;
            return k - NXG_CLASSIC_CLUMP_EXPRS * module;
        }
        return k;
    };
    uint64_t total = 0u;
    PackPlan plans[NXG_CLASSIC_EXPR_COUNT];
    for (uint32_t k = 0; k < NXG_CLASSIC_EXPR_COUNT; ++k) {
        const nxg_expr_program *p = plan->expressions[k];
        uint32_t module;
        const uint32_t role = role_of(k, module);
        const bool noise = role >= NXG_CLASSIC_EXPR_NOISE_MASK;
        const bool clump = !noise && role >= NXG_CLASSIC_EXPR_CLUMP_MASK;
        const bool used = role == NXG_CLASSIC_EXPR_CUT_AMOUNT ? module < cuts
                          : noise                            ? module < noises
                          : clump                            ? module < plan->clump_count
                                                             : true;
        if (!p || !used) continue;
        if (!validate_classic_program(*p)) return NXG_ERROR_INVALID_PROGRAM;
        // the noise frequency and correlation are evaluated once per guide
        if (clump &&
            (role == NXG_CLASSIC_EXPR_CLUMP_NOISE_FREQUENCY || role == NXG_CLASSIC_EXPR_CLUMP_NOISE_CORRELATION) &&
            p->input_count != 0u)
            return NXG_ERROR_INVALID_PROGRAM;
        plans[k] = plan_program(*p, true);
        if (plans[k].slot_count > NXG_CLASSIC_MAX_PROGRAM_SLOTS) return NXG_ERROR_INVALID_PROGRAM;
        r.present |= uint64_t{1} << k;
        r.program_offset[k] = static_cast<uint32_t>(total);
        total += program_bytes(*p);
        for (uint32_t i = 0; i < p->input_count; ++i) {
            const uint32_t b = p->input_bindings[i];
            if (b >= NXG_CLASSIC_VAR_COLUMN0 && b - NXG_CLASSIC_VAR_COLUMN0 + 1u > r.columns_required)
                r.columns_required = b - NXG_CLASSIC_VAR_COLUMN0 + 1u;
            if (b == NXG_CLASSIC_VAR_CLENGTH) r.length_readers |= uint64_t{1} << k;
            if (b == NXG_CLASSIC_VAR_CLENGTH &&
                (k == NXG_CLASSIC_EXPR_TAPER || k == NXG_CLASSIC_EXPR_TAPER_START || k == NXG_CLASSIC_EXPR_WIDTH_RAMP))
                r.final_length = 1u;
            if (b == NXG_CLASSIC_VAR_CLENGTH && clump &&
                (role == NXG_CLASSIC_EXPR_CLUMP || role == NXG_CLASSIC_EXPR_CLUMP_SCALE))
                r.clump_length |= 1u << module;
        }
        if (p->random_call_count != 0u) r.needs_random_prefix = 1u;
        bool uniform = p->input_count == 0u;
        for (uint32_t i = 0; i < p->instruction_count; ++i)
            if (p->instructions[i].op == NXG_OP_RANDOM) uniform = false;
        // NoiseFX programs run per strand; the table holds one CutFX amount (module 0's)
        if (uniform && !noise && k < 32u) r.uniform |= 1u << k;
    }
    r.byte_size = total ? total : 16u;
    *info = r;
    if (!out) return NXG_OK;
    if (capacity < r.byte_size) return NXG_ERROR_BUFFER_TOO_SMALL;
    uint8_t *base = static_cast<uint8_t *>(out);
    std::memset(base, 0, r.byte_size);
    for (uint32_t k = 0; k < NXG_CLASSIC_EXPR_COUNT; ++k)
        if ((r.present >> k) & 1u) write_program(base, r.program_offset[k], *plan->expressions[k], plans[k]);
    return NXG_OK;
}

}  // extern "C"
