"""Float expression programs for per-root attributes.

Program is a builder for the SSA IR of include/nanoxgen_amd.h (the
reference's XgenFloatExpressionProgram, xgen_expression.h:83-100, plus the
pow/mix/noise3/texel extensions).  ExprSet packs up to four programs with
nxg_expr_pack and uploads the image; evaluate_root_attributes runs them on
the GPU for every root.

random_attribute_program builds the seeded 40-op programs of the
expression-attribute workload: ops drawn from add/mul/mix/clamp/pow/
smoothstep, four noise3 calls and two texel reads of a 16x16 per-face map,
with every intermediate value kept non-negative and finite, and the result
mapped into a stated attribute range.
"""
import ctypes as C

import numpy as np

from . import abi


class Program:
    def __init__(self):
        self.instructions = []
        self.operands = []
        self.bindings = []
        self.ramp_points = []
        self.ramps = []
        self.result = None
        self._struct = None
        self.op_count = 0  # instructions other than constant/input

    def _emit(self, op, args=(), aux=0, imm=0.0):
        for a in args:
            if not (0 <= a < len(self.instructions)):
                raise ValueError("operand must refer to an earlier value")
        self.instructions.append((abi.OP[op], len(self.operands), len(args), aux, float(imm)))
        self.operands.extend(args)
        if op not in ("constant", "input"):
            self.op_count += 1
        self._struct = None
        return len(self.instructions) - 1

    def const(self, v):
        return self._emit("constant", imm=v)

    def input(self, var):
        slot = abi.ROOT_VAR[var]
        if slot not in self.bindings:
            self.bindings.append(slot)
        return self._emit("input", aux=self.bindings.index(slot))

    def op(self, name, *args, aux=0, imm=0.0):
        return self._emit(name, args, aux=aux, imm=imm)

    def ramp(self, points):
        """points: [(position, value, interpolation)]"""
        self.ramps.append((len(self.ramp_points), len(points)))
        self.ramp_points.extend(points)
        return self._emit("ramp", aux=len(self.ramps) - 1)

    def finish(self, result):
        self.result = result
        self._struct = None
        return self

    @property
    def struct(self):
        if self._struct is None:
            ins = np.zeros(len(self.instructions), dtype=np.dtype([("op", "u1"), ("reserved", "u1"), ("count", "<u2"),
                                                                  ("offset", "<u4"), ("aux", "<u4"), ("imm", "<f4")]))
            for i, (op, off, cnt, aux, imm) in enumerate(self.instructions):
                ins[i] = (op, 0, cnt, off, aux, imm)
            self._ins = ins
            self._ops = np.ascontiguousarray(self.operands or [0], dtype=np.uint32)
            self._bind = np.ascontiguousarray(self.bindings or [0], dtype=np.uint32)
            rp = np.zeros(max(1, len(self.ramp_points)), dtype=np.dtype([("p", "<f4"), ("v", "<f4"), ("i", "<u4")]))
            for i, pt in enumerate(self.ramp_points):
                rp[i] = pt
            self._rp = rp
            self._ramps = np.ascontiguousarray(self.ramps or [(0, 0)], dtype=np.uint32).reshape(-1, 2)
            self._struct = abi.ExprProgram(
                instructions=self._ins.ctypes.data, operands=self._ops.ctypes.data,
                ramp_points=self._rp.ctypes.data, ramps=self._ramps.ctypes.data,
                input_bindings=self._bind.ctypes.data, instruction_count=len(self.instructions),
                operand_count=len(self.operands), ramp_point_count=len(self.ramp_points),
                ramp_count=len(self.ramps), input_count=len(self.bindings),
                result=self.result if self.result is not None else len(self.instructions) - 1,
                expression_seed=0, random_call_count=0)
        return self._struct


ATTRIBUTE_RANGES = {"length": (0.6, 1.4), "width": (0.5, 1.5), "noise": (0.3, 1.7), "clump": (0.0, 1.0)}


def random_attribute_program(seed, lo, hi, n_ops=40, noise_calls=4, texel_reads=2, position_extent=1.1,
                             frequency_scale=1.0):
    """A seeded program of exactly n_ops ops (noise3 and texel calls and the
    arithmetic that feeds them included); constants and inputs are not ops.
    Every value stays in [0, 2]; the result is mix(lo, hi, clamp(x, 0, 1)).
    frequency_scale multiplies the noise3 argument scales (drawn from
    U(0.8, 3.0)); at 1.0 the program is unchanged bit for bit."""
    rng = np.random.default_rng(seed)
    p = Program()
    pool = []  # (value index, max)
    # texel reads at barycentric coordinates of the root's face
    face, b1, b2 = p.input("face"), p.input("bary1"), p.input("bary2")
    pool.append((p.op("texel", face, b1, b2, aux=0), 1.0))
    if texel_reads > 1:
        pool.append((p.op("texel", face, b2, b1, aux=0), 1.0))
    # noise3 over scaled inputs (3 muls + 1 call each)
    sources = ["position_x", "position_y", "position_z", "normal_x", "normal_y", "normal_z"]
    for _ in range(noise_calls):
        args = []
        for axis in range(3):
            src = p.input(sources[axis] if rng.random() < 0.7 else sources[rng.integers(0, 6)])
            args.append(p.op("multiply", src, p.const(rng.uniform(0.8, 3.0) * frequency_scale)))
        pool.append((p.op("noise3", *args), 1.0))
    # texcoords and normalised positions as further [0,1] leaves
    for name in ("u", "v"):
        pool.append((p.input(name), 1.0))
    remaining = n_ops - p.op_count - 2  # the final clamp + mix
    while remaining > 0:
        kind = rng.choice(["add", "mul", "mix", "clamp", "pow", "smoothstep"])
        a, am = pool[rng.integers(0, len(pool))]
        b, bm = pool[rng.integers(0, len(pool))]
        if kind == "add":
            if am + bm > 2.0:
                continue
            v, m = p.op("add", a, b), am + bm
        elif kind == "mul":
            v, m = p.op("multiply", a, b), am * bm
        elif kind == "mix":
            t, tm = pool[rng.integers(0, len(pool))]
            if tm > 1.0:
                continue
            v, m = p.op("mix", a, b, t), max(am, bm)
        elif kind == "clamp":
            v, m = p.op("clamp", a, p.const(0.0), p.const(1.0)), min(am, 1.0)
        elif kind == "pow":
            e = rng.uniform(0.5, 2.0)
            v, m = p.op("pow", a, p.const(e)), am ** e
        else:
            lo_e = rng.uniform(0.0, 0.4)
            v, m = p.op("smoothstep", a, p.const(lo_e), p.const(lo_e + rng.uniform(0.3, 0.8))), 1.0
        pool.append((v, m))
        remaining -= 1
    x = p.op("clamp", pool[-1][0], p.const(0.0), p.const(1.0))
    return p.finish(p.op("mix", p.const(lo), p.const(hi), x))


def attribute_programs(seed=3, n_ops=40, clump=False, frequency_scale=1.0):
    """The per-root length, width and noise-magnitude programs (and, with
    clump=True, the clump weight) of the expression-attribute workloads."""
    names = ["length", "width", "noise"] + (["clump"] if clump else [])
    return [random_attribute_program(seed * 1000 + k, *ATTRIBUTE_RANGES[n], n_ops=n_ops,
                                     frequency_scale=frequency_scale) for k, n in enumerate(names)]


def texel_map(faces, resolution=16, seed=4):
    """Synthetic per-face texel map with U(0,1) values, [faces, R, R]."""
    rng = np.random.default_rng(seed)
    return rng.uniform(0.0, 1.0, size=(faces, resolution, resolution)).astype(np.float32)


class ExprSet:
    """Packed programs uploaded to the device (nxg_expr_pack + copy)."""

    def __init__(self, programs, device="cuda"):
        import torch
        from . import lib, check
        L = lib()
        structs = (abi.ExprProgram * len(programs))(*[p.struct for p in programs])
        info = abi.ExprSet()
        check(L.nxg_expr_pack(structs, len(programs), None, 0, C.byref(info)), "nxg_expr_pack")
        host = np.zeros(info.byte_size, dtype=np.uint8)
        check(L.nxg_expr_pack(structs, len(programs), host.ctypes.data, host.nbytes, C.byref(info)), "nxg_expr_pack")
        self.host = host
        self.programs = programs
        self.device = torch.from_numpy(host).to(device)
        info.image = self.device.data_ptr()
        self.info = info


class TexelMaps:
    def __init__(self, maps, device="cuda"):
        import torch
        self.tensors = [torch.from_numpy(np.ascontiguousarray(m, dtype=np.float32)).to(device) for m in maps]
        self.structs = (abi.TexelMap * max(1, len(maps)))(
            *[abi.TexelMap(t.data_ptr(), t.shape[0], t.shape[1]) for t in self.tensors])
        self.count = len(maps)


def evaluate_root_attributes(asset, params, exprs, maps, attributes, error_word=None, stream=None):
    """Enqueue nxg_evaluate_root_attributes; attributes is a float32 [N, 4]
    device tensor, error_word an optional int32 [1] device tensor."""
    from . import lib, check, _stream_handle
    check(lib().nxg_evaluate_root_attributes(
        C.byref(asset.desc), C.byref(params), C.byref(exprs.info), maps.structs if maps else None,
        maps.count if maps else 0, C.c_void_p(attributes.data_ptr()), attributes.numel(),
        C.c_void_p(error_word.data_ptr()) if error_word is not None else None, _stream_handle(stream)),
        "nxg_evaluate_root_attributes")


_INSTR = np.dtype([("op", "u1"), ("reserved", "u1"), ("count", "<u2"), ("offset", "<u4"), ("aux", "<u4"),
                   ("imm", "<f4")])
_RPT = np.dtype([("p", "<f4"), ("v", "<f4"), ("i", "<u4")])

# Default binding of expression $-variables to per-root variables.
DEFAULT_BINDINGS = {"u": "u", "v": "v", "id": "id", "faceId": "face", "Px": "position_x", "Py": "position_y",
                    "Pz": "position_z", "Nx": "normal_x", "Ny": "normal_y", "Nz": "normal_z"}


class CompiledProgram:
    """A program compiled from expression text (nxg_expr_compile), usable
    wherever a Program is (ExprSet, oracle.expr_eval)."""

    def __init__(self, arrays, counts, inputs, result, expression_seed, random_call_count):
        # arrays hold at least one element each so their pointers are valid;
        # counts are the lengths the program actually has.
        self.ins, self.ops, self.pts, self.ramps_arr = arrays
        self.counts = counts
        self.inputs = inputs
        self.result = result
        self.expression_seed = expression_seed
        self.random_call_count = random_call_count
        self.bindings = list(range(len(inputs)))
        self._struct = None

    @property
    def instructions(self):
        """(op, operand_offset, operand_count, auxiliary, immediate) per instruction, as Program holds them."""
        return [(int(i["op"]), int(i["offset"]), int(i["count"]), int(i["aux"]), float(i["imm"]))
                for i in self.ins[:self.counts[0]]]

    def bind(self, bindings):
        """Program input name -> root variable name (abi.ROOT_VARS)."""
        missing = [n for n in self.inputs if n not in bindings]
        if missing:
            raise KeyError(f"no root variable bound to ${', $'.join(missing)}")
        self.bindings = [abi.ROOT_VAR[bindings[n]] for n in self.inputs]
        self._struct = None
        return self

    def bind_indices(self, indices):
        """Input i reads variable indices[i] of the caller's binding space
        (e.g. NXG_CLASSIC_VAR_* for a Classic plan)."""
        if len(indices) != len(self.inputs):
            raise ValueError(f"{len(self.inputs)} inputs, {len(indices)} bindings")
        self.bindings = [int(i) for i in indices]
        self._struct = None
        return self

    @property
    def struct(self):
        if self._struct is None:
            self._bind = np.ascontiguousarray(self.bindings or [0], dtype=np.uint32)
            self._struct = abi.ExprProgram(
                instructions=self.ins.ctypes.data, operands=self.ops.ctypes.data, ramp_points=self.pts.ctypes.data,
                ramps=self.ramps_arr.ctypes.data, input_bindings=self._bind.ctypes.data,
                instruction_count=self.counts[0], operand_count=self.counts[1], ramp_point_count=self.counts[2],
                ramp_count=self.counts[3], input_count=len(self.inputs), result=self.result,
                expression_seed=self.expression_seed, random_call_count=self.random_call_count)
        return self._struct


def compile_expression(source, name="", object_type="", *, extensions=False, limits=(0, 0, 0, 0, 0)):
    """compile_xgen_scalar_expression + make_xgen_float_expression_program
    (src/xgen_expression.cpp:771-818) through nxg_expr_compile.  Raises
    ValueError with the reference's diagnostic."""
    from . import lib
    src = source.encode()
    opt = abi.ExprCompileOptions(expression_name=name.encode(), object_type=object_type.encode(),
                                 max_source_bytes=limits[0], max_tokens=limits[1], max_instructions=limits[2],
                                 max_locals=limits[3], max_ramp_points=limits[4], extensions=1 if extensions else 0)
    out = abi.ExprCompiled()
    err = C.create_string_buffer(512)
    st = lib().nxg_expr_compile(src, len(src), C.byref(opt), C.byref(out), err, 512)
    if st == abi.INVALID_PROGRAM:
        raise ValueError(err.value.decode())
    ins = np.zeros(max(1, out.instruction_count), _INSTR)
    ops = np.zeros(max(1, out.operand_count), np.uint32)
    pts = np.zeros(max(1, out.ramp_point_count), _RPT)
    ramps = np.zeros((max(1, out.ramp_count), 2), np.uint32)
    names = C.create_string_buffer(max(1, out.input_name_bytes))
    out = abi.ExprCompiled(instructions=ins.ctypes.data, operands=ops.ctypes.data, ramp_points=pts.ctypes.data,
                           ramps=ramps.ctypes.data, input_names=C.cast(names, C.c_void_p),
                           instruction_capacity=len(ins), operand_capacity=len(ops), ramp_point_capacity=len(pts),
                           ramp_capacity=len(ramps), input_name_capacity=len(names))
    st = lib().nxg_expr_compile(src, len(src), C.byref(opt), C.byref(out), err, 512)
    if st != 0:
        raise ValueError(err.value.decode() or lib().nxg_status_message(st).decode())
    inputs = [x.decode() for x in names.raw[:out.input_name_bytes].split(b"\0")[:-1]] if out.input_count else []
    counts = (out.instruction_count, out.operand_count, out.ramp_point_count, out.ramp_count)
    arrays = (ins[:max(1, counts[0])], ops[:max(1, counts[1])], pts[:max(1, counts[2])], ramps[:max(1, counts[3])])
    return CompiledProgram(arrays, counts, inputs, out.result, out.expression_seed, out.random_call_count)
