"""Classic CutFX and primitive expressions on device curves (host side of
nxg_classic_cut_apply and nxg_classic_runtime_apply).

The reference runs a Classic description's CutFX module inside its float
runtime (apply_xgen_classic_float_runtime_plan_cpu,
include/nanoxgen/xgen_classic_runtime.h:185-190): per strand cLength, the
width/length culling, cut_from_tip_and_rebuild with the module's amount
expression, then an in-order compaction of the kept strands
(src/xgen_classic_runtime.cpp:1648-1678, 1806-1821, 1893-1918).  Here the
amount is amount_scale * cLength + amount_offset, optionally times a
per-strand multiplier (e.g. one column of the attribute runtime's output),
and the whole chain runs on the GPU into separate output buffers.

ClassicPlan / classic_runtime cover the same function with a plan of
SplinePrimitive expressions (length, width, taper, taperStart, widthRamp)
and a CutFX amount expression, compiled from their text as the reference's
plan compiler does (compile_xgen_classic_float_runtime_plan,
src/xgen_classic_runtime.cpp:1189-1237) and bound to the Classic variables
($id, $cLength, $cWidth and the per-strand input row).
"""
import ctypes as C

import numpy as np

from . import abi, check, lib, _stream_handle, _torch
from . import expr as _expr


class ClassicCutOutputs:
    """Compacted outputs sized for every input strand, plus the device
    kept-count and status words."""

    def __init__(self, strands, cvs, device="cuda", point_counts=True, roots=True, root_uvs=True):
        torch = _torch()
        self.strands, self.cvs = strands, cvs
        self.points = torch.empty((max(1, strands * cvs), 4), dtype=torch.float32, device=device)
        self.point_counts = torch.empty(max(1, strands), dtype=torch.int32, device=device) if point_counts else None
        self.roots = torch.empty((max(1, strands), 12), dtype=torch.int32, device=device) if roots else None
        self.root_uvs = torch.empty((max(1, strands), 2), dtype=torch.float32, device=device) if root_uvs else None
        self.words = torch.zeros(64, dtype=torch.int32, device=device)  # [0] kept count, [32] status

    def descriptor(self):
        t = lambda x: None if x is None else x.data_ptr()
        n = lambda x, k: 0 if x is None else self.strands * k
        return abi.ClassicOutput(points=self.points.data_ptr(), point_counts=t(self.point_counts),
                                 roots=t(self.roots), root_uvs=t(self.root_uvs),
                                 point_capacity=self.strands * self.cvs, point_count_capacity=n(self.point_counts, 1),
                                 root_capacity=n(self.roots, 1), root_uv_capacity=n(self.root_uvs, 1),
                                 kept_count=self.words.data_ptr(), status=self.words.data_ptr() + 128)

    def kept(self):
        """Kept strand count (synchronises)."""
        return int(self.words[0].item())

    def status(self):
        """0, or the NXG_CLASSIC_NONFINITE_* bits of the call (synchronises)."""
        return int(self.words[32].item())

    def to_host(self):
        """The kept strands as numpy arrays (synchronises)."""
        k = self.kept()
        out = dict(strand_count=k, status=self.status(), points=self.points[:k * self.cvs].cpu().numpy())
        if self.point_counts is not None:
            out["point_counts"] = self.point_counts[:k].cpu().numpy().view(np.uint32)
        if self.roots is not None:
            out["roots"] = self.roots[:k].cpu().numpy().view(abi.ROOT_DTYPE).reshape(-1)
        if self.root_uvs is not None:
            out["root_uvs"] = self.root_uvs[:k].cpu().numpy()
        return out


def workspace_bytes(strands):
    return int(lib().nxg_classic_cut_workspace_size(strands))


def classic_cut(points, cvs, amount_scale, amount_offset, *, point_counts=None, roots=None, root_uvs=None,
                multiplier=None, multiplier_stride=1, radius_scale=1.0, outputs=None, workspace=None,
                stream=None):
    """Enqueue the CutFX chain on `stream` over device curves: `points`
    float32 [N * cvs, 4]; optional `point_counts` int32 [N], `roots` int32
    [N, 12] (RootSample records), `root_uvs` float32 [N, 2] and `multiplier`
    float32 (strand s reads multiplier.flatten()[s * multiplier_stride]).
    Returns the ClassicCutOutputs (its kept() and to_host() synchronise)."""
    torch = _torch()
    n = points.shape[0] // cvs
    if outputs is None:
        outputs = ClassicCutOutputs(n, cvs, device=points.device, point_counts=True, roots=roots is not None,
                                    root_uvs=root_uvs is not None)
    if workspace is None:
        workspace = torch.empty(workspace_bytes(n), dtype=torch.uint8, device=points.device)
    if points.dtype != torch.float32 or points.dim() != 2 or points.shape[1] != 4 or not points.is_contiguous():
        raise ValueError("points must be a contiguous float32 [N * cvs, 4] tensor")
    if multiplier is not None:
        reach = (n - 1) * multiplier_stride + 1 if n else 0
        have = multiplier.untyped_storage().nbytes() // 4 - multiplier.storage_offset()
        if multiplier.dtype != torch.float32 or multiplier_stride < 1 or have < reach:
            raise ValueError(f"multiplier must be float32 with {reach} elements from its first at stride "
                             f"{multiplier_stride}")
    for name, x, k in (("point_counts", point_counts, 1), ("roots", roots, 12), ("root_uvs", root_uvs, 2)):
        if x is not None and (x.numel() < n * k or not x.is_contiguous()):
            raise ValueError(f"{name} must be contiguous with {n * k} elements")
    t = lambda x: None if x is None else x.data_ptr()
    curves = abi.ClassicCurves(points=points.data_ptr(), point_counts=t(point_counts), roots=t(roots),
                               root_uvs=t(root_uvs), strand_count=n, cvs_per_strand=cvs)
    cut = abi.ClassicCutParams(amount_scale=amount_scale, amount_offset=amount_offset, radius_scale=radius_scale,
                               multiplier_stride=multiplier_stride if multiplier is not None else 0,
                               multiplier=t(multiplier))
    out = outputs.descriptor()
    check(lib().nxg_classic_cut_apply(C.byref(curves), C.byref(cut), C.byref(out),
                                      C.c_void_p(workspace.data_ptr()), workspace.numel(), _stream_handle(stream)),
          "nxg_classic_cut_apply")
    return outputs


def add_endpoints(points, cvs, *, strands=None, device_strand_count=None, out=None, point_counts=None, stream=None):
    """Renderer endpoints on device curves (nxg_classic_add_endpoints):
    `points` float32 [N * cvs, 4] -> `out` float32 [N * (cvs + 2), 4].
    `device_strand_count` (an int32 device tensor, e.g. ClassicCutOutputs.words)
    bounds the strands on the device without synchronising."""
    torch = _torch()
    if points.dtype != torch.float32 or points.dim() != 2 or points.shape[1] != 4 or not points.is_contiguous():
        raise ValueError("points must be a contiguous float32 [N * cvs, 4] tensor")
    n = points.shape[0] // cvs if strands is None else strands
    if out is None:
        out = torch.empty((max(1, n * (cvs + 2)), 4), dtype=torch.float32, device=points.device)
    t = lambda x: None if x is None else x.data_ptr()
    check(lib().nxg_classic_add_endpoints(points.data_ptr(), n, cvs, t(device_strand_count), out.data_ptr(),
                                          out.shape[0] if n else 0, t(point_counts),
                                          0 if point_counts is None else point_counts.numel(),
                                          _stream_handle(stream)), "nxg_classic_add_endpoints")
    return out


def clump_workspace_bytes(guide_count, cvs):
    return int(lib().nxg_classic_clump_workspace_size(guide_count, cvs))


def clump(points, cvs, strand_guides, guides, params, *, amount_multiplier=None, amount_stride=1, out=None,
          status=None, workspace=None, stream=None):
    """ClumpingFX with constant module expressions (nxg_classic_clump_apply)
    on device curves: `points` float32 [N * cvs, 4]; `strand_guides` int32 [N]
    (-1: no guide); `guides` dict of device float32 tensors axes [G * cvs, 3],
    normals [G, 3], tangents [G, 3] and optionally reference_positions [G, 3];
    `params` dict with mask, amount, clump_scale, noise, noise_scale,
    noise_frequency, noise_correlation.  `out` defaults to a new tensor; pass
    `points` itself to clump in place; `status` (int32 device tensor) receives
    the NXG_CLASSIC_* error bits."""
    torch = _torch()
    if points.dtype != torch.float32 or points.dim() != 2 or points.shape[1] != 4 or not points.is_contiguous():
        raise ValueError("points must be a contiguous float32 [N * cvs, 4] tensor")
    n = points.shape[0] // cvs
    axes = guides["axes"]
    g = axes.shape[0] // cvs
    for name in ("axes", "normals", "tangents", "reference_positions"):
        x = guides.get(name)
        if x is not None and (x.dtype != torch.float32 or not x.is_contiguous() or x.shape[-1] != 3):
            raise ValueError(f"guides[{name!r}] must be a contiguous float32 [..., 3] tensor")
    if strand_guides.numel() < n or not strand_guides.is_contiguous():
        raise ValueError("strand_guides must hold one index per strand")
    if amount_multiplier is not None:
        reach = (n - 1) * amount_stride + 1 if n else 0
        if amount_multiplier.untyped_storage().nbytes() // 4 - amount_multiplier.storage_offset() < reach:
            raise ValueError("amount_multiplier is too short for its stride")
    if out is None:
        out = torch.empty_like(points)
    if workspace is None:
        workspace = torch.empty(clump_workspace_bytes(g, cvs), dtype=torch.uint8, device=points.device)
    t = lambda x: None if x is None else x.data_ptr()
    gd = abi.ClumpGuides(axes=axes.data_ptr(), normals=guides["normals"].data_ptr(),
                         tangents=guides["tangents"].data_ptr(),
                         reference_positions=t(guides.get("reference_positions")), guide_count=g)
    pr = abi.ClumpParams(amount_stride=amount_stride if amount_multiplier is not None else 0,
                         amount_multiplier=t(amount_multiplier),
                         **{k: float(params[k]) for k in ("mask", "amount", "clump_scale", "noise", "noise_scale",
                                                          "noise_frequency", "noise_correlation")})
    check(lib().nxg_classic_clump_apply(points.data_ptr(), n, cvs, strand_guides.data_ptr(), C.byref(gd),
                                        C.byref(pr), out.data_ptr(), out.shape[0], t(status),
                                        C.c_void_p(workspace.data_ptr()), workspace.numel(),
                                        _stream_handle(stream)), "nxg_classic_clump_apply")
    return out


# plan slot -> (attribute, object type): the compile options the reference
# compiles each expression under (expression_name, object_type), which seed
# rand() call sites (src/xgen_classic_runtime.cpp:420-426)
CLASSIC_ATTRIBUTES = {
    "length": ("length", "SplinePrimitive"),
    "width": ("width", "SplinePrimitive"),
    "cut_amount": ("amount", "CutFXModule"),
    "taper": ("taper", "SplinePrimitive"),
    "taper_start": ("taperStart", "SplinePrimitive"),
    "width_ramp": ("widthRamp", "SplinePrimitive"),
    "clump_mask": ("mask", "ClumpingFXModule"),
    "clump": ("clump", "ClumpingFXModule"),
    "clump_scale": ("clumpScale", "ClumpingFXModule"),
    "clump_noise": ("noise", "ClumpingFXModule"),
    "clump_noise_scale": ("noiseScale", "ClumpingFXModule"),
    "clump_noise_frequency": ("noiseFrequency", "ClumpingFXModule"),
    "clump_noise_correlation": ("noiseCorrelation", "ClumpingFXModule"),
    "noise_mask": ("mask", "NoiseFXModule"),
    "noise_magnitude": ("magnitude", "NoiseFXModule"),
    "noise_magnitude_scale": ("magnitudeScale", "NoiseFXModule"),
    "noise_frequency": ("frequency", "NoiseFXModule"),
    "noise_correlation": ("correlation", "NoiseFXModule"),
    "noise_preserve_length": ("preserveLength", "NoiseFXModule"),
}
# ClassicPlan(noise=...) keys, the NoiseFX module's attribute names
NOISE_KEYS = {"mask": "noise_mask", "magnitude": "noise_magnitude", "magnitudeScale": "noise_magnitude_scale",
              "frequency": "noise_frequency", "correlation": "noise_correlation",
              "preserveLength": "noise_preserve_length"}
# ClassicPlan(clump=...) keys, the module's attribute names
CLUMP_KEYS = {"mask": "clump_mask", "clump": "clump", "clumpScale": "clump_scale", "noise": "clump_noise",
              "noiseScale": "clump_noise_scale", "noiseFrequency": "clump_noise_frequency",
              "noiseCorrelation": "clump_noise_correlation"}
_CLASSIC_INPUT_PREFIXES = ("__nxg_map_", "__nxg_custom_", "__nxg_pref_noise_")


def classic_binding(name, ptex_count=0, custom_count=0):
    """The NXG_CLASSIC_VAR_* index of a Classic runtime variable, as the
    reference's input_value binds it (src/xgen_classic_runtime.cpp:504-553):
    $id, $cLength, $cWidth, then the input row [PTEX, custom, $Prefg noise]."""
    fixed = {"id": abi.CLASSIC_VAR_ID, "cLength": abi.CLASSIC_VAR_CLENGTH, "cWidth": abi.CLASSIC_VAR_CWIDTH}
    if name in fixed:
        return fixed[name]
    bases = (0, ptex_count, ptex_count + custom_count)
    for prefix, base in zip(_CLASSIC_INPUT_PREFIXES, bases):
        if name.startswith(prefix) and name[len(prefix):].isdigit():
            return abi.CLASSIC_VAR_COLUMN0 + base + int(name[len(prefix):])
    raise KeyError(f"Classic runtime variable is not bound: ${name}")


class ClassicPlan:
    """A ClassicFloatRuntimePlan (include/nanoxgen/xgen_classic_runtime.h:
    130-165) of SplinePrimitive expressions and FX modules, packed for the
    device (nxg_classic_plan_pack).

    `effects` is the plan's effect list in object order (the reference's
    plan.effects): ("clump", dict of the module's seven expressions by
    attribute name, mask ... noiseCorrelation), ("noise", dict of the six
    NoiseFX expressions, mask ... preserveLength) or ("cut", the CutFX amount
    expression), up to three clumps, two NoiseFX and two CutFX.  Without it
    the short form applies: `clump` (one module's dict or a list of them),
    an optional `noise` module and an optional `cut_amount`, the NoiseFX and
    the CutFX after all clumps but the last `clumps_after_cut`.

    Each expression is text (compiled here with its attribute name and
    object type) or a CompiledProgram; ptex_count and custom_count place the
    custom and $Prefg-noise inputs in the input row."""

    def __init__(self, *, length=None, width=None, taper=None, taper_start=None, width_ramp=None,
                 cut_amount=None, clump=None, description_id=0, description_name="", ptex_count=0,
                 custom_count=0, clumps_after_cut=0, noise=None, effects=None):
        given = dict(length=length, width=width, taper=taper, taper_start=taper_start, width_ramp=width_ramp)
        if effects is not None:
            if clump is not None or noise is not None or cut_amount is not None or clumps_after_cut:
                raise ValueError("effects= replaces clump=, noise=, cut_amount= and clumps_after_cut=")
            modules = dict(clump=[], noise=[], cut=[])
            order = []
            for kind, module in effects:
                if kind not in modules:
                    raise ValueError(f"unknown effect type {kind!r}")
                order.append((kind, len(modules[kind])))
                modules[kind].append(module)
        else:
            clumps = [] if clump is None else [clump] if isinstance(clump, dict) else list(clump)
            before = len(clumps) - clumps_after_cut
            order = ([("clump", m) for m in range(before)] + ([("noise", 0)] if noise is not None else [])
                     + ([("cut", 0)] if cut_amount is not None else [])
                     + [("clump", m) for m in range(before, len(clumps))])
            modules = dict(clump=clumps, noise=[] if noise is None else [noise],
                           cut=[] if cut_amount is None else [cut_amount])
        limits = dict(clump=abi.CLASSIC_MAX_CLUMPS, noise=abi.CLASSIC_MAX_NOISES, cut=abi.CLASSIC_MAX_CUTS)
        for kind, limit in limits.items():
            if len(modules[kind]) > limit:
                raise ValueError(f"at most {limit} {kind} modules")
        for m, module in enumerate(modules["clump"]):
            missing = set(CLUMP_KEYS) - set(module)
            if missing:
                raise ValueError(f"ClumpingFX module {m} needs {sorted(missing)}")
            given.update({abi.clump_key(m, CLUMP_KEYS[k]): v for k, v in module.items()})
        for n, module in enumerate(modules["noise"]):
            missing = set(NOISE_KEYS) - set(module)
            if missing:
                raise ValueError(f"NoiseFX module {n} needs {sorted(missing)}")
            given.update({abi.noise_key(n, NOISE_KEYS[k]): v for k, v in module.items()})
        for c, amount in enumerate(modules["cut"]):
            given[abi.cut_key(c)] = amount
        self.programs = {}
        for key in abi.CLASSIC_EXPRS:
            src = given.get(key)
            if src is None:
                continue
            attribute, object_type = CLASSIC_ATTRIBUTES[key.split("@")[0]]
            prog = src if isinstance(src, _expr.CompiledProgram) else _expr.compile_expression(
                src, attribute, object_type)
            self.programs[key] = prog.bind_indices(
                [classic_binding(name, ptex_count, custom_count) for name in prog.inputs])
        self.effects = order  # [(kind, module)] in application order
        self.clump_count = len(modules["clump"])
        self.noise_count = len(modules["noise"])
        self.cut_count = len(modules["cut"])
        self.has_cut = self.cut_count > 0
        self.has_noise = self.noise_count > 0
        self.clumps_after_cut = clumps_after_cut
        self.description_id = description_id
        self.name_seed = int(lib().nxg_string_seed(description_name.rsplit(":", 1)[-1].encode()))
        if effects is not None:
            self.struct = abi.ClassicPlan(clump_count=self.clump_count, noise_count=self.noise_count,
                                          cut_count=self.cut_count, effect_count=len(order),
                                          description_id=description_id, description_name_seed=self.name_seed)
            for e, (kind, module) in enumerate(order):
                self.struct.effects[e] = abi.ClassicEffect(abi.CLASSIC_EFFECT_TYPES[kind], module)
        else:
            self.struct = abi.ClassicPlan(has_cut=1 if self.has_cut else 0, clump_count=self.clump_count,
                                          clumps_after_cut=clumps_after_cut, has_noise=1 if self.has_noise else 0,
                                          description_id=description_id, description_name_seed=self.name_seed)
        for k, key in enumerate(abi.CLASSIC_EXPRS):
            if key in self.programs:
                self.struct.expressions[k] = C.addressof(self.programs[key].struct)
        self.info = abi.ClassicPlanInfo()
        check(lib().nxg_classic_plan_pack(C.byref(self.struct), None, 0, C.byref(self.info)), "nxg_classic_plan_pack")
        self.image = np.zeros(self.info.byte_size, np.uint8)
        check(lib().nxg_classic_plan_pack(C.byref(self.struct), self.image.ctypes.data, self.image.nbytes,
                                          C.byref(self.info)), "nxg_classic_plan_pack")
        self._device = {}

    def device_image(self, device):
        """The packed image in device memory (uploaded once per device)."""
        key = str(device)
        if key not in self._device:
            self._device[key] = _torch().from_numpy(self.image).to(device)
        return self._device[key]


def runtime_workspace_bytes(strands, cvs=2, guides=0, plan=None):
    info = None if plan is None else C.byref(plan.info)
    return int(lib().nxg_classic_runtime_workspace_size(info, strands, cvs, guides))


def classic_runtime(points, cvs, plan, *, roots, point_counts=None, root_uvs=None, primitive_ids=None,
                    random_prefixes=None, columns=None, radius_scale=1.0, strand_guides=None, guides=None,
                    guide_uvs=None, guide_random_prefixes=None, clumps=None, surface_tangents=None,
                    noise_domain_positions=None, root_relative=False, outputs=None, workspace=None, stream=None):
    """Enqueue the Classic runtime chain of `plan` (a ClassicPlan) on
    `stream` over device curves: `points` float32 [N * cvs, 4], `roots` int32
    [N, 12] (RootSample records: the context's u, v and face); optional
    `point_counts` int32 [N], `root_uvs` float32 [N, 2], `primitive_ids`
    int32 [N] ($id), `random_prefixes` int32 [N] (SeExpr prefixes, required
    when a program calls rand()) and `columns` float32 [N, K] (the input
    row).  Each ClumpingFX module of the plan takes a binding: `clumps`, a
    list of dicts with `strand_guides` int32 [N], `guides` (dict of float32
    device tensors axes [G * cvs, 3], normals and tangents [G, 3], optional
    reference [G, 3]), `guide_uvs` float32 [G, 2] and `guide_random_prefixes`
    int32 [G], optionally `local_axes` float32 [G * cvs, 3] (translation-free
    guide CVs); for a single module the first four may be passed as keywords
    instead.  A NoiseFX takes `surface_tangents` float32 [N, 3] and optional
    `noise_domain_positions` float32 [N, 3].  `root_relative` takes clump
    goals about each strand's root.  Returns the ClassicCutOutputs (its
    kept(), status() and to_host() synchronise)."""
    torch = _torch()
    if points.dtype != torch.float32 or points.dim() != 2 or points.shape[1] != 4 or not points.is_contiguous():
        raise ValueError("points must be a contiguous float32 [N * cvs, 4] tensor")
    n = points.shape[0] // cvs
    if clumps is None and guides is not None:
        clumps = [dict(strand_guides=strand_guides, guides=guides, guide_uvs=guide_uvs,
                       guide_random_prefixes=guide_random_prefixes)]
    clumps = list(clumps or [])
    t = lambda x: None if x is None else x.data_ptr()
    guide_count = 0
    keep = []  # the ctypes structs the call points at
    bindings = (abi.ClassicClumpBinding * max(len(clumps), 1))()
    for m, binding in enumerate(clumps):
        gs = binding["guides"]
        g = gs["normals"].shape[0]
        for name, x, k in (("axes", gs["axes"], g * cvs * 3), ("normals", gs["normals"], g * 3),
                           ("tangents", gs["tangents"], g * 3), ("reference", gs.get("reference"), g * 3),
                           ("guide_uvs", binding["guide_uvs"], g * 2),
                           ("guide_random_prefixes", binding["guide_random_prefixes"], g),
                           ("strand_guides", binding["strand_guides"], n),
                           ("local_axes", binding.get("local_axes"), g * cvs * 3)):
            if x is not None and (x.numel() < k or not x.is_contiguous() or x.element_size() != 4):
                raise ValueError(f"clump {m}: {name} must be contiguous 32-bit with {k} elements")
        gd = abi.ClumpGuides(axes=gs["axes"].data_ptr(), normals=gs["normals"].data_ptr(),
                             tangents=gs["tangents"].data_ptr(), reference_positions=t(gs.get("reference")),
                             guide_count=g)
        keep.append(gd)
        bindings[m] = abi.ClassicClumpBinding(strand_guides=t(binding["strand_guides"]), guides=C.addressof(gd),
                                              guide_uvs=t(binding["guide_uvs"]),
                                              guide_random_prefixes=t(binding["guide_random_prefixes"]),
                                              local_axes=t(binding.get("local_axes")))
        if m < plan.clump_count:
            guide_count += g
    if outputs is None:
        outputs = ClassicCutOutputs(n, cvs, device=points.device, point_counts=True, roots=True,
                                    root_uvs=root_uvs is not None)
    if workspace is None:
        workspace = torch.empty(runtime_workspace_bytes(n, cvs, guide_count, plan),
                                dtype=torch.uint8, device=points.device)
    column_count = 0 if columns is None else (columns.shape[1] if columns.dim() == 2 else 0)
    for name, x, k in (("roots", roots, 12), ("point_counts", point_counts, 1), ("root_uvs", root_uvs, 2),
                       ("primitive_ids", primitive_ids, 1), ("random_prefixes", random_prefixes, 1),
                       ("columns", columns, column_count)):
        if x is not None and (x.numel() < n * k or not x.is_contiguous() or x.element_size() != 4):
            raise ValueError(f"{name} must be contiguous 32-bit with {n * k} elements")
    curves = abi.ClassicCurves(points=points.data_ptr(), point_counts=t(point_counts), roots=t(roots),
                               root_uvs=t(root_uvs), strand_count=n, cvs_per_strand=cvs)
    image = plan.device_image(points.device)
    inputs = abi.ClassicRuntimeInputs(plan_image=image.data_ptr(), plan=C.addressof(plan.info),
                                      primitive_ids=t(primitive_ids), random_prefixes=t(random_prefixes),
                                      columns=t(columns) if column_count else None, column_count=column_count,
                                      radius_scale=radius_scale,
                                      clumps=C.addressof(bindings) if clumps else None,
                                      surface_tangents=t(surface_tangents),
                                      noise_domain_positions=t(noise_domain_positions),
                                      root_relative=1 if root_relative else 0)
    out = outputs.descriptor()
    check(lib().nxg_classic_runtime_apply(C.byref(curves), C.byref(inputs), C.byref(out),
                                          C.c_void_p(workspace.data_ptr()), workspace.numel(),
                                          _stream_handle(stream)),
          "nxg_classic_runtime_apply")
    return outputs
