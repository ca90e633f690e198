"""Synthetic, seeded descriptions for the strand-generation workloads.

Each function returns a BuildInputArrays (the flattened AssetBuildInput of
include/nanoxgen_amd.h) plus the generation parameters and noise octaves of
the workload.  Everything is derived from numpy's PCG64 with a fixed seed, so
the same call yields the same bytes on every machine.

Workloads (sizes as stated for the benchmark configurations):
  icosphere_description   icosphere r=1 at subdivision 3 (1,280 triangles),
                          radial jitter U(-0.05, 0.05); 64 guides at
                          Fibonacci-sphere roots, 8 CVs along the normal of
                          length 0.3 with lateral offsets N(0, 0.02); K=4.
  plane_description       512x512-quad displaced plane (524,288 triangles)
                          with log-normal(0, 0.5) per-triangle area weights;
                          4,096 guides of 16 CVs, length U(0.5, 1.5), curl
                          from a seeded sine sweep; K=8.
  config5                 nine descriptions of uneven size for the batched
                          launch (see its docstring).
Every builder takes its sizes as arguments, so tests can run the same
construction at reduced size.
"""
from dataclasses import dataclass, field

import numpy as np

from .abi import BuildInputArrays, GenerationParams, NoiseOctaves


@dataclass
class Workload:
    name: str
    build: BuildInputArrays
    params: GenerationParams
    octaves: NoiseOctaves = None
    info: dict = field(default_factory=dict)


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def icosphere(subdivisions):
    """Unit icosphere: 10*4^s+2 vertices, 20*4^s outward-wound triangles."""
    v, f = _icosahedron()
    verts = [tuple(x) for x in v]
    for _ in range(subdivisions):
        cache = {}
        faces = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = (np.array(verts[a]) + np.array(verts[b])) * 0.5
                m /= np.linalg.norm(m)
                cache[key] = len(verts)
                verts.append(tuple(m))
            return cache[key]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            faces += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = np.array(faces, dtype=np.int64)
    return np.array(verts, dtype=np.float64), f


def _fibonacci_sphere(n):
    i = np.arange(n, dtype=np.float64)
    y = 1.0 - 2.0 * (i + 0.5) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - y * y))
    phi = i * np.pi * (3.0 - 5.0 ** 0.5)
    return np.stack([r * np.cos(phi), y, r * np.sin(phi)], axis=1)


def _tangent_frame(n):
    """Two unit vectors orthogonal to each unit normal in n [G,3]."""
    ref = np.where(np.abs(n[:, 1:2]) < 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    t = np.cross(ref, n)
    t /= np.linalg.norm(t, axis=1, keepdims=True)
    b = np.cross(n, t)
    return t, b


def icosphere_description(seed=1, subdivisions=3, guides=64, guide_cvs=8, guide_length=0.3,
                          jitter=0.05, lateral_sigma=0.02, stencil_k=4):
    rng = np.random.default_rng(seed)
    v, f = icosphere(subdivisions)
    radial = 1.0 + rng.uniform(-jitter, jitter, size=len(v))
    positions = v * radial[:, None]
    # spherical texcoords of the undisplaced direction
    u = 0.5 + np.arctan2(v[:, 2], v[:, 0]) / (2.0 * np.pi)
    w = np.arccos(np.clip(v[:, 1], -1.0, 1.0)) / np.pi
    texcoords = np.stack([u, w], axis=1)
    roots_dir = _fibonacci_sphere(guides)
    roots = roots_dir * 1.0
    t, b = _tangent_frame(roots_dir)
    k = np.arange(guide_cvs, dtype=np.float64) / (guide_cvs - 1)
    lateral = rng.normal(0.0, lateral_sigma, size=(guides, guide_cvs, 2))
    lateral[:, 0, :] = 0.0  # the first CV is the root itself
    cvs = (roots[:, None, :] + roots_dir[:, None, :] * (guide_length * k)[None, :, None]
           + t[:, None, :] * lateral[..., 0:1] + b[:, None, :] * lateral[..., 1:2])
    build = BuildInputArrays(positions, f, cvs.reshape(-1, 3), np.full(guides, guide_cvs),
                             texcoords=texcoords, guide_root_normals=roots_dir,
                             stencil_k=stencil_k)
    return build


def config1(seed=1, strands=10_000):
    """Plain method, small: icosphere, 10,000 roots, 8 CVs, K=4, one octave."""
    build = icosphere_description(seed=seed)
    params = GenerationParams(strand_count=strands, cvs_per_strand=8, seed=seed, root_width=0.02,
                              tip_width=0.002, noise_amplitude=0.05, noise_frequency=5.0)
    return Workload("icosphere-1280tri-10k-strands-8cv", build, params)


def displaced_plane(quads, extent=2.0, seed=2):
    """(quads+1)^2 vertices on [-extent, extent]^2 (y up), displaced in y by a
    seeded sum of sines; 2*quads^2 triangles wound so normals point +y."""
    rng = np.random.default_rng(seed)
    n = quads + 1
    g = np.linspace(-extent, extent, n)
    x, z = np.meshgrid(g, g, indexing="xy")
    y = np.zeros_like(x)
    for _ in range(4):
        kx, kz = rng.uniform(0.3, 1.5, size=2)
        ph = rng.uniform(0, 2 * np.pi)
        amp = rng.uniform(0.02, 0.08)
        y += amp * np.sin(kx * x * np.pi + ph) * np.cos(kz * z * np.pi - ph)
    positions = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    uv = np.stack([(x + extent) / (2 * extent), (z + extent) / (2 * extent)], axis=-1).reshape(-1, 2)
    i, j = np.meshgrid(np.arange(quads), np.arange(quads), indexing="xy")
    v00 = (j * n + i).ravel()
    v10 = v00 + 1
    v01 = v00 + n
    v11 = v01 + 1
    # (v00, v01, v10): cross((v01-v00), (v10-v00)) = (+z) x (+x) = +y
    tri = np.empty((2 * len(v00), 3), dtype=np.int64)
    tri[0::2] = np.stack([v00, v01, v10], axis=1)
    tri[1::2] = np.stack([v10, v01, v11], axis=1)
    return positions, uv, tri


def plane_description(seed=2, quads=512, guides_per_side=64, guide_cvs=16, stencil_k=8):
    rng = np.random.default_rng(seed)
    positions, uv, tri = displaced_plane(quads, seed=seed)
    weights = rng.lognormal(mean=0.0, sigma=0.5, size=len(tri))
    # guides on a jittered grid, each bound to the triangle under it
    G = guides_per_side * guides_per_side
    gi, gj = np.meshgrid(np.arange(guides_per_side), np.arange(guides_per_side), indexing="xy")
    cell = quads / guides_per_side
    fx = (gi.ravel() + rng.uniform(0.1, 0.9, size=G)) * cell
    fz = (gj.ravel() + rng.uniform(0.1, 0.9, size=G)) * cell
    qi = np.minimum(fx.astype(np.int64), quads - 1)
    qj = np.minimum(fz.astype(np.int64), quads - 1)
    lx, lz = fx - qi, fz - qj
    upper = lx + lz > 1.0  # second triangle of the quad
    tri_index = 2 * (qj * quads + qi) + upper
    t = tri[tri_index]
    p0, p1, p2 = positions[t[:, 0]], positions[t[:, 1]], positions[t[:, 2]]
    # barycentrics (weights of vertices y and z) of the grid point
    b1 = np.where(upper, 1.0 - lx, lz)   # v01 weight
    b2 = np.where(upper, lx + lz - 1.0, lx)  # v10 or v11 weight
    b0 = 1.0 - b1 - b2
    roots = p0 * b0[:, None] + p1 * b1[:, None] + p2 * b2[:, None]
    nrm = np.cross(p1 - p0, p2 - p0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tang, bit = _tangent_frame(nrm)
    length = rng.uniform(0.5, 1.5, size=G)
    freq = rng.uniform(0.5, 2.0, size=G)
    phase = rng.uniform(0.0, 2.0 * np.pi, size=G)
    curl = rng.uniform(0.05, 0.2, size=G)
    s = np.arange(guide_cvs, dtype=np.float64) / (guide_cvs - 1)
    ang = 2.0 * np.pi * freq[:, None] * s[None, :] + phase[:, None]
    off_t = curl[:, None] * (np.sin(ang) - np.sin(phase)[:, None])
    off_b = curl[:, None] * (np.cos(ang) - np.cos(phase)[:, None])
    cvs = (roots[:, None, :] + nrm[:, None, :] * (length[:, None] * s[None, :])[..., None]
           + tang[:, None, :] * off_t[..., None] + bit[:, None, :] * off_b[..., None])
    uv_root = uv[t[:, 0]] * b0[:, None] + uv[t[:, 1]] * b1[:, None] + uv[t[:, 2]] * b2[:, None]
    build = BuildInputArrays(positions, tri, cvs.reshape(-1, 3), np.full(G, guide_cvs),
                             texcoords=uv, guide_root_normals=nrm, guide_root_uvs=uv_root,
                             guide_triangles=tri_index, guide_barycentrics=np.stack([b1, b2], 1),
                             stencil_k=stencil_k, triangle_weights=weights)
    return build


def config2(seed=2, strands=2_000_000, quads=512, guides_per_side=64):
    """Plain method, single large description: 2M strands x 16 CVs, K=8,
    two stacked octaves (0.1 @ 4, 0.05 @ 8)."""
    build = plane_description(seed=seed, quads=quads, guides_per_side=guides_per_side)
    params = GenerationParams(strand_count=strands, cvs_per_strand=16, seed=seed, root_width=0.01,
                              tip_width=0.001, noise_amplitude=0.1, noise_frequency=4.0)
    octaves = NoiseOctaves.of([(0.1, 4.0), (0.05, 8.0)])
    name = f"plane-{2 * quads * quads}tri-{guides_per_side ** 2}guides-{strands}-strands-16cv-2oct"
    return Workload(name, build, params, octaves)


def config3(seed=1, strands=1_000_000, program_seed=3, map_seed=4):
    """Plain method with expressions: the icosphere description, 1,000,000
    roots x 24 CVs, per-root length/width/noise-magnitude programs of 40 ops
    (4 noise3, 2 texel reads of a 16x16 per-face U(0,1) map)."""
    from . import expr
    build = icosphere_description(seed=seed)
    params = GenerationParams(strand_count=strands, cvs_per_strand=24, seed=seed, root_width=0.02,
                              tip_width=0.002, noise_amplitude=0.05, noise_frequency=5.0)
    programs = expr.attribute_programs(seed=program_seed)
    maps = [expr.texel_map(len(build.triangles), 16, seed=map_seed)]
    return Workload(f"icosphere-1280tri-{strands}-strands-24cv-3x40op-expr", build, params, None,
                    info=dict(programs=programs, maps=maps))


# Noise-frequency scale of the surrogate workload's programs.  Its four
# programs are built like the expression workload's (40 ops, 4 noise3 calls
# of scaled root positions and normals, 2 texel reads), with the noise3
# argument scales multiplied by this factor: 0.08-0.3 cycles per unit, under
# one and a half cycles across the 4-unit plane, the smoothly varying
# per-root attributes a groom's length, width, noise and clump maps carry.
# At the expression workload's scales (factor 1: up to 14 cycles across) a
# 10 -> 64x4 -> 4 network cannot meet a 1e-3 relative bound anywhere.
SURROGATE_FREQUENCY_SCALE = 0.1


def config4(seed=2, strands=1_000_000, quads=512, guides_per_side=64, program_seed=5, map_seed=6,
            frequency_scale=SURROGATE_FREQUENCY_SCALE):
    """Learned-surrogate form of the plane description: 1,000,000 roots x 24
    CVs, two octaves, per-root length/width/noise-magnitude/clump-weight
    programs of 40 ops (4 noise3, 2 texel reads of a 16x16 per-face map),
    noise3 argument scales multiplied by frequency_scale.  The surrogate
    trains on roots drawn with seed + 1000 (disjoint)."""
    from . import expr
    build = plane_description(seed=seed, quads=quads, guides_per_side=guides_per_side)
    params = GenerationParams(strand_count=strands, cvs_per_strand=24, seed=seed, root_width=0.01,
                              tip_width=0.001, noise_amplitude=0.1, noise_frequency=4.0)
    octaves = NoiseOctaves.of([(0.1, 4.0), (0.05, 8.0)])
    programs = expr.attribute_programs(seed=program_seed, clump=True, frequency_scale=frequency_scale)
    maps = [expr.texel_map(len(build.triangles), 16, seed=map_seed)]
    return Workload(f"plane-{2 * quads * quads}tri-{strands}-strands-24cv-2oct-4x40op-f{frequency_scale:g}-surrogate",
                    build, params, octaves,
                    info=dict(programs=programs, maps=maps, train_seed=seed + 1000, frequency_scale=frequency_scale))


# Batched uneven form: (strands, CVs per strand, mesh, stencil K, octaves,
# preserve-length, program ops).  CVs are paired with strand counts in
# ascending order (description i gets the i-th CV count), which gives
# 104,492,896 output points; meshes span 2,048 to 199,712 triangles and
# guides 256 to 4,096.
CONFIG5_DESCRIPTIONS = [
    (5_000, 4, ("plane", 32, 16), 4, 1, 0.0, 12),
    (20_000, 8, ("sphere", 4, 256), 8, 1, 0.0, 24),
    (60_000, 12, ("plane", 64, 24), 4, 2, 0.0, 16),
    (120_000, 16, ("sphere", 5, 1024), 8, 1, 0.4, 32),
    (200_000, 20, ("plane", 128, 32), 4, 2, 0.0, 20),
    (300_000, 24, ("sphere", 6, 2048), 8, 1, 0.0, 40),
    (400_000, 32, ("plane", 200, 48), 4, 2, 0.0, 24),
    (550_000, 48, ("plane", 256, 56), 8, 1, 0.0, 40),
    (801_139, 64, ("plane", 316, 64), 8, 2, 0.0, 48),
]


def config5(seed=50, strand_divisor=1, mesh_divisor=1, descriptions=None, frequency_scale=1.0):
    """Batched uneven form: nine descriptions, each with its own seeded mesh,
    guides, stencil K, octaves and per-root length/width/noise programs
    (plus a 4x4 per-face texel map), generated in one batched launch.
    strand_divisor / mesh_divisor shrink strand counts and mesh resolution
    (plane quads and guides per side, sphere subdivision) for tests;
    frequency_scale multiplies the programs' noise3 argument scales (1.0
    leaves them unchanged bit for bit)."""
    from . import expr
    out = []
    table = CONFIG5_DESCRIPTIONS if descriptions is None else [CONFIG5_DESCRIPTIONS[i] for i in descriptions]
    for i, (strands, cvs, mesh, k, octs, preserve, ops) in enumerate(table):
        dseed = seed + 7 * i
        if mesh[0] == "plane":
            quads = max(8, mesh[1] // mesh_divisor)
            side = max(4, min(mesh[2] // mesh_divisor, quads))
            build = plane_description(seed=dseed, quads=quads, guides_per_side=side, guide_cvs=16, stencil_k=k)
            base_len, root_w, tip_w = 1.0, 0.01, 0.001
        else:
            sub = max(2, mesh[1] - (mesh_divisor.bit_length() - 1))
            guides = max(32, mesh[2] // mesh_divisor)
            build = icosphere_description(seed=dseed, subdivisions=sub, guides=guides, guide_cvs=8, stencil_k=k)
            base_len, root_w, tip_w = 0.3, 0.02, 0.002
        params = GenerationParams(strand_count=max(1, strands // strand_divisor), cvs_per_strand=cvs, seed=dseed,
                                  root_width=root_w, tip_width=tip_w, noise_amplitude=0.08 * base_len,
                                  noise_frequency=4.0, noise_preserve_length=preserve)
        octaves = NoiseOctaves.of([(0.08 * base_len, 4.0), (0.04 * base_len, 8.0)]) if octs == 2 else None
        noise_calls = 4 if ops >= 24 else 2
        programs = [expr.random_attribute_program(dseed * 1000 + j, *expr.ATTRIBUTE_RANGES[n], n_ops=ops,
                                                  noise_calls=noise_calls, frequency_scale=frequency_scale)
                    for j, n in enumerate(["length", "width", "noise"])]
        maps = [expr.texel_map(len(build.triangles), 4, seed=dseed + 1)]
        name = f"d{i}-{mesh[0]}-{len(build.triangles)}tri-{params.strand_count}x{cvs}cv-K{k}-{octs}oct-{ops}op"
        if frequency_scale != 1.0:
            name += f"-f{frequency_scale:g}"
        out.append(Workload(name, build, params, octaves, info=dict(programs=programs, maps=maps, ops=ops)))
    return out
