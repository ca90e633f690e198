"""GPU parity of the gfx950 strand generator (through the C ABI) against the
CPU oracle on identical seeded inputs.

Bar (stated per test): root identities (the whole 48-byte RootSample),
stencil slots, point counts and radii bit-exact; positions within 1e-5 of
the bounding-box diagonal of the generated groom.  The error of both the GPU
and the oracle against the FP64 twin is printed for the record."""
import numpy as np
import pytest

from conftest import aligned_copy, params_from
import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads

pytestmark = pytest.mark.gpu
POSITION_TOL = 1e-5  # fraction of the groom bounding-box diagonal
WEIGHT_ULPS = 2      # stencil weights: pow from the device math library


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def run_gpu(blob, params, *, octaves=None, deformed=None, roots=None, stencil=True, radius_scale=1.0):
    import torch
    asset = nx.DeviceAsset(blob)
    out = nx.PackedOutputs(params.strand_count, params.cvs_per_strand, stencil=stencil and roots is None,
                           radius_scale=radius_scale)
    dev_roots = None
    if roots is not None:
        dev_roots = torch.from_numpy(np.ascontiguousarray(roots).view(np.int32).reshape(-1, 12)).cuda()
    ddesc, keep = None, []
    if deformed is not None:
        ddesc = abi.DeformedGeometry()
        for name, cnt in (("positions", "position_count"), ("normals", "normal_count"),
                          ("guide_cvs", "guide_cv_count")):
            if deformed.get(name) is not None:
                t = torch.from_numpy(np.ascontiguousarray(deformed[name], dtype=np.float32)).cuda()
                keep.append(t)
                setattr(ddesc, name, t.data_ptr())
                setattr(ddesc, cnt, t.shape[0])
    nx.generate_packed(asset, params, out, octaves=octaves, deformed=ddesc, roots=dev_roots)
    torch.cuda.synchronize()
    return out.to_host()


def compare(gpu, orc_out, f64=None, label=""):
    pts = orc_out["points"]
    lo, hi = pts[:, :3].min(0), pts[:, :3].max(0)
    diag = float(np.linalg.norm(hi.astype(np.float64) - lo))
    assert np.array_equal(gpu["roots"].view(np.uint8), orc_out["roots"].view(np.uint8)), "root records differ"
    assert np.array_equal(gpu["root_uvs"].view(np.uint32), orc_out["root_uvs"].view(np.uint32))
    assert np.array_equal(gpu["point_counts"], orc_out["point_counts"])
    if "stencil_indices" in gpu and "stencil_indices" in orc_out:
        assert np.array_equal(gpu["stencil_indices"], orc_out["stencil_indices"]), "stencil slots differ"
    if "stencil_weights" in gpu and "stencil_weights" in orc_out:
        # w = pow(1 - sqrt(d2) / r, p) (generate.h:232-240): every operand is
        # bit-exact, pow comes from the device math library, so the bar is a
        # distance in ulps (weights are >= 0: the bit patterns order them)
        gw = gpu["stencil_weights"].view(np.int32).astype(np.int64)
        ow = orc_out["stencil_weights"].view(np.int32).astype(np.int64)
        ulps = np.abs(gw - ow)
        print(f"[{label}] stencil weights: {int((ulps == 0).sum())} of {ulps.size} bit-exact, "
              f"max {int(ulps.max()) if ulps.size else 0} ulp")
        assert ulps.size == 0 or ulps.max() <= WEIGHT_ULPS, "stencil weights beyond the ulp bound"
    assert np.array_equal(gpu["points"][:, 3].view(np.uint32), pts[:, 3].view(np.uint32)), "radii differ"
    err = np.abs(gpu["points"][:, :3].astype(np.float64) - pts[:, :3]).max() if len(pts) else 0.0
    msg = f"[{label}] max |gpu - oracle| = {err:.3e} (tol {POSITION_TOL * diag:.3e}, diag {diag:.3f})"
    if f64 is not None:
        e_gpu = np.abs(gpu["points"][:, :3] - f64["points"][:, :3]).max()
        e_cpu = np.abs(pts[:, :3] - f64["points"][:, :3]).max()
        msg += f"; vs FP64: gpu {e_gpu:.3e}, oracle {e_cpu:.3e}"
    print(msg)
    assert err <= POSITION_TOL * diag, msg
    return err


def test_config1_full(torch, orc):
    """Icosphere, 10,000 strands x 8 CVs, K=4, noise 0.05 @ 5."""
    w = workloads.config1()
    blob = nx.build_asset(w.build)
    gpu = run_gpu(blob, w.params)
    ref = orc.generate(blob, w.params, stencil=True)
    compare(gpu, ref, orc.generate(blob, w.params, precision=64), "config1")


@pytest.mark.parametrize("variant", [
    dict(noise_amplitude=0.0),
    dict(noise_preserve_length=0.4, noise_correlation=0.3, noise_mask=0.83),
    dict(noise_preserve_length=1.0, guide_weight_power=0.5, guide_support_scale=1.3),
    dict(cvs_per_strand=2), dict(cvs_per_strand=13), dict(cvs_per_strand=64, strand_count=777),
    dict(strand_count=1), dict(strand_count=257, seed=99), dict(normal_rejection_cos=0.2),
    dict(length_scale=0.0), dict(guide_support_scale=0.0),
])
def test_config1_variants(torch, orc, variant):
    w = workloads.config1(strands=3000)
    blob = nx.build_asset(w.build)
    p = w.params.copy(**variant)
    gpu = run_gpu(blob, p, radius_scale=1.5)
    compare(gpu, orc.generate(blob, p, stencil=True, radius_scale=1.5), label=str(variant))


def test_reference_golden_cases(torch, orc, golden):
    """The reference's own outputs (fixtures written by the reference core)."""
    for case in ("noise", "plain", "preserve", "cv13"):
        blob = aligned_copy(golden["ico8_blob"])
        p = params_from(golden[f"ico8_{case}_params"])
        gpu = run_gpu(blob, p)
        ref = dict(points=golden[f"ico8_{case}_points"], roots=golden[f"ico8_{case}_roots"],
                   root_uvs=golden[f"ico8_{case}_root_uvs"], point_counts=golden[f"ico8_{case}_point_counts"])
        compare(gpu, ref, label=f"golden {case}")


def test_plane_two_octaves(torch, orc):
    """Config-2 construction at reduced size: K=8, 16 CVs, two octaves."""
    w = workloads.config2(strands=20000, quads=64, guides_per_side=16)
    blob = nx.build_asset(w.build)
    gpu = run_gpu(blob, w.params, octaves=w.octaves)
    ref = orc.generate(blob, w.params, octaves=w.octaves, stencil=True)
    compare(gpu, ref, orc.generate(blob, w.params, octaves=w.octaves, precision=64), "plane 2 octaves")
    # one octave given as an octave list equals the plain single octave bitwise
    one = abi.NoiseOctaves.of([(w.params.noise_amplitude, w.params.noise_frequency)])
    a = run_gpu(blob, w.params, octaves=one)
    b = run_gpu(blob, w.params)
    assert np.array_equal(a["points"].view(np.uint32), b["points"].view(np.uint32))


def test_deformed_geometry(torch, orc, golden):
    """Deformed patch + guides and patch-only deformation (generate.h:86-187)."""
    blob = aligned_copy(golden["plane_blob"])
    p = params_from(golden["plane_params"])
    h = abi.header_of(blob)
    nrm = blob[h.normals_offset:h.normals_offset + 12 * h.vertex_count].view(np.float32).reshape(-1, 3)
    gcv = blob[h.guide_cvs_offset:h.guide_cvs_offset + 12 * h.guide_cv_count].view(np.float32).reshape(-1, 3)
    dpos = golden["plane_deformed_positions"]
    shift = np.array([2.0, 3.0, -4.0], dtype=np.float32)
    for name, d in (("plane_deformed", dict(positions=dpos, normals=nrm, guide_cvs=gcv + shift)),
                    ("plane_patchonly", dict(positions=dpos))):
        gpu = run_gpu(blob, p, deformed=d)
        ref = orc.generate(blob, p, deformed=d, stencil=True)
        compare(gpu, ref, label=name)
        assert np.array_equal(ref["points"].view(np.uint32), golden[f"{name}_points"].view(np.uint32))


def test_explicit_roots(torch, orc, golden):
    """generate_packed_roots (generate.h:644-676): the GPU's own roots fed
    back reproduce its output bitwise; golden explicit-root case in tolerance."""
    w = workloads.config1(strands=5000)
    blob = nx.build_asset(w.build)
    a = run_gpu(blob, w.params)
    b = run_gpu(blob, w.params, roots=a["roots"])
    assert np.array_equal(a["points"].view(np.uint32), b["points"].view(np.uint32))
    pblob = aligned_copy(golden["plane_blob"])
    p = params_from(golden["plane_params"])
    gpu = run_gpu(pblob, p, roots=golden["plane_explicit_roots"])
    ref = dict(points=golden["plane_explicit_points"], roots=golden["plane_explicit_roots"],
               root_uvs=golden["plane_explicit_root_uvs"], point_counts=golden["plane_explicit_point_counts"])
    compare(gpu, ref, label="explicit roots")


def test_config2_full_size_properties(torch, orc):
    """The benchmark workload at full size (524,288 triangles, 4,096 guides,
    2,000,000 strands x 16 CVs, two octaves): size-independent properties
    everywhere, and the oracle on 4,096 strands spread over the range."""
    w = workloads.config2()
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    out = nx.PackedOutputs(w.params.strand_count, 16, stencil=True)
    nx.generate_packed(asset, w.params, out, octaves=w.octaves)
    out2 = nx.PackedOutputs(w.params.strand_count, 16, stencil=False)
    nx.generate_packed(asset, w.params, out2, octaves=w.octaves)
    torch.cuda.synchronize()
    assert torch.equal(out.points, out2.points), "generation must be deterministic"
    assert bool((out.point_counts == 16).all())
    roots = out.roots.cpu().numpy().view(abi.ROOT_DTYPE).reshape(-1)
    bary = roots["barycentric"]
    assert np.all(bary >= 0) and np.all(bary.sum(1) <= 1.00001)
    pts = out.points.view(-1, 16, 4)
    first = pts[:, 0, :3].cpu().numpy()
    assert np.array_equal(first.view(np.uint32), roots["position"].view(np.uint32)), "first CV is the root"
    t = np.arange(16, dtype=np.float32) / np.float32(15)
    radius = np.float32(0.5) * (np.float32(0.01) * (np.float32(1) - t) + np.float32(0.001) * t)
    assert np.array_equal(pts[0, :, 3].cpu().numpy().view(np.uint32), radius.view(np.uint32))
    assert bool((pts[:, :, 3] == pts[0:1, :, 3]).all())
    sample = np.linspace(0, w.params.strand_count - 1, 4096).astype(np.uint32)
    ref = orc.generate(blob, w.params, octaves=w.octaves, strands=sample, stencil=True)
    gpu = dict(points=out.points.view(-1, 16, 4)[torch.from_numpy(sample.astype(np.int64)).cuda()].reshape(-1, 4).cpu().numpy(),
               roots=roots[sample], root_uvs=out.root_uvs.cpu().numpy()[sample],
               point_counts=out.point_counts.cpu().numpy().view(np.uint32)[sample],
               stencil_indices=out.stencil_indices.cpu().numpy().view(np.uint32)[sample])
    compare(gpu, ref, orc.generate(blob, w.params, octaves=w.octaves, strands=sample, precision=64), "config2 sample")


def test_motion_samples(torch, orc, golden):
    """Motion samples (device_contract.h:42-52, 256-288): sample k's positions
    equal generate_packed under sample k's deformation, bit for bit, and the
    oracle's deformed generation within tolerance; the rest pose (no
    deformation) is a valid sample too."""
    blob = aligned_copy(golden["plane_blob"])
    p = params_from(golden["plane_params"])
    h = abi.header_of(blob)
    nrm = blob[h.normals_offset:h.normals_offset + 12 * h.vertex_count].view(np.float32).reshape(-1, 3)
    gcv = blob[h.guide_cvs_offset:h.guide_cvs_offset + 12 * h.guide_cv_count].view(np.float32).reshape(-1, 3)
    dpos = golden["plane_deformed_positions"]
    shift = np.array([2.0, 3.0, -4.0], dtype=np.float32)
    poses = [None, dict(positions=dpos, normals=nrm, guide_cvs=gcv + shift), dict(positions=dpos)]
    keep, samples = [], []
    for t, d in zip((0.0, 0.5, 1.0), poses):
        dd = abi.DeformedGeometry()
        for name, cnt in (("positions", "position_count"), ("normals", "normal_count"), ("guide_cvs", "guide_cv_count")):
            if d is not None and d.get(name) is not None:
                x = torch.from_numpy(np.ascontiguousarray(d[name], dtype=np.float32)).cuda()
                keep.append(x)
                setattr(dd, name, x.data_ptr())
                setattr(dd, cnt, x.shape[0])
        samples.append((t, dd))
    asset = nx.DeviceAsset(blob)
    N, Cv = p.strand_count, p.cvs_per_strand
    pts = torch.full((3 * N * Cv, 3), float("nan"), device="cuda")
    nx.generate_motion(asset, p, samples, pts)
    torch.cuda.synchronize()
    got = pts.cpu().numpy().reshape(3, N * Cv, 3)
    for k, d in enumerate(poses):
        single = run_gpu(blob, p, deformed=d)
        assert np.array_equal(got[k].view(np.uint32), single["points"][:, :3].view(np.uint32)), f"sample {k}"
        ref = orc.generate(blob, p, deformed=d)
        diag = float(np.linalg.norm(ref["points"][:, :3].max(0).astype(np.float64) - ref["points"][:, :3].min(0)))
        err = np.abs(got[k].astype(np.float64) - ref["points"][:, :3]).max()
        print(f"motion sample {k}: max |gpu - oracle| = {err:.3e} (tol {POSITION_TOL * diag:.3e})")
        assert err <= POSITION_TOL * diag
