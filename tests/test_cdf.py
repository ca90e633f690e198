"""The opt-in inverse-CDF root sampler's restatement (oracle/oracle_cdf.c) on
the CPU, pinned without retyping it:
- its prefix sums against a vectorised numpy construction of the same
  masses, and against what they must equal: the last sum is T * 2^32, and
  each triangle's share is its area * weight share of the mesh (the
  distribution the reference's alias table was built from, asset.cpp:43-73)
  to within the alias table's float probabilities;
- its triangle choice against numpy's searchsorted on the same targets;
- its empirical distribution against the masses (chi-square over 2M draws);
- its roots against the alias mode's where the two draws pick the same
  triangle (the barycentric draw and interpolation are shared)."""
import numpy as np

import nanoxgen_amd as nx
from nanoxgen_amd import abi, workloads


def hash32(x):
    x = np.asarray(x, np.uint32).copy()
    x ^= x >> np.uint32(16)
    x *= np.uint32(0x7FEB352D)
    x ^= x >> np.uint32(15)
    x *= np.uint32(0x846CA68B)
    x ^= x >> np.uint32(16)
    return x


def targets(seed, strands, T):
    s = np.arange(strands, dtype=np.uint64)
    h = [hash32(np.uint32(seed) ^ hash32((s + np.uint64(0x9E3779B9) * np.uint64(d + 1)).astype(np.uint32)))
         .astype(np.uint64) for d in (0, 1)]
    return h[0] * np.uint64(T) + ((h[1] * np.uint64(T)) >> np.uint64(32))


def alias_table(blob):
    h = abi.header_of(blob)
    raw = blob[h.alias_table_offset:h.alias_table_offset + 8 * h.triangle_count]
    return raw.view(np.float32)[0::2].copy(), raw.view(np.uint32)[1::2].copy()


def numpy_masses(blob):
    prob, alias = alias_table(blob)
    T = len(prob)
    k = np.where(prob >= 1.0, 1 << 24, np.where(prob > 0.0, np.ceil(prob.astype(np.float64) * 2.0 ** 24), 0))
    keep = k.astype(np.uint64) << np.uint64(8)
    mass = keep.copy()
    to = np.where(alias < T, alias, np.arange(T, dtype=np.uint32))
    np.add.at(mass, to, (np.uint64(1) << np.uint64(32)) - keep)
    return mass


def blobs():
    return [nx.build_asset(workloads.config1().build), nx.build_asset(workloads.config2().build)]


def test_prefix_sums_match_numpy_and_close_at_T_times_2_32(orc):
    for blob in blobs():
        T = abi.header_of(blob).triangle_count
        pre = orc.cdf_prefix(blob)
        assert np.array_equal(pre, np.cumsum(numpy_masses(blob), dtype=np.uint64))
        assert int(pre[-1]) == T << 32
        assert np.all(np.diff(pre.astype(np.float64)) >= 0)


def test_masses_are_the_area_weight_distribution(orc):
    """mass_t / 2^32 = T * w_t / sum(w) up to the alias table's float
    probabilities: each bucket contributes to a triangle within 2^-24."""
    for w in (workloads.config1(), workloads.config2()):
        b = w.build
        blob = nx.build_asset(b)
        pre = orc.cdf_prefix(blob)
        T = len(pre)
        mass = np.diff(pre, prepend=np.uint64(0)).astype(np.float64) / 2.0 ** 32
        p = b.positions.astype(np.float32)
        t = b.triangles
        c = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
        area = (np.float32(0.5) * np.sqrt((c * c).sum(1, dtype=np.float32))).astype(np.float64)
        weight = area * (b.triangle_weights.astype(np.float64) if b.triangle_weights is not None else 1.0)
        want = T * weight / weight.sum()
        _, alias = alias_table(blob)
        contributors = 1 + np.bincount(alias[alias < T], minlength=T)
        err = np.abs(mass - want)
        print(f"{w.name}: max |mass - T w / sum w| = {err.max():.3e} (bound 2^-24 per contributing bucket)")
        assert np.all(err <= contributors * 2.0 ** -24 + 1e-9)


def test_triangle_is_the_first_prefix_sum_above_the_target(orc):
    for blob in blobs():
        h = abi.header_of(blob)
        pre = orc.cdf_prefix(blob)
        n = 100_000
        roots = orc.cdf_roots(blob, 7, n, pre)
        want = np.searchsorted(pre, targets(7, n, h.triangle_count), side="right")
        assert np.array_equal(roots["triangle_index"], want.astype(np.uint32))


def test_empirical_distribution_follows_the_masses(orc):
    blob = nx.build_asset(workloads.config1().build)
    pre = orc.cdf_prefix(blob)
    T, n = len(pre), 2_000_000
    tri = np.searchsorted(pre, targets(3, n, T), side="right")
    counts = np.bincount(tri, minlength=T).astype(np.float64)
    expected = n * np.diff(pre, prepend=np.uint64(0)).astype(np.float64) / float(T << 32)
    chi2 = float(((counts - expected) ** 2 / expected).sum())
    z = (chi2 - (T - 1)) / np.sqrt(2.0 * (T - 1))
    print(f"chi-square {chi2:.1f} over {T - 1} degrees of freedom (z = {z:.2f})")
    assert abs(z) < 5.0


def test_roots_equal_the_alias_roots_on_the_same_triangle(orc):
    w = workloads.config1(strands=20_000)
    blob = nx.build_asset(w.build)
    cdf = orc.cdf_roots(blob, w.params.seed, w.params.strand_count)
    alias = orc.generate(blob, w.params)["roots"]
    same = cdf["triangle_index"] == alias["triangle_index"]
    assert same.sum() > 10  # a few draws land on the alias draw's triangle
    assert np.array_equal(cdf[same].view(np.uint8), alias[same].view(np.uint8))
    # and the draws differ in general (a different sampler, same distribution)
    assert same.mean() < 0.5
