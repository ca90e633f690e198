"""GPU attribute surrogate through the C ABI: MFMA forward and backward
against the FP64 oracle evaluated at the same rounding points (BF16 weights,
activations as two BF16 terms, BF16 pre-activation gradients), the Adam
update, the root features, the sampled exact verification, and the trained
network at the 1e-3 gate on the surrogate workload's programs (what it gains
over the exact runtime it replaces: the tiles it serves).

Tolerances: forward outputs within 2e-5 of the oracle's output scale (the
device's FP32 accumulation order and hardware exp2/reciprocal move an
activation across a BF16 rounding boundary now and then); gradients within
GRAD_TOL relative in norm, per layer and separately for weights and biases;
features bit-exact; verified attributes bit-exact with the exact runtime on
re-evaluated tiles."""
import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, expr, surrogate, workloads
from oracle import mlp

pytestmark = pytest.mark.gpu
FORWARD_TOL = 2e-5  # of the output scale
GRAD_TOL = 2e-3     # relative, in norm, per layer and parameter kind


@pytest.fixture(scope="module")
def torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def random_problem(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(size=(n, 10)).astype(np.float32) * 2.0 + 0.5
    t = rng.normal(size=(n, 4)).astype(np.float32) * 0.3 + 1.0
    return x, t


def per_layer_errors(g_gpu, g_orc):
    """[(layer, kind, relative error in norm)] over the flat layout."""
    out, at = [], 0
    for l, (r, c) in enumerate(mlp.layer_shapes()):
        for kind, size in (("W", r * c), ("b", c)):
            a, b = g_gpu[at:at + size], g_orc[at:at + size]
            out.append((l, kind, float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))))
            at += size
    return out


def test_forward_matches_oracle(torch):
    net = surrogate.Surrogate(seed=7)
    net.set_norm(np.full(10, 0.5), np.full(10, 2.0), np.array([0.5, 0.3, 0.2, 0.1]), np.array([1.0, 1.0, 1.0, 0.5]))
    x, _ = random_problem(1000, 1)
    out = net.forward(torch.from_numpy(x).cuda()).cpu().numpy()
    y, _ = mlp.forward(net.params(), x, net.norm.astype(np.float64), bf16_round=True)
    ref = y * net.norm[20:24] + net.norm[24:28]
    err = np.abs(out - ref).max()
    scale = np.abs(y).max() * net.norm[20:24].max()
    print(f"forward: max |gpu - oracle| = {err:.3e} (output scale {scale:.3f})")
    assert err <= FORWARD_TOL * max(scale, 1.0)
    # the BF16 copy is the rounded master
    assert np.array_equal(net.params_bf16(), (mlp.bf16(net.params()).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16))


def test_fragments_hold_the_rounded_weights(torch):
    """The inference fragments are a permutation of each layer's BF16
    weights plus zero padding (their order is pinned by the forward test)."""
    net = surrogate.Surrogate(seed=4)
    frags = net.fragments()
    flat = net.params_bf16()
    per_layer = [4 * 512, 8 * 512, 8 * 512, 8 * 512, 2 * 512]  # fragments x 512 BF16
    at_f = at_p = 0
    for l, (r, c) in enumerate(mlp.layer_shapes()):
        f = frags[at_f:at_f + per_layer[l]]
        w = flat[at_p:at_p + r * c]
        assert np.array_equal(np.sort(f[f != 0]), np.sort(w[w != 0])), f"layer {l}"
        at_f += per_layer[l]
        at_p += r * c + c


def test_gradients_match_oracle(torch):
    net = surrogate.Surrogate(seed=8)
    net.set_norm(np.zeros(10), np.full(10, 2.0), np.full(4, 0.3), np.ones(4))
    x, t = random_problem(320, 2)
    net.gradients(torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda(), 0, 320)
    part = net.partials(5).cpu().numpy().astype(np.float64)
    g_gpu = part[:, :-1].sum(0)
    loss_gpu = part[:, -1].sum() / (320 * 4)
    tn = (t - net.norm[24:28]) / net.norm[20:24]
    loss, g = mlp.loss_and_grads(net.params(), x, tn, net.norm.astype(np.float64), bf16_round=True)
    errs = per_layer_errors(g_gpu, g)
    print("gradients: " + ", ".join(f"{k}{l} {e:.2e}" for l, k, e in errs) + f"; loss gpu {loss_gpu:.6f} oracle {loss:.6f}")
    for l, kind, e in errs:
        assert e <= GRAD_TOL, f"layer {l} {kind}: {e:.3e}"
    assert abs(loss_gpu - loss) <= 1e-5 * loss
    # padding parameters receive exactly zero gradient
    mask = np.concatenate([np.concatenate([mlp._pad_mask(i).ravel(), mlp._bias_mask(i)]) for i in range(5)])
    assert not g_gpu[mask == 0].any()


def test_adam_step_matches_oracle(torch):
    net = surrogate.Surrogate(seed=9)
    net.set_norm(np.zeros(10), np.ones(10), np.ones(4), np.zeros(4))
    x, t = random_problem(8192, 3)
    p0 = net.params().astype(np.float64)
    xd, td = torch.from_numpy(x).cuda(), torch.from_numpy(t).cuda()
    net.gradients(xd, td, 0, 8192)
    g = net.partials(128).cpu().numpy().astype(np.float64)[:, :-1].sum(0)
    net.train_step(xd, td, 0, 8192, 1e-3)
    p1 = net.params().astype(np.float64)
    want, _, _ = mlp.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, 1e-3)
    assert np.abs(p1 - want).max() <= 1e-6
    # both BF16 copies follow the update
    assert np.array_equal(net.params_bf16(), (mlp.bf16(net.params()).astype(np.float32).view(np.uint32) >> 16).astype(np.uint16))


def test_root_features_are_exact(torch, orc):
    w = workloads.config4(strands=5000, quads=32, guides_per_side=8)
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    maps = expr.TexelMaps(w.info["maps"])
    f = surrogate.root_features(asset, w.params, maps).cpu().numpy()
    tm = w.info["maps"][0]
    for s in range(0, 5000, 97):
        r = abi.RootSample()
        orc.lib().orc_sample_root_f32(blob.ctypes.data, w.params.seed, s, __import__("ctypes").byref(r))
        want = [r.position.x, r.position.y, r.position.z, r.normal.x, r.normal.y, r.normal.z, r.uv.x, r.uv.y]
        assert np.array_equal(f[s, :8], np.array(want, np.float32))
        b1, b2 = r.barycentric.x, r.barycentric.y
        pick = lambda u, v: tm[r.triangle_index, min(15, int(np.floor(np.float32(v) * np.float32(16)))),
                               min(15, int(np.floor(np.float32(u) * np.float32(16))))]
        assert f[s, 8] == pick(b1, b2) and f[s, 9] == pick(b2, b1)


def test_verify_extremes(torch, orc):
    """threshold 0: every tile is re-evaluated, attributes equal the exact
    runtime bitwise; threshold inf: only the samples change."""
    w = workloads.config4(strands=3000, quads=32, guides_per_side=8)
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    es, maps = expr.ExprSet(w.info["programs"]), expr.TexelMaps(w.info["maps"])
    exact = torch.empty((3000, 4), device="cuda")
    expr.evaluate_root_attributes(asset, w.params, es, maps, exact)
    flags = torch.zeros((3000 + 63) // 64, dtype=torch.int32, device="cuda")
    stats = torch.zeros(2, dtype=torch.int32, device="cuda")
    guess = torch.full((3000, 4), 0.25, device="cuda")
    surrogate.verify_attributes(asset, w.params, es, maps, guess, 0.0, 16, flags, stats)
    torch.cuda.synchronize()
    assert torch.equal(guess, exact)
    assert stats.tolist() == [len(flags), len(range(8, 3000, 16))]
    guess = torch.full((3000, 4), 0.25, device="cuda")
    surrogate.verify_attributes(asset, w.params, es, maps, guess, float("inf"), 16, flags, stats)
    torch.cuda.synchronize()
    sampled = torch.zeros(3000, dtype=torch.bool, device="cuda")
    sampled[8::16] = True
    assert torch.equal(guess[sampled], exact[sampled]) and bool((guess[~sampled] == 0.25).all())
    assert stats[0].item() == 0


@pytest.fixture(scope="module")
def trained(torch):
    """The surrogate workload's construction on a 64-quad plane: a network
    trained on the GPU on 2^17 exact-runtime pairs of seed-disjoint roots
    (300 epochs of 1,024-pair batches, about 4 s)."""
    w = workloads.config4(strands=65536, quads=64, guides_per_side=16)
    blob = nx.build_asset(w.build)
    asset = nx.DeviceAsset(blob)
    es, maps = expr.ExprSet(w.info["programs"]), expr.TexelMaps(w.info["maps"])
    train_p = w.params.copy(seed=w.info["train_seed"], strand_count=1 << 17)
    feats = surrogate.root_features(asset, train_p, maps)
    targets = torch.empty((1 << 17, 4), device="cuda")
    expr.evaluate_root_attributes(asset, train_p, es, maps, targets)
    net = surrogate.Surrogate(seed=1)
    net.fit_norm(feats, targets)
    hist = net.train(feats, targets, epochs=300, batch=1024, lr=2e-3, log_every=2000)
    torch.cuda.synchronize()
    print(f"trained: loss {hist[0][1]:.4f} -> {hist[-1][1]:.2e}")
    return dict(w=w, blob=blob, asset=asset, es=es, maps=maps, net=net, targets=targets, hist=hist)


def test_surrogate_gate_at_1e3(torch, trained):
    """The fused surrogate, then the exact check of one root per 64-root tile
    at 1e-3 relative: every re-evaluated tile equals the exact runtime
    bitwise, every checked root of a served tile is within 1e-3, point counts
    are exact; most tiles are served (the exact runtime they skip is what
    the network gains), and the unchecked roots of served tiles are reported
    against the exact runtime."""
    t = trained
    w, asset, es, maps, net = t["w"], t["asset"], t["es"], t["maps"], t["net"]
    N, stride, threshold = w.params.strand_count, 64, 1e-3
    attrs = torch.empty((N, 4), device="cuda")
    surrogate.surrogate_attributes(asset, w.params, net, maps, attrs)
    served_guess = attrs.clone()
    flags = torch.zeros((N + 63) // 64, dtype=torch.int32, device="cuda")
    stats = torch.zeros(2, dtype=torch.int32, device="cuda")
    surrogate.verify_attributes(asset, w.params, es, maps, attrs, threshold, stride, flags, stats)
    exact = torch.empty((N, 4), device="cuda")
    expr.evaluate_root_attributes(asset, w.params, es, maps, exact)
    torch.cuda.synchronize()
    tiles = len(flags)
    redo = flags.cpu().numpy() != 0
    assert stats.tolist() == [int(redo.sum()), len(range(stride // 2, N, stride))]
    a, e, g = attrs.cpu().numpy(), exact.cpu().numpy(), served_guess.cpu().numpy()
    tile_of = np.arange(N) // 64
    re_rows = redo[tile_of]
    assert np.array_equal(a[re_rows].view(np.uint32), e[re_rows].view(np.uint32))
    rel = np.abs(g - e) / np.maximum(np.abs(e), 1e-6)
    sampled = np.zeros(N, bool)
    sampled[stride // 2::stride] = True
    ok_rows = ~re_rows
    assert (rel[ok_rows & sampled] <= threshold).all()
    # served rows: the surrogate's values, the checked ones replaced by exact
    assert np.array_equal(a[ok_rows & sampled].view(np.uint32), e[ok_rows & sampled].view(np.uint32))
    assert np.array_equal(a[ok_rows & ~sampled].view(np.uint32), g[ok_rows & ~sampled].view(np.uint32))
    served = 1.0 - redo.mean()
    unchecked = rel[ok_rows & ~sampled].max(1)
    print(f"gate 1e-3: {tiles - redo.sum()} of {tiles} tiles served ({served:.3f}); unchecked served roots within "
          f"1e-3: {(unchecked <= threshold).mean():.4f}, max rel {unchecked.max() if unchecked.size else 0:.2e}; "
          f"all roots within 1e-3 before the gate: {(rel.max(1) <= threshold).mean():.4f}")
    assert served >= 0.5
    assert unchecked.size == 0 or unchecked.max() <= 2e-2
    # point counts do not depend on attributes: generation with the gated
    # attributes counts exactly what the exact attributes count
    out_s = nx.PackedOutputs(N, w.params.cvs_per_strand)
    out_e = nx.PackedOutputs(N, w.params.cvs_per_strand)
    nx.generate_packed(asset, w.params, out_s, octaves=w.octaves, attributes=attrs)
    nx.generate_packed(asset, w.params, out_e, octaves=w.octaves, attributes=exact)
    torch.cuda.synchronize()
    assert torch.equal(out_s.point_counts, out_e.point_counts)
    assert bool((out_e.point_counts == w.params.cvs_per_strand).all())


def test_training_beats_mean_predictor(torch, trained):
    """On held-out roots the trained network's RMS attribute error is far
    below that of predicting the training mean."""
    t = trained
    w, asset, es, maps, net = t["w"], t["asset"], t["es"], t["maps"], t["net"]
    test_p = w.params.copy(strand_count=8192)
    pred = torch.empty((8192, 4), device="cuda")
    surrogate.surrogate_attributes(asset, test_p, net, maps, pred)
    exact = torch.empty((8192, 4), device="cuda")
    expr.evaluate_root_attributes(asset, test_p, es, maps, exact)
    torch.cuda.synchronize()
    rms = ((pred - exact) ** 2).mean(0).sqrt().cpu().numpy()
    rms_mean = ((t["targets"].mean(0, keepdim=True) - exact) ** 2).mean(0).sqrt().cpu().numpy()
    print(f"rms per attribute {rms}, mean-predictor rms {rms_mean}")
    assert t["hist"][-1][1] < 0.01 * t["hist"][0][1]
    assert np.all(rms < 0.05 * rms_mean)


def test_fused_kernel_equals_forward_on_features(torch, trained):
    """nxg_surrogate_attributes (features computed in the kernel) equals
    nxg_root_features + nxg_mlp_forward bit for bit."""
    t = trained
    w, asset, maps, net = t["w"], t["asset"], t["maps"], t["net"]
    p = w.params.copy(strand_count=10_001)
    fused = torch.empty((p.strand_count, 4), device="cuda")
    surrogate.surrogate_attributes(asset, p, net, maps, fused)
    feats = surrogate.root_features(asset, p, maps)
    plain = net.forward(feats)
    torch.cuda.synchronize()
    assert torch.equal(fused, plain)


def test_surrogate_item_in_a_batch(torch, trained):
    """A batch holding a surrogate item (the network, then the exact check
    of one root per 64-root tile), an exact item with the same programs and
    a plain item: the surrogate item's attributes, tile flags and counts
    equal nxg_surrogate_attributes + nxg_verify_attributes alone bit for
    bit, and every item's outputs equal its single-description launch from
    those attributes; a second run (and a graph replay) repeats them."""
    from nanoxgen_amd import batch
    t = trained
    w, asset, es, maps, net = t["w"], t["asset"], t["es"], t["maps"], t["net"]
    p_sur = w.params.copy(strand_count=20_001)
    p_exact = w.params.copy(strand_count=5_000, seed=w.params.seed + 11)
    p_plain = w.params.copy(strand_count=777, seed=3)
    b = batch.Batch()
    outs = [nx.PackedOutputs(p.strand_count, p.cvs_per_strand, stencil=True) for p in (p_sur, p_exact, p_plain)]
    b.add(asset, p_sur, outs[0], octaves=w.octaves, exprs=es, maps=maps, surrogate=net,
          surrogate_threshold=1e-3, surrogate_stride=64)
    b.add(asset, p_exact, outs[1], octaves=w.octaves, exprs=es, maps=maps)
    b.add(asset, p_plain, outs[2], octaves=w.octaves)
    b.generate()
    torch.cuda.synchronize()
    n = p_sur.strand_count
    attrs = torch.empty((n, 4), device="cuda")
    surrogate.surrogate_attributes(asset, p_sur, net, maps, attrs)
    flags = torch.zeros((n + 63) // 64, dtype=torch.int32, device="cuda")
    stats = torch.zeros(2, dtype=torch.int32, device="cuda")
    surrogate.verify_attributes(asset, p_sur, es, maps, attrs, 1e-3, 64, flags, stats)
    torch.cuda.synchronize()
    e0 = b.entries[0]
    assert torch.equal(e0["attributes"], attrs)
    assert torch.equal(e0["surrogate_flags"], flags)
    assert b.stats(0) == tuple(stats.tolist())
    redo = int(flags.ne(0).sum())
    print(f"batch surrogate item: {len(flags) - redo} of {len(flags)} tiles served")
    exact = torch.empty((p_exact.strand_count, 4), device="cuda")
    expr.evaluate_root_attributes(asset, p_exact, es, maps, exact)
    torch.cuda.synchronize()
    assert torch.equal(b.entries[1]["attributes"], exact)
    fields = ("points", "roots", "root_uvs", "point_counts", "stencil_indices", "stencil_weights")
    want = []
    for p, out, a in zip((p_sur, p_exact, p_plain), outs, (attrs, exact, None)):
        single = nx.PackedOutputs(p.strand_count, p.cvs_per_strand, stencil=True)
        nx.generate_packed(asset, p, single, octaves=w.octaves, attributes=a)
        torch.cuda.synchronize()
        for f in fields:
            assert torch.equal(getattr(out, f), getattr(single, f)), f"{p.strand_count} strands: {f}"
        want.append(out.points.clone())
    # the flags and counters are cleared by the check itself on each run
    for run in ("eager", "graph"):
        for out in outs:
            out.points.zero_()
        e0["attributes"].fill_(float("nan"))
        torch.cuda.synchronize()
        if run == "eager":
            b.generate()
        else:
            g = b.graph()
            g.launch()
        torch.cuda.synchronize()
        assert torch.equal(e0["attributes"], attrs), run
        assert torch.equal(e0["surrogate_flags"], flags), run
        assert b.stats(0) == tuple(stats.tolist()), run
        for out, pts in zip(outs, want):
            assert torch.equal(out.points, pts), run
        if run == "graph":
            g.close()
