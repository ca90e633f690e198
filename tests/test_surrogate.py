"""The attribute surrogate on the CPU: the FP64 oracle's gradients against
finite differences, the sensitivity of the GPU test's per-layer gradient
check (mistakes planted in the oracle's backward pass are flagged; another
valid accumulation order is not), the device layout query, host-side
refusals."""
import ctypes as C

import numpy as np
import pytest

import nanoxgen_amd as nx
from nanoxgen_amd import abi, surrogate
from oracle import mlp


def test_oracle_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    flat = mlp.init_params(1).astype(np.float64)
    x = rng.normal(size=(37, 10))
    t = rng.normal(size=(37, 4)) * 0.3
    norm = np.concatenate([rng.normal(size=10) * 0.1, rng.uniform(0.5, 2.0, 10), np.ones(4), np.zeros(4)])
    loss, g = mlp.loss_and_grads(flat, x, t, norm)
    idx = rng.choice(np.flatnonzero(g != 0), 80, replace=False)
    for i in idx:
        e = 1e-6
        a, b = flat.copy(), flat.copy()
        a[i] += e
        b[i] -= e
        fd = (mlp.loss_and_grads(a, x, t, norm)[0] - mlp.loss_and_grads(b, x, t, norm)[0]) / (2 * e)
        assert abs(fd - g[i]) <= 1e-5 * max(abs(fd), 1e-6)
    # padding parameters carry no gradient
    assert g.size == mlp.param_count() == 14608


def test_he_init_matches_oracle_shapes():
    flat = surrogate.he_init(3)
    assert flat.size == mlp.param_count()
    for (W, b), (r, c) in zip(mlp.unpack(flat), mlp.layer_shapes()):
        assert W.shape == (r, c) and not b.any()
    W0 = mlp.unpack(flat)[0][0]
    assert not W0[10:].any() and W0[:10].std() > 0.1
    W4 = mlp.unpack(flat)[4][0]
    assert not W4[:, 4:].any()


def test_layout():
    L = surrogate.layout()
    assert L.param_count == 14608 and L.max_batch == 8192
    offs = [L.params_offset, L.adam_m_offset, L.adam_v_offset, L.params_bf16_offset, L.norm_offset,
            L.loss_offset, L.partials_offset]
    assert offs == sorted(offs) and all(o % 256 == 0 for o in offs)
    assert L.total_bytes >= L.partials_offset + (8192 // 64) * (14608 + 1) * 4


def test_refusals_before_device_work():
    L = nx.lib()
    buf = np.zeros(64, np.float32)
    p = buf.ctypes.data
    assert L.nxg_mlp_train_step(None, p, p, 10, 0, 64, 1e-3, 1, None) == abi.NULL_ARGUMENT
    assert L.nxg_mlp_train_step(p, p, p, 10, 0, 9000, 1e-3, 1, None) == abi.INVALID_MLP_REQUEST
    assert L.nxg_mlp_train_step(p, p, p, 10, 0, 64, 0.0, 1, None) == abi.INVALID_MLP_REQUEST
    assert L.nxg_mlp_train_step(p, p, p, 10, 0, 64, 1e-3, 0, None) == abi.INVALID_MLP_REQUEST
    assert L.nxg_mlp_gradients(p, p, p, 0, 0, 64, None) == abi.INVALID_MLP_REQUEST
    assert L.nxg_verify_attributes(None, None, None, None, 0, None, 0, 1e-3, 16, None, 0, None, None) == \
        abi.NULL_ARGUMENT


def _problem():
    """Random nonzero parameters (biases too), quantised to float32, and a
    320-example batch like the GPU gradient test's."""
    rng = np.random.default_rng(8)
    flat = mlp.init_params(8)
    flat += (rng.normal(size=flat.size) * 0.05).astype(np.float32)
    x = (rng.normal(size=(320, 10)) * 2.0 + 0.5).astype(np.float32)
    t = (rng.normal(size=(320, 4)) * 0.3 + 1.0).astype(np.float32)
    norm = np.concatenate([np.zeros(10), np.full(10, 2.0), np.full(4, 0.3), np.ones(4)])
    return flat, x, (t - norm[24:28]) / norm[20:24], norm


PLANTED = [("dsilu_term", l) for l in range(4)] + [("bias", l) for l in range(5)] + \
          [("unit", l) for l in range(1, 5)] + [("transposed", l) for l in range(1, 4)] + [("example", 0)]


@pytest.mark.parametrize("plant,layer", PLANTED)
def test_gradient_check_detects_planted_mistakes(plant, layer):
    """The GPU gradient test's per-layer check (relative error in norm per
    layer and parameter kind, at the same BF16 rounding points) flags every
    planted mistake with at least twice its tolerance, down to one missing
    example of 320 (6.6e-3) or one hidden unit's path of 64 (6.6e-3 at
    layer 3)."""
    from test_gpu_surrogate import GRAD_TOL, per_layer_errors
    flat, x, tn, norm = _problem()
    _, good = mlp.loss_and_grads(flat, x, tn, norm, bf16_round=True)
    _, bad = mlp.loss_and_grads(flat, x, tn, norm, bf16_round=True, plant=plant, layer=layer)
    worst = max(e for _, _, e in per_layer_errors(bad, good))
    assert worst > 2 * GRAD_TOL, f"{plant} at layer {layer}: worst per-layer error {worst:.2e}"


def test_gradient_check_passes_another_accumulation_order():
    """The same gradients evaluated in FP32 arithmetic (a valid, different
    rounding of the accumulations) stay well inside the per-layer bar."""
    from test_gpu_surrogate import GRAD_TOL, per_layer_errors
    flat, x, tn, norm = _problem()
    _, good = mlp.loss_and_grads(flat, x, tn, norm, bf16_round=True)
    _, f32 = mlp.loss_and_grads(flat, x, tn, norm, bf16_round=True, dtype=np.float32)
    worst = max(e for _, _, e in per_layer_errors(f32, good))
    assert worst < 0.1 * GRAD_TOL, worst
