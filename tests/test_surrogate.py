"""The attribute surrogate on the CPU: the FP64 oracle's gradients against
finite differences, the device layout query, host-side refusals."""
import ctypes as C

import numpy as np

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
