"""Learned surrogate for per-root attributes (host side of the C ABI's
nxg_mlp_* / nxg_surrogate_attributes / nxg_verify_attributes).

Surrogate: a 10 -> 64 -> 64 -> 64 -> 64 -> 4 SiLU MLP with BF16 weights and
activations and FP32 accumulation, trained on the GPU by the library's own
forward/backward/Adam kernels on (root features -> attributes) pairs that
the exact expression runtime generates for roots drawn with a disjoint seed.
At inference the exact runtime stays the arbiter: nxg_verify_attributes
re-evaluates a sample of roots exactly and sends every 64-root tile with a
sample outside the relative threshold back to the exact runtime.
"""
import ctypes as C

import numpy as np

from . import abi, check, lib, _stream_handle

INPUTS, HIDDEN, LAYERS, OUTPUTS = 10, 64, 4, 4
IN_PAD, OUT_PAD = 16, 16


class Layout(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("param_count", "params_offset", "adam_m_offset", "adam_v_offset",
                                          "params_bf16_offset", "fragments_offset", "norm_offset", "loss_offset", "partials_offset",
                                          "max_batch", "total_bytes")]


def layout():
    L = Layout()
    lib().nxg_mlp_get_layout(C.byref(L))
    return L


def layer_shapes():
    return [(IN_PAD, HIDDEN)] + [(HIDDEN, HIDDEN)] * (LAYERS - 1) + [(HIDDEN, OUT_PAD)]


def he_init(seed=0):
    """Flat FP32 parameters: He-normal weights on the real rows/columns,
    zero biases and zero padding; the output layer is scaled by 0.5."""
    rng = np.random.default_rng(seed)
    parts = []
    for i, (r, c) in enumerate(layer_shapes()):
        W = np.zeros((r, c), np.float32)
        rows = INPUTS if i == 0 else HIDDEN
        cols = OUTPUTS if i == LAYERS else HIDDEN
        W[:rows, :cols] = rng.normal(0.0, np.sqrt(2.0 / rows), size=(rows, cols)) * (0.5 if i == LAYERS else 1.0)
        parts += [W.ravel(), np.zeros(c, np.float32)]
    return np.concatenate(parts)


class Surrogate:
    """Device state of one surrogate network."""

    def __init__(self, device="cuda", seed=0):
        import torch
        self.L = layout()
        self.torch = torch
        self.state = torch.zeros(self.L.total_bytes, dtype=torch.uint8, device=device)
        self.set_params(he_init(seed))
        self.set_norm(np.zeros(INPUTS), np.ones(INPUTS), np.ones(OUTPUTS), np.zeros(OUTPUTS))
        self.step = 0

    def _view(self, offset, count, dtype):
        t = self.state[offset:offset + count * np.dtype(dtype).itemsize]
        return t.view(getattr(self.torch, np.dtype(dtype).name))

    def ptr(self):
        return C.c_void_p(self.state.data_ptr())

    def set_params(self, flat):
        flat = np.ascontiguousarray(flat, np.float32)
        assert flat.size == self.L.param_count
        self._view(self.L.params_offset, flat.size, np.float32).copy_(self.torch.from_numpy(flat))
        self._view(self.L.adam_m_offset, flat.size, np.float32).zero_()
        self._view(self.L.adam_v_offset, flat.size, np.float32).zero_()
        check(lib().nxg_mlp_sync(self.ptr(), _stream_handle(None)), "nxg_mlp_sync")
        self.step = 0

    def params(self):
        return self._view(self.L.params_offset, self.L.param_count, np.float32).cpu().numpy()

    def params_bf16(self):
        return self._view(self.L.params_bf16_offset, self.L.param_count, np.uint16).cpu().numpy()

    def fragments(self):
        """The BF16 weights in MFMA operand order (30 fragments x 64 lanes x 8)."""
        n = (self.L.norm_offset - self.L.fragments_offset) // 2
        return self._view(self.L.fragments_offset, min(n, 30 * 64 * 8), np.uint16).cpu().numpy()

    def set_norm(self, in_mean, in_std, out_scale, out_offset):
        norm = np.concatenate([in_mean, 1.0 / np.maximum(np.asarray(in_std, np.float64), 1e-6),
                               out_scale, out_offset]).astype(np.float32)
        self._view(self.L.norm_offset, 28, np.float32).copy_(self.torch.from_numpy(norm))
        self.norm = norm

    def fit_norm(self, features, targets):
        """Input standardisation and output scaling from training data (device tensors)."""
        f = features.double()
        t = targets.double()
        self.set_norm(f.mean(0).cpu().numpy(), f.std(0).cpu().numpy(), np.maximum(t.std(0).cpu().numpy(), 1e-6),
                      t.mean(0).cpu().numpy())

    def loss(self):
        return float(self._view(self.L.loss_offset, 1, np.float32).item())

    def partials(self, blocks):
        return self._view(self.L.partials_offset, blocks * (self.L.param_count + 1), np.float32).view(
            blocks, self.L.param_count + 1)

    def gradients(self, features, targets, first, batch, stream=None):
        check(lib().nxg_mlp_gradients(self.ptr(), C.c_void_p(features.data_ptr()), C.c_void_p(targets.data_ptr()),
                                      features.shape[0], first, batch, _stream_handle(stream)), "nxg_mlp_gradients")

    def train_step(self, features, targets, first, batch, lr, stream=None):
        self.step += 1
        check(lib().nxg_mlp_train_step(self.ptr(), C.c_void_p(features.data_ptr()), C.c_void_p(targets.data_ptr()),
                                       features.shape[0], first, batch, lr, self.step, _stream_handle(stream)),
              "nxg_mlp_train_step")

    def train(self, features, targets, epochs=40, batch=4096, lr=3e-3, lr_final=3e-5, log_every=0):
        """Adam with cosine-decayed learning rate; samples visited in a fixed
        order (the training roots are already a random draw)."""
        n = features.shape[0]
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        history = []
        for k in range(total):
            lr_k = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * k / total))
            self.train_step(features, targets, (k % steps_per_epoch) * batch, batch, lr_k)
            if log_every and (k % log_every == 0 or k == total - 1):
                history.append((k, self.loss()))
        return history

    def forward(self, features, out=None, stream=None):
        if out is None:
            out = self.torch.empty((features.shape[0], OUTPUTS), dtype=self.torch.float32, device=features.device)
        check(lib().nxg_mlp_forward(self.ptr(), C.c_void_p(features.data_ptr()), C.c_void_p(out.data_ptr()),
                                    features.shape[0], _stream_handle(stream)), "nxg_mlp_forward")
        return out


def root_features(asset, params, maps, out=None, stream=None):
    import torch
    if out is None:
        out = torch.empty((params.strand_count, INPUTS), dtype=torch.float32, device=asset.tensor.device)
    check(lib().nxg_root_features(C.byref(asset.desc), C.byref(params), C.byref(maps.structs[0]),
                                  C.c_void_p(out.data_ptr()), out.numel(), _stream_handle(stream)),
          "nxg_root_features")
    return out


def surrogate_attributes(asset, params, net, maps, out, stream=None, prepared=None):
    """Enqueue the fused surrogate; `prepared` (optional) is a workspace that
    nanoxgen_amd.prepare_packed filled for the same asset, seed
    and strand count (the roots then run grouped by triangle)."""
    ws = C.c_void_p(prepared.data_ptr()) if prepared is not None else None
    check(lib().nxg_surrogate_attributes(C.byref(asset.desc), C.byref(params), net.ptr(), C.byref(maps.structs[0]),
                                         C.c_void_p(out.data_ptr()), out.numel(), ws,
                                         prepared.numel() if prepared is not None else 0, _stream_handle(stream)),
          "nxg_surrogate_attributes")
    return out


def verify_attributes(asset, params, exprs, maps, attributes, threshold, stride, tile_flags, stats=None,
                      stream=None):
    check(lib().nxg_verify_attributes(C.byref(asset.desc), C.byref(params), C.byref(exprs.info), maps.structs,
                                      maps.count, C.c_void_p(attributes.data_ptr()), attributes.numel(), threshold,
                                      stride, C.c_void_p(tile_flags.data_ptr()), tile_flags.numel(),
                                      C.c_void_p(stats.data_ptr()) if stats is not None else None,
                                      _stream_handle(stream)), "nxg_verify_attributes")
    return attributes
