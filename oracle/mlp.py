"""ORACLE (test infrastructure only): the attribute surrogate MLP in numpy.

Network (no reference counterpart; the generator's own learned form):
  x (10 features) -> normalise ((x - mean) * inv_std)
  -> 4 hidden layers of width 64: h = silu(h_prev @ W + b)
  -> output layer: y = h4 @ W5 + b5 (4 values) -> attribute = y * scale + offset
Loss: mean over batch and outputs of (y - t_norm)^2 with t_norm = (t - offset) / scale.

The device kernels compute in BF16 operands with FP32 accumulation.  The
weights are rounded to BF16; the normalised input and every hidden
activation are carried as two BF16 terms, hi = bf16(h) and lo = bf16(h - hi)
(so their value is hi + lo, about 16 significant bits); in the backward pass
the pre-activation gradient dz is rounded to BF16 for both products and the
bias gradient sums the unrounded dz.  `forward(..., bf16_round=True)` and
`loss_and_grads(..., bf16_round=True)` apply the same rounding so the device
results can be compared to FP64 arithmetic at the same rounding points;
`bf16_round=False` is the plain FP64 network, pinned by finite differences.
"""
import numpy as np

INPUTS, HIDDEN, LAYERS, OUTPUTS = 10, 64, 4, 4
IN_PAD, OUT_PAD = 16, 16


def layer_shapes():
    """(rows, cols) of the padded weight matrices, input-major."""
    return [(IN_PAD, HIDDEN)] + [(HIDDEN, HIDDEN)] * (LAYERS - 1) + [(HIDDEN, OUT_PAD)]


def param_count():
    return sum(r * c + c for r, c in layer_shapes())


def unpack(flat):
    """Flat FP32 parameter vector -> [(W, b)] (views)."""
    out, at = [], 0
    for r, c in layer_shapes():
        W = flat[at:at + r * c].reshape(r, c)
        at += r * c
        b = flat[at:at + c]
        at += c
        out.append((W, b))
    return out


def init_params(seed=0):
    """He-normal weights, zero biases; padding rows/cols zero."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(param_count(), dtype=np.float32)
    layers = unpack(flat)
    fan_in = [INPUTS] + [HIDDEN] * LAYERS
    for i, (W, b) in enumerate(layers):
        rows = INPUTS if i == 0 else HIDDEN
        cols = OUTPUTS if i == LAYERS else HIDDEN
        W[:rows, :cols] = rng.normal(0.0, np.sqrt(2.0 / fan_in[i]), size=(rows, cols)) * (0.5 if i == LAYERS else 1.0)
    return flat


def bf16(x):
    """Round float values to bfloat16 (nearest even), returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    bits = a.view(np.uint32).astype(np.uint64)
    rounded = ((bits + 0x7FFF + ((bits >> 16) & 1)) >> 16) << 16
    return rounded.astype(np.uint32).view(np.float32).astype(np.float64)


def split(x):
    """hi + lo of the two BF16 terms an activation is carried as."""
    x = np.asarray(x, np.float64)
    hi = bf16(x)
    return hi + bf16(x - hi)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def normalise(x, norm):
    mean, inv_std = norm[:INPUTS], norm[INPUTS:2 * INPUTS]
    return (np.asarray(x, np.float64) - mean) * inv_std


def forward(flat, x, norm, bf16_round=False):
    """Returns (y [B,4] normalised outputs, cache for backward)."""
    plain = lambda v: np.asarray(v, np.float64)
    rw, ra = (bf16, split) if bf16_round else (plain, plain)
    layers = unpack(np.asarray(flat, np.float32 if bf16_round else np.float64))
    h = np.zeros((len(x), IN_PAD))
    h[:, :INPUTS] = normalise(x, norm)
    h = ra(h)
    cache = []
    for i, (W, b) in enumerate(layers):
        Wr = rw(W)
        z = h @ Wr + b.astype(np.float64)
        cache.append((h, Wr, z))
        if i < LAYERS:
            h = ra(silu(z))
        else:
            h = z
    return h[:, :OUTPUTS], cache


# Deliberate mistakes loss_and_grads can plant, so that tests can show what
# a gradient check at a given tolerance detects (each at layer `layer`):
#   "dsilu_term"  silu's derivative without its z (1 - s) term (hidden layer)
#   "bias"        the layer's bias gradient left at zero
#   "unit"        the gradient through hidden unit 0 of the layer's input
#                 dropped (layers 1..4)
#   "transposed"  dh = dz @ W instead of dz @ W^T (square layers 1..3)
#   "example"     the last example's output gradient dropped (any layer)
PLANTS = ("dsilu_term", "bias", "unit", "transposed", "example")


def loss_and_grads(flat, x, targets_norm, norm, bf16_round=False, plant=None, layer=0, dtype=np.float64):
    """Mean-squared loss and its gradient w.r.t. the flat parameters.
    plant/layer inject one of PLANTS (None: the correct gradient); dtype
    float32 evaluates the backward pass in FP32 arithmetic (another valid
    accumulation order, for the same tests)."""
    assert plant is None or plant in PLANTS
    y, cache = forward(flat, x, norm, bf16_round)
    B = len(x)
    diff = y - targets_norm
    loss = float(np.mean(diff ** 2))
    dy = np.zeros((B, OUT_PAD), dtype)
    dy[:, :OUTPUTS] = 2.0 * diff / (B * OUTPUTS)
    if plant == "example":
        dy[-1] = 0.0
    grads = []
    dz = dy
    for i in range(LAYERS, -1, -1):
        h, Wr, z = (np.asarray(a, dtype) for a in cache[i])
        # the device feeds BF16 copies of dz to both products and sums the
        # FP32 dz for the bias
        dzr = np.asarray(bf16(dz), dtype) if bf16_round else dz
        dW = h.T @ dzr
        db = dz.sum(0)
        if plant == "bias" and i == layer:
            db = np.zeros_like(db)
        grads.append((dW, db))
        if i > 0:
            dh = (dzr @ Wr) if (plant == "transposed" and i == layer) else dzr @ Wr.T
            if plant == "unit" and i == layer:
                dh[:, 0] = 0.0
            zp = np.asarray(cache[i - 1][2], dtype)
            if plant == "dsilu_term" and i - 1 == layer:
                d = 1.0 / (1.0 + np.exp(-zp))
            else:
                d = dsilu(zp)
            dz = dh * d
    grads.reverse()
    g = np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads]).astype(np.float64)
    # padding entries carry no gradient in the packed layout
    mask = np.concatenate([np.concatenate([_pad_mask(i).ravel(), _bias_mask(i)]) for i in range(LAYERS + 1)])
    return loss, g * mask


def _pad_mask(i):
    r, c = layer_shapes()[i]
    m = np.zeros((r, c))
    rows = INPUTS if i == 0 else HIDDEN
    cols = OUTPUTS if i == LAYERS else HIDDEN
    m[:rows, :cols] = 1.0
    return m


def _bias_mask(i):
    c = layer_shapes()[i][1]
    m = np.zeros(c)
    m[:OUTPUTS if i == LAYERS else HIDDEN] = 1.0
    return m


def adam(flat, m, v, g, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step (FP64 restatement of the device update)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v
