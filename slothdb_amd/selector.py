"""The learned strategy selector: parameters, GPU training and inference
through the C ABI (gbagg_selector_*), and attachment to an operator.

The model is a 32 -> 64 -> 64 -> 64 -> 5 ReLU MLP in FP32. Outputs: the
distinct-group count as log2((distinct + 1) / (default + 1)) / 4, a correction
to the sample's own estimate (default_estimate), and logits of the paths
DIRECT, LDS_HASH, GLOBAL_HASH, RADIX. The parameter vector follows include/slothdb_gpu.h
(GBAGG_SELECTOR_PARAMS floats)."""
import ctypes as C
import json
import math
import os

import numpy as np

from . import gbagg as G

N_IN, N_H, N_OUT = 32, 64, 5
LAYERS = [(N_H, N_IN), (N_H, N_H), (N_H, N_H), (N_OUT, N_H)]
N_W = sum(o * i + o for o, i in LAYERS)
N_PARAMS = N_W + 2 * N_IN
PATHS = [G.DIRECT, G.LDS_HASH, G.GLOBAL_HASH, G.RADIX]
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_weights.json")


def init_params(mu, sigma, seed=1):
    """He-normal weights, zero biases, the given input standardisation."""
    rng = np.random.default_rng(seed)
    parts = []
    for o, i in LAYERS:
        parts.append(rng.normal(0.0, math.sqrt(2.0 / i), (o, i)).astype(np.float32).ravel())
        parts.append(np.zeros(o, np.float32))
    p = np.concatenate(parts + [np.asarray(mu, np.float32), np.asarray(sigma, np.float32)])
    assert p.size == N_PARAMS
    return p


def unpack(p):
    """[(W, b)] per layer, then (mu, sigma)."""
    out, off = [], 0
    for o, i in LAYERS:
        W = p[off:off + o * i].reshape(o, i)
        off += o * i
        b = p[off:off + o]
        off += o
        out.append((W, b))
    return out, (p[off:off + N_IN], p[off + N_IN:off + 2 * N_IN])


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def train(x, y, params, epochs=300, lr=1e-3, lam=0.5):
    """Minibatch Adam on the GPU (gbagg_selector_train); returns (params, epoch losses)."""
    x, y, p = _f32(x), _f32(y), _f32(params).copy()
    loss = np.zeros(epochs, np.float32)
    G._check(G.lib().gbagg_selector_train(x.ctypes.data, y.ctypes.data, len(x), epochs, lr, lam,
                                          p.ctypes.data, p.size, loss.ctypes.data))
    return p, loss


def infer(params, x):
    """Forward pass on the GPU (gbagg_selector_infer): n x N_OUT outputs."""
    x, p = _f32(x), _f32(params)
    out = np.zeros((len(x), N_OUT), np.float32)
    if len(x):
        G._check(G.lib().gbagg_selector_infer(p.ctypes.data, p.size, x.ctypes.data, len(x), out.ctypes.data))
    return out


def default_estimate(f):
    """The sample's own distinct estimate from its features: keys seen plus
    the singleton share of the unsampled rows (the operator's default sizing)."""
    f = np.asarray(f, np.float64)
    rows = np.exp2(32 * f[..., 0]) - 1
    n = np.exp2(32 * f[..., 1]) - 1
    d = np.exp2(32 * f[..., 2]) - 1
    f1 = f[..., 4] * d
    rest = np.where(n > 0, f1 / np.maximum(n, 1) * rows, 0.0)
    return np.maximum(1.0, np.minimum(rows, d + rest))


def target(distinct, f):
    """Training target of output 0 for a query with this many groups."""
    return (np.log2(np.asarray(distinct, np.float64) + 1) - np.log2(default_estimate(f) + 1)) / 4


def distinct_from_output(o, f):
    return np.maximum(1.0, (default_estimate(f) + 1) * np.exp2(4 * np.asarray(o, np.float64)) - 1)


def save(params, path=WEIGHTS, meta=None):
    with open(path, "w") as f:
        json.dump({"shape": [N_IN, N_H, N_H, N_H, N_OUT], "params": [float(v) for v in params],
                   "meta": meta or {}}, f)


def load(path=WEIGHTS):
    with open(path) as f:
        d = json.load(f)
    p = np.array(d["params"], np.float32)
    if p.size != N_PARAMS:
        raise ValueError(f"{path}: {p.size} parameters, expected {N_PARAMS}")
    return p, d.get("meta", {})


def attach(op, params):
    """Plans of this operator (AUTO strategy) consult the selector."""
    p = _f32(params)
    if op._h is None:
        op.Init()
    G._check(G.lib().gbagg_selector_set(op._h, p.ctypes.data, p.size))
    op._selector_keep = p


def features(op, with_outputs=False):
    """Features (and selector outputs) of the operator's last finalize."""
    f = np.zeros(N_IN, np.float32)
    o = np.zeros(N_OUT, np.float32)
    G._check(G.lib().gbagg_plan_features(op._h, f.ctypes.data, o.ctypes.data if with_outputs else None))
    return (f, o) if with_outputs else f
