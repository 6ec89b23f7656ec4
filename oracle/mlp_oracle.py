"""CPU oracle of the learned form's network -- TEST INFRASTRUCTURE ONLY.

Only tests/ may import this module, as the checker of the HIP kernels in
coqui_amd/csrc/mlp.hip.  The network has no counterpart in the reference (it is the learned
part of the Krylov Dyson solve); this restatement follows the network's own definition:

    h0 = silu(x W0^T + b0),  h1 = silu(h0 W1^T + b1),  y = h1 W2^T + b2,   silu(z) = z / (1 + e^-z)

with the parameters in one flat array W0, b0, W1, b1, W2, b2 (torch.nn.Linear layout).
The GPU products read BF16 operands (round to nearest even from FP32) and accumulate in
FP32; with round_bf16=True the oracle rounds the same operands the same way and sums in
float64, so the two differ only by FP32 summation order and by the rare BF16 rounding that
lands on the other side because of it.

Pins (tests/test_oracle_pins.py): the backward pass against central finite differences of
the forward pass in float64 (round_bf16=False); the Adam step against its closed form on a
scalar sequence.
"""
import numpy as np


def bf16(x):
    """Round to bfloat16 (nearest even), via float32 as the GPU does; returns float64."""
    f = np.ascontiguousarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(n_in, width, n_out):
    """Offsets of W0, b0, W1, b1, W2, b2 in the flat parameter array and its length."""
    o = {}
    o["W0"] = 0
    o["b0"] = o["W0"] + width * n_in
    o["W1"] = o["b0"] + width
    o["b1"] = o["W1"] + width * width
    o["W2"] = o["b1"] + width
    o["b2"] = o["W2"] + n_out * width
    return o, o["b2"] + n_out


def unpack(params, n_in, width, n_out):
    o, n = layout(n_in, width, n_out)
    p = np.asarray(params, np.float64)
    assert p.size == n
    return dict(W0=p[o["W0"]:o["b0"]].reshape(width, n_in), b0=p[o["b0"]:o["W1"]],
                W1=p[o["W1"]:o["b1"]].reshape(width, width), b1=p[o["b1"]:o["W2"]],
                W2=p[o["W2"]:o["b2"]].reshape(n_out, width), b2=p[o["b2"]:])


def _silu(z):
    return z / (1.0 + np.exp(-z))


def _dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, X, round_bf16=True):
    """Returns (Y, cache) for the batch X (rows)."""
    n_in, width, n_out = dims
    P = unpack(params, n_in, width, n_out)
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    X = np.asarray(X, np.float64)
    h0p = r(X) @ r(P["W0"]).T + P["b0"]
    h0 = _silu(h0p)
    h1p = r(h0) @ r(P["W1"]).T + P["b1"]
    h1 = _silu(h1p)
    Y = r(h1) @ r(P["W2"]).T + P["b2"]
    return Y, dict(X=X, h0p=h0p, h0=h0, h1p=h1p, h1=h1, P=P)


def backward(cache, dims, dY, round_bf16=True):
    """Gradient of sum(Y * dY) with respect to the flat parameters."""
    n_in, width, n_out = dims
    r = bf16 if round_bf16 else (lambda a: np.asarray(a, np.float64))
    P = cache["P"]
    dY = np.asarray(dY, np.float64)
    g = {}
    g["W2"] = r(dY).T @ r(cache["h1"])
    g["b2"] = dY.sum(0)
    d1 = (r(dY) @ r(P["W2"])) * _dsilu(cache["h1p"])
    g["W1"] = r(d1).T @ r(cache["h0"])
    g["b1"] = d1.sum(0)
    d2 = (r(d1) @ r(P["W1"])) * _dsilu(cache["h0p"])
    g["W0"] = r(d2).T @ r(cache["X"])
    g["b0"] = d2.sum(0)
    return np.concatenate([g[k].ravel() for k in ("W0", "b0", "W1", "b1", "W2", "b2")])


def adam(w, m, v, g, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma and Ba 2015, Algorithm 1); step counts from 1.  Returns (w, m, v)."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    mh = m / (1.0 - b1 ** step)
    vh = v / (1.0 - b2 ** step)
    return w - lr * mh / (np.sqrt(vh) + eps), m, v
