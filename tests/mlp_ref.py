"""CPU reference of the viscosity surrogate network (test infrastructure).

A 7 -> 64 -> 64 -> 1 multilayer perceptron with SiLU activations, in float64 numpy,
with the parameter layout of the device kernels (geoflac_amd/csrc/gf_mlp.h):

    W1[64][7] | b1[64] | W2[64][64] | b2[64] | W3[64] | b3[1]      (4737 floats)

loss(w) = mean over the batch of (y - t)^2.  The backward pass is written out by hand
and pinned by central finite differences in tests/test_mlp_cpu.py.
"""
import numpy as np

NIN, NH = 7, 64
SIZES = [("W1", (NH, NIN)), ("b1", (NH,)), ("W2", (NH, NH)), ("b2", (NH,)), ("W3", (NH,)), ("b3", (1,))]
NPARAM = sum(int(np.prod(s)) for _, s in SIZES)


def unpack(w):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(w[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def init(seed=0):
    """Glorot-uniform weights, zero biases (the device trainer's initialisation)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SIZES:
        if name.startswith("W"):
            fan_in = shape[-1] if len(shape) == 2 else NH
            fan_out = shape[0] if len(shape) == 2 else 1
            lim = np.sqrt(6.0 / (fan_in + fan_out))
            d[name] = rng.uniform(-lim, lim, size=shape)
        else:
            d[name] = np.zeros(shape)
    return pack(d)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def forward(w, x):
    p = unpack(w)
    z1 = x @ p["W1"].T + p["b1"]
    h1 = silu(z1)
    z2 = h1 @ p["W2"].T + p["b2"]
    h2 = silu(z2)
    y = h2 @ p["W3"] + p["b3"][0]
    return y, (z1, h1, z2, h2)


def loss_grad(w, x, t):
    """Mean squared error and its gradient with respect to the packed parameters."""
    p = unpack(w)
    y, (z1, h1, z2, h2) = forward(w, x)
    n = x.shape[0]
    r = y - t
    loss = float(np.mean(r * r))
    dy = 2.0 * r / n
    g = {"W3": h2.T @ dy, "b3": np.array([dy.sum()])}
    dz2 = np.outer(dy, p["W3"]) * dsilu(z2)
    g["W2"] = dz2.T @ h1
    g["b2"] = dz2.sum(axis=0)
    dh1 = dz2 @ p["W2"]
    dz1 = dh1 * dsilu(z1)
    g["W1"] = dz1.T @ x
    g["b1"] = dz1.sum(axis=0)
    return loss, pack(g)


def bf16(a):
    """Round float32 values to bfloat16 (nearest, ties to even) and back."""
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def forward_bf16(w, x):
    """The inference the device runs on the matrix cores: inputs, weights and hidden
    activations rounded to bfloat16, products accumulated in float32, biases and
    activations in float32."""
    p = {k: np.asarray(v, dtype=np.float32) for k, v in unpack(w).items()}
    x = bf16(x)
    z1 = bf16(x) @ bf16(p["W1"]).T + p["b1"]
    h1 = bf16(silu(z1.astype(np.float64)).astype(np.float32))
    z2 = h1 @ bf16(p["W2"]).T + p["b2"]
    h2 = bf16(silu(z2.astype(np.float64)).astype(np.float32))
    return h2 @ bf16(p["W3"]) + p["b3"][0]
