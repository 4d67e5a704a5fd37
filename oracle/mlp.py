"""CPU oracle of the learned source-leg surrogate (numpy, FP64).

TEST INFRASTRUCTURE ONLY: the checker of the surrogate's device kernels.
Nothing in the product path imports this module.

The network is this build's own (the reference has no learned component;
its traveltimes come only from the table files, src/q-psdm.cu:2382-2622):
an MLP 5 -> 128 -> 128 -> 128 -> 128 -> 3 with SiLU, inputs
(sx, sy, x, y, z) mapped from a box to [-1, 1], outputs the residuals of
(t, t*, log A) over the anomaly-free straight-ray closed form, scaled.

`forward_backward(..., bf16=True)` rounds to bfloat16 (round to nearest
even) exactly where the device kernels do: the normalised inputs, every
weight matrix, the SiLU activations, the loss gradient and each layer's
dL/dZ; everything else is FP64.  With bf16=False it is the exact FP64
network, whose gradient the tests pin against central finite differences.
"""
import numpy as np

H, NIN, NOUT = 128, 5, 3
SHAPES = [("W1", (H, NIN)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (H, H)), ("b4", (H,)), ("W5", (NOUT, H)), ("b5", (NOUT,))]
NPARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(p):
    out, k = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = np.asarray(p[k:k + n], np.float64).reshape(shp)
        k += n
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], np.float64).ravel() for name, _ in SHAPES])


def bf16(x):
    """Round to bfloat16 (nearest even) via float32, returned as float64."""
    f = np.asarray(x, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def _id(x):
    return np.asarray(x, np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def base_leg(model, X):
    """(t, t*, -log sigma) of the straight ray through the model without its
    anomaly: closed form for linear v(z), exact layer pieces otherwise."""
    X = np.asarray(X, np.float64)
    sx, sy, x, y, z = X.T
    L = np.sqrt((x - sx) ** 2 + (y - sy) ** 2 + z ** 2)
    if model["kind"] == 0:
        v0, k, q0 = model["v0"], model["kgrad"], model["q0"]
        kz = k * z
        small = kz <= 1e-6 * v0
        kzs = np.where(small, 1.0, kz)
        t = np.where(small, L / v0, L / kzs * np.log1p(kzs / v0))
        sg = L * (v0 + 0.5 * np.where(small, 0.0, kz))
        return np.stack([t, t / q0, -np.log(sg)], 1)
    vl, ql, dz = np.asarray(model["vl"], float), np.asarray(model["ql"], float), model["layer_dz"]
    nl = len(vl)
    at = np.zeros_like(z)
    aq = np.zeros_like(z)
    av = np.zeros_like(z)
    prev = np.zeros_like(z)
    for l in range(nl):
        nxt = np.ones_like(z)
        if l + 1 < nl:
            zb = (l + 1) * dz
            nxt = np.where(zb < z, zb / np.where(z > 0, z, 1.0), 1.0)
        d = np.maximum(nxt - prev, 0.0)
        zm = 0.5 * (prev + nxt) * z
        li = np.clip((zm / dz).astype(int), 0, nl - 1)
        at += d / vl[li]
        aq += d / (vl[li] * ql[li])
        av += d * vl[li]
        prev = np.maximum(prev, nxt)
    return np.stack([at * L, aq * L, -np.log(av * L)], 1)


def normalise(X, lo, hi):
    lo, hi = np.asarray(lo, float), np.asarray(hi, float)
    return (np.asarray(X, np.float64) - lo) * (2.0 / (hi - lo)) - 1.0


def forward(p, X, lo, hi, model, out_scale, bf16_round=True):
    """Outputs (t, t*, log A) at X [n][5]."""
    r = bf16 if bf16_round else _id
    w = unpack(p)
    h = r(normalise(X, lo, hi)).T
    for l in range(1, 5):
        h = r(silu(r(w[f"W{l}"]) @ h + w[f"b{l}"][:, None]))
    y = (r(w["W5"]) @ h + w["b5"][:, None]).T
    return base_leg(model, X) + y * np.asarray(out_scale, float)


def forward_backward(p, X, T, lo, hi, model, out_scale, loss_w=(1.0, 1.0, 1.0), bf16_round=True):
    """Loss = mean over points of sum_o w_o (y_o - R_o)^2 with the residual
    targets R = (T - base) / scale; returns (loss, gradient [NPARAMS])."""
    r = bf16 if bf16_round else _id
    w = unpack(p)
    n = X.shape[0]
    Wb = {l: r(w[f"W{l}"]) for l in range(1, 6)}
    x0 = r(normalise(X, lo, hi)).T                      # [5][n]
    hs, zs = [x0], []
    for l in range(1, 5):
        z = Wb[l] @ hs[-1] + w[f"b{l}"][:, None]
        zs.append(z)
        hs.append(r(silu(z)))
    y = Wb[5] @ hs[-1] + w["b5"][:, None]               # [3][n]
    R = ((np.asarray(T, np.float64) - base_leg(model, X)) / np.asarray(out_scale, float)).T
    lw = np.asarray(loss_w, float)[:, None]
    d = y - R
    loss = float(np.sum(lw * d * d) / n)
    G = r(2.0 * lw * d / n)
    grads = {"W5": G @ hs[4].T, "b5": G.sum(1)}
    dz = r((Wb[5].T @ G) * dsilu(zs[3]))
    for l in range(4, 0, -1):
        grads[f"W{l}"] = dz @ hs[l - 1].T
        grads[f"b{l}"] = dz.sum(1)
        if l > 1:
            dz = r((Wb[l].T @ dz) * dsilu(zs[l - 2]))
    return loss, pack(grads)


def init_params(seed=0):
    """Xavier-uniform weights, zero biases (any initialisation serves the
    oracle; the device seeds its own)."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shp in SHAPES:
        if name.startswith("W"):
            a = np.sqrt(6.0 / (shp[0] + shp[1])) * (0.1 if name == "W5" else 1.0)
            d[name] = rng.uniform(-a, a, shp)
        else:
            d[name] = np.zeros(shp)
    return pack(d)
