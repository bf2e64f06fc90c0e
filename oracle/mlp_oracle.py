"""CPU restatement of the learned start's network (test infrastructure only).

The engine's MLP (fdt_amd/csrc/fdt_mlp.hip, fdt_api.hip): input = the voxel's
signal over the mean of its b <= 50 volumes; layers N -> 512 -> 512 -> 256 ->
28 with W[l] stored in x out, GELU (erf form) after the first three; BF16
operands (weights, activations, back-propagated dZ) and FP32 accumulation.
`emulate_bf16=True` rounds at the same points as the GPU so the gradients can
be compared closely; `emulate_bf16=False` is the exact FP64 network that the
finite-difference test pins.  The loss is the one of sg_loss_kernel: squared
error on f_1..3, d and d_std targets, the sign-free ||p p^T - x x^T||^2 on the
directions weighted by the target f, and 0.1 x squared error on the log widths
of used slots (targets >= 1e29 mark unused ones).  Nothing here is part of
the product path; the engine never imports it.
"""
import numpy as np
from scipy.special import erf

SG_OUT = 28
B0_MAX = 50.0
W_LW = 0.1


def bf16(x):
    """round float32 values to bfloat16 (nearest even), returned as float64"""
    x = np.asarray(x, dtype=np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def normalise(signals, bvals):
    s = np.asarray(signals, dtype=np.float64)
    b0 = np.asarray(bvals) <= B0_MAX
    s0 = s[:, b0].mean(1) if b0.any() else np.ones(s.shape[0])
    s0 = np.where(s0 > 0, s0, 1.0)
    return (s.astype(np.float32) * (1.0 / s0.astype(np.float32))[:, None]).astype(np.float64)


def forward(x, layers, emulate_bf16=True):
    """x: normalised inputs [b][N]; returns (outputs, cache)"""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    a = q(x)
    acts, zs = [a], []
    for l, (W, bias) in enumerate(layers):
        z = a @ q(W) + np.asarray(bias, dtype=np.float64)
        zs.append(z)
        if l < len(layers) - 1:
            a = q(gelu(z))
            acts.append(a)
    return zs[-1], (acts, zs)


def loss_grad(o, t):
    """mean loss over rows and its gradient w.r.t. the outputs"""
    o = np.asarray(o, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    M = o.shape[0]
    g = np.zeros_like(o)
    e = o[:, 0:3] - t[:, 0:3]
    l = (e * e).sum(1)
    g[:, 0:3] = 2 * e
    for k in range(3):
        p = o[:, 3 + 3 * k:6 + 3 * k]
        x = t[:, 3 + 3 * k:6 + 3 * k]
        w = np.maximum(t[:, k], 0.0)
        pp = (p * p).sum(1)
        px = (p * x).sum(1)
        l += w * (pp * pp - 2 * px * px + 1)
        g[:, 3 + 3 * k:6 + 3 * k] = w[:, None] * (4 * pp[:, None] * p - 4 * px[:, None] * x)
    e = o[:, 12:14] - t[:, 12:14]
    l += (e * e).sum(1)
    g[:, 12:14] = 2 * e
    used = t[:, 14:] < 1e29
    e = np.where(used, o[:, 14:] - t[:, 14:], 0.0)
    l += W_LW * (e * e).sum(1)
    g[:, 14:] = 2 * W_LW * e
    return l.mean(), g / M


def backward(x, layers, t, emulate_bf16=True):
    """gradients [(dW, db) per layer] of the mean loss and the loss"""
    q = bf16 if emulate_bf16 else (lambda v: np.asarray(v, dtype=np.float64))
    o, (acts, zs) = forward(x, layers, emulate_bf16)
    loss, dO = loss_grad(o, t)
    dZ = q(dO)
    grads = [None] * len(layers)
    for l in range(len(layers) - 1, -1, -1):
        W = q(layers[l][0])
        grads[l] = (acts[l].T @ dZ, dZ.sum(0))
        if l == 0:
            break
        dA = dZ @ W.T
        dZ = q(dA * gelu_grad(zs[l - 1]))
    return grads, loss


def flat(grads):
    return np.concatenate([np.concatenate([dW.ravel(), db.ravel()]) for dW, db in grads])


def init_layers(n_in, seed=0):
    g = np.random.default_rng(seed)
    dims = [n_in, 512, 512, 256, SG_OUT]
    out = []
    for l in range(4):
        s = np.sqrt((1.0 if l == 3 else 2.0) / dims[l])
        out.append((g.standard_normal((dims[l], dims[l + 1])) * s, np.zeros(dims[l + 1])))
    return out
