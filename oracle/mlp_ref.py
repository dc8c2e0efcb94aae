"""Float64 numpy restatement of the cost model's MLP (ORACLE ONLY).

Same contract as pmx_mlp_forward / pmx_mlp_backward / pmx_mlp_adam in
include/pmx_rt.h: ReLU hidden layers, linear output, loss =
sum((y - t)^2) / 2 / batch, parameters packed W_0, b_0, W_1, b_1, ...
with W_l of shape [dims[l+1], dims[l]].  Pinned by a finite-difference
check of its own gradient (tests/test_costmodel.py)."""

import numpy as np


def unpack(params, dims):
    out = []
    off = 0
    for l in range(len(dims) - 1):
        W = params[off:off + dims[l] * dims[l + 1]].reshape(dims[l + 1], dims[l])
        off += dims[l] * dims[l + 1]
        b = params[off:off + dims[l + 1]]
        off += dims[l + 1]
        out.append((W, b))
    return out


def forward(params, dims, x):
    layers = unpack(np.asarray(params, np.float64), dims)
    acts = [np.asarray(x, np.float64)]
    for l, (W, b) in enumerate(layers):
        y = acts[-1] @ W.T + b
        if l + 1 < len(layers):
            y = np.maximum(y, 0.0)
        acts.append(y)
    return acts


def loss_and_grad(params, dims, x, t):
    acts = forward(params, dims, x)
    layers = unpack(np.asarray(params, np.float64), dims)
    batch = x.shape[0]
    y = acts[-1]
    loss = 0.5 * np.sum((y - t) ** 2) / batch
    delta = (y - t) / batch
    grads = [None] * len(layers)
    for l in reversed(range(len(layers))):
        W, b = layers[l]
        gW = delta.T @ acts[l]
        gb = delta.sum(0)
        grads[l] = (gW, gb)
        if l > 0:
            delta = (delta @ W) * (acts[l] > 0)
    flat = np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])
    return loss, flat


def adam(params, grads, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * grads
    v = b2 * v + (1 - b2) * grads * grads
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return params - lr * mh / (np.sqrt(vh) + eps), m, v
