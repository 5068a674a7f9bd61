"""Float64 numpy restatement of the roundoff-surrogate MLP (test oracle).

Test infrastructure only: the GPU product trains and evaluates the network
with its own HIP kernels; this module is the checker for them.

Network: tanh hidden layers, one linear output, loss = mean squared error.
Parameters are one flat vector: for each layer l, W_l (out x in, row-major)
then b_l (out).  The reference has no such component (it picks FFT lengths
by fixed power laws, CUDALucas.cu:1027-1068, and reacts to roundoff after
the fact, CUDALucas.cu:2589-2648); the network is pinned by mathematics:
finite differences of the loss, the closed-form least-squares gradient of a
single linear layer, and hand-expanded Adam steps (tests/test_selector.py).
"""
import numpy as np


def layout(sizes):
    """[(w_offset, b_offset, in, out)] per layer and the total count."""
    out, off = [], 0
    for i, o in zip(sizes[:-1], sizes[1:]):
        out.append((off, off + i * o, i, o))
        off += i * o + o
    return out, off


def unpack(params, sizes):
    lay, total = layout(sizes)
    params = np.asarray(params, dtype=np.float64)
    assert params.shape == (total,)
    return [(params[w:b].reshape(o, i), params[b:b + o]) for w, b, i, o in lay]


def forward(params, sizes, x):
    """Returns (y[batch], activations per layer)."""
    a = np.asarray(x, dtype=np.float64)
    acts = [a]
    layers = unpack(params, sizes)
    for l, (W, b) in enumerate(layers):
        z = a @ W.T + b
        a = np.tanh(z) if l + 1 < len(layers) else z
        acts.append(a)
    return a[:, 0], acts


def loss_and_grad(params, sizes, x, t):
    """Mean squared error over the batch and its gradient (flat vector)."""
    t = np.asarray(t, dtype=np.float64)
    y, acts = forward(params, sizes, x)
    B = len(t)
    loss = float(np.mean((y - t) ** 2))
    layers = unpack(params, sizes)
    lay, total = layout(sizes)
    g = np.zeros(total)
    dz = (2.0 * (y - t) / B)[:, None]
    for l in range(len(layers) - 1, -1, -1):
        W, _ = layers[l]
        w_off, b_off, i, o = lay[l]
        g[w_off:b_off] = (dz.T @ acts[l]).ravel()
        g[b_off:b_off + o] = dz.sum(axis=0)
        if l:
            da = dz @ W
            dz = da * (1.0 - acts[l] ** 2)
    return loss, g


def adam_step(params, grad, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1, with the bias
    corrections folded into the step size); t counts from 1."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    lr_t = lr * np.sqrt(1 - b2 ** t) / (1 - b1 ** t)
    return params - lr_t * m / (np.sqrt(v) + eps), m, v
