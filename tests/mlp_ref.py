"""numpy restatement of the cost model's network (test infrastructure): the
16 -> 64 -> 64 -> 4 ReLU MLP of include/kifmm_mlp.h in float64, its
mean-squared-error gradient and Adam steps, to check the HIP kernels."""
import numpy as np

DIMS = (16, 64, 64, 4)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    out, o = [], 0
    for din, dout in zip(DIMS[:-1], DIMS[1:]):
        W = p[o:o + din * dout].reshape(din, dout)
        o += din * dout
        b = p[o:o + dout]
        o += dout
        out.append((W, b))
    assert o == p.size
    return out


def forward(p, X):
    h = np.asarray(X, dtype=np.float64)
    cache = [h]
    layers = unpack(p)
    for k, (W, b) in enumerate(layers):
        a = h @ W + b
        h = np.maximum(a, 0.0) if k < len(layers) - 1 else a
        cache.append(a)
    return h, cache


def loss_and_grad(p, X, T):
    layers = unpack(p)
    Y, cache = forward(p, X)
    n, dout = Y.shape
    d = Y - T
    loss = float((d * d).sum() / (n * dout))
    g = 2.0 * d / (n * dout)
    grads = []
    h_in = [cache[0]] + [np.maximum(a, 0.0) for a in cache[1:-1]]
    for k in range(len(layers) - 1, -1, -1):
        W, _ = layers[k]
        grads.append((h_in[k].T @ g, g.sum(0)))
        if k > 0:
            g = (g @ W.T) * (cache[k] > 0)
    grads.reverse()
    return loss, np.concatenate([np.concatenate([dW.ravel(), db]) for dW, db in grads])


def adam(p, X, T, epochs, lr, b1=0.9, b2=0.999, eps=1e-8):
    p = np.asarray(p, dtype=np.float64).copy()
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    losses = []
    for t in range(1, epochs + 1):
        loss, g = loss_and_grad(p, X, T)
        losses.append(loss)
        m = b1 * m + (1 - b1) * g
        v = b2 * v + (1 - b2) * g * g
        p -= lr * (m / (1 - b1 ** t)) / (np.sqrt(v / (1 - b2 ** t)) + eps)
    return p, np.asarray(losses)
