"""CPU restatement of the learned-guess network (test infrastructure only).

The per atom-pair density network n_in -> n_hidden -> n_hidden -> n_out with
SiLU after the hidden layers, and its loss (per-sample mean squared error over
the valid entries of a density block, averaged over the batch).  With
bf16=True the operands are rounded to BF16 (round to nearest even) at the
points where the GPU kernels round them: the input, the weight copies, the
hidden activations and the gradients that enter the matrix products; all
arithmetic is then float64, against the GPU's FP32 accumulation.  With
bf16=False it is the exact float64 network, whose gradient
tests/test_mlp.py checks by finite differences.

The network has no reference counterpart (the reference's SCF starts from
the SAD guess, src/pymodule/sadguessdriver.py:841-971).
"""
import numpy as np


def bf16_round(x):
    """float32 values rounded to BF16 (nearest even), returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params, n_in, n_hidden, n_out):
    """(W0, b0, W1, b1, W2, b2) of the packed parameter vector."""
    shapes = [(n_in, n_hidden), (n_hidden,), (n_hidden, n_hidden), (n_hidden,), (n_hidden, n_out), (n_out,)]
    out, off = [], 0
    for s in shapes:
        n = int(np.prod(s))
        out.append(np.asarray(params[off:off + n], dtype=np.float64).reshape(s))
        off += n
    return out


def _sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def _silu(x):
    return x * _sig(x)


def _dsilu(x):
    s = _sig(x)
    return s * (1.0 + x * (1.0 - s))


def mask(shape, n_out, block_cols):
    """(m, n_out) weights 1/(m ni nj) on the valid block entries, 0 elsewhere."""
    m = len(shape)
    o = np.arange(n_out)
    rows, cols = o // block_cols, o % block_cols
    valid = (rows[None, :] < shape[:, 0:1]) & (cols[None, :] < shape[:, 1:2])
    return valid / (m * shape[:, 0:1] * shape[:, 1:2]).astype(np.float64)


def forward(params, x, n_in, n_hidden, n_out, bf16=True):
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W0, b0, W1, b1, W2, b2 = unpack(params, n_in, n_hidden, n_out)
    xb = r(x)
    h1 = xb @ r(W0) + b0
    a1 = r(_silu(h1))
    h2 = a1 @ r(W1) + b1
    a2 = r(_silu(h2))
    y = a2 @ r(W2) + b2
    return y, (xb, h1, a1, h2, a2)


def loss_grad(params, x, target, shape, n_in, n_hidden, n_out, block_cols, bf16=True):
    """(loss, packed gradient, y) of one batch."""
    r = bf16_round if bf16 else (lambda a: np.asarray(a, dtype=np.float64))
    W0, b0, W1, b1, W2, b2 = unpack(params, n_in, n_hidden, n_out)
    y, (xb, h1, a1, h2, a2) = forward(params, x, n_in, n_hidden, n_out, bf16)
    w = mask(np.asarray(shape), n_out, block_cols)
    res = y - target
    loss = float(np.sum(w * res * res))
    dy = 2.0 * w * res
    dyb = r(dy)
    gW2, gb2 = a2.T @ dyb, dy.sum(axis=0)
    dh2 = (dyb @ r(W2).T) * _dsilu(h2)
    dh2b = r(dh2)
    gW1, gb1 = a1.T @ dh2b, dh2.sum(axis=0)
    dh1 = (dh2b @ r(W1).T) * _dsilu(h1)
    dh1b = r(dh1)
    gW0, gb0 = xb.T @ dh1b, dh1.sum(axis=0)
    grad = np.concatenate([g.ravel() for g in (gW0, gb0, gW1, gb1, gW2, gb2)])
    return loss, grad, y


def adam(params, grad, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (bias-corrected); returns (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    p = params - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps)
    return p, m, v
