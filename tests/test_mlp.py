"""The learned-guess network: the CPU oracle (oracle/mlp.py) pinned by finite differences and
closed forms; the GPU kernels (BF16 on the matrix cores, FP32 accumulation) against it.

GPU bar: the oracle rounds the same operands to BF16 and accumulates in float64, so the GPU's
outputs and gradients agree with it to FP32 accumulation order plus rare one-ulp BF16 rounding
differences of an activation: relative 2e-2 of each parameter block's largest gradient, 1e-2 of
the largest output.  A dropped SiLU derivative, a transposed operand or a wrong bias gradient
moves these by O(1).
"""
import numpy as np
import pytest

from oracle import mlp as orc


def _problem(m, n_in, n_hidden, n_out, block_cols, seed):
    rng = np.random.Generator(np.random.PCG64(seed))
    x = rng.normal(size=(m, n_in)).astype(np.float32)
    t = (0.3 * rng.normal(size=(m, n_out))).astype(np.float32)
    rows = n_out // block_cols
    shape = np.stack([rng.integers(1, rows + 1, m), rng.integers(1, block_cols + 1, m)], axis=1).astype(np.int32)
    dims = [(n_in, n_hidden), (n_hidden, n_hidden), (n_hidden, n_out)]
    params = np.concatenate([np.concatenate([rng.normal(size=fi * fo) * np.sqrt(1.0 / fi), 0.1 * rng.normal(size=fo)])
                             for fi, fo in dims]).astype(np.float32)
    return x, t, shape, params


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 1.0 + 2 ** -9, 1.0 + 2 ** -8 + 2 ** -20],
                 dtype=np.float32)
    r = orc.bf16_round(x)
    # 8 significant bits; ties go to even, just above a tie rounds up
    assert list(r) == [1.0, 1.0, 1.0 + 4 * 2 ** -8, -2.5, 0.0, 1.0, 1.0 + 2 ** -7]


def test_oracle_gradient_matches_finite_differences():
    n_in, nh, n_out, bc = 5, 7, 12, 4
    x, t, shape, params = _problem(6, n_in, nh, n_out, bc, 1)
    p = params.astype(np.float64)
    loss, grad, _ = orc.loss_grad(p, x, t, shape, n_in, nh, n_out, bc, bf16=False)
    h = 1e-6
    for k in np.random.Generator(np.random.PCG64(2)).choice(p.size, 40, replace=False):
        e = np.zeros_like(p)
        e[k] = h
        lp = orc.loss_grad(p + e, x, t, shape, n_in, nh, n_out, bc, bf16=False)[0]
        lm = orc.loss_grad(p - e, x, t, shape, n_in, nh, n_out, bc, bf16=False)[0]
        assert abs((lp - lm) / (2 * h) - grad[k]) < 1e-7 * max(1.0, abs(grad[k])) + 1e-9, k


def test_oracle_loss_closed_form():
    """A network with zero weights outputs its last bias: the loss is the masked mean squared error."""
    n_in, nh, n_out, bc = 3, 4, 6, 3
    x, t, shape, params = _problem(5, n_in, nh, n_out, bc, 3)
    p = np.zeros(params.size)
    p[-n_out:] = np.arange(n_out)
    loss, _, y = orc.loss_grad(p, x, t, shape, n_in, nh, n_out, bc, bf16=False)
    assert np.allclose(y, np.arange(n_out)[None, :])
    expect = 0.0
    for s in range(5):
        ni, nj = shape[s]
        blk = (np.arange(n_out)[None, :] - t[s:s + 1].astype(np.float64)).reshape(-1)
        sel = [o for o in range(n_out) if o // bc < ni and o % bc < nj]
        expect += np.mean(blk[sel] ** 2) / 5
    assert abs(loss - expect) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize('m,n_in,nh,n_out', [(100, 70, 256, 196), (64, 96, 256, 196), (300, 40, 128, 50)])
def test_gpu_forward_and_gradient_match_oracle(m, n_in, nh, n_out):
    from veloxchem_amd.densitynet import DensityNet
    bc = 14 if n_out == 196 else 10
    x, t, shape, params = _problem(m, n_in, nh, n_out, bc, 5)
    net = DensityNet(n_in, nh, n_out, bc, max_batch=512)
    net.set_params(params)
    y = net.forward(x).cpu().numpy()
    y_ref, _ = orc.forward(params.astype(np.float64), x, n_in, nh, n_out)
    err_y = np.max(np.abs(y - y_ref)) / np.max(np.abs(y_ref))
    loss, g = net.loss_grad(x, t, shape, want_grad=True)
    loss_ref, g_ref, _ = orc.loss_grad(params.astype(np.float64), x, t, shape, n_in, nh, n_out, bc)
    print(f'm {m}: forward rel {err_y:.2e}, loss {loss:.6e} vs {loss_ref:.6e}', flush=True)
    assert err_y < 1e-2
    assert abs(loss - loss_ref) < 1e-3 * loss_ref
    off = 0
    for fi, fo in [(n_in, nh), (nh, nh), (nh, n_out)]:
        for n in (fi * fo, fo):
            blk, ref = g[off:off + n], g_ref[off:off + n]
            rel = np.max(np.abs(blk - ref)) / np.max(np.abs(ref))
            print(f'  block of {n}: rel {rel:.2e}', flush=True)
            assert rel < 2e-2
            off += n


@pytest.mark.gpu
def test_gpu_adam_step_matches_oracle_and_training_lowers_loss():
    from veloxchem_amd.densitynet import DensityNet
    n_in, nh, n_out, bc = 32, 64, 36, 6
    x, t, shape, params = _problem(256, n_in, nh, n_out, bc, 7)
    net = DensityNet(n_in, nh, n_out, bc, max_batch=256)
    net.set_params(params)
    _, g = net.loss_grad(x, t, shape, want_grad=True)
    net.adam_step(1e-3)
    p1 = net.get_params()
    p_ref, _, _ = orc.adam(params.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, 1e-3)
    assert np.max(np.abs(p1 - p_ref)) < 1e-6
    losses = []
    for _ in range(200):
        loss, _ = net.loss_grad(x, t, shape)
        net.adam_step(3e-3)
        losses.append(loss)
    print(f'loss {losses[0]:.4e} -> {losses[-1]:.4e}', flush=True)
    assert losses[-1] < 0.5 * losses[0]
