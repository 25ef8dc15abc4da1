"""The learned-proposal network's HIP kernels against the CPU oracle
(oracle/mlp.py): the MFMA tile layout, initial weights, loss and gradient,
Adam training steps (batch draws included) and inference."""
import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    from sx6502 import core
    d = core.Device(0)
    yield d
    d.close()


def bf16_bits(x):
    return (mlp.bf16(np.asarray(x, np.float32)).view(np.uint32) >> 16).astype(np.uint16)


def bits_to_f32(b):
    return (b.astype(np.uint32) << 16).view(np.float32)


@pytest.mark.parametrize("M,N,K", [(64, 64, 16), (70, 33, 50), (128, 192, 96), (1, 2863, 256)])
def test_mfma_tile_layout(dev, M, N, K):
    print(f"mfma gemm M={M} N={N} K={K}", flush=True)
    rng = np.random.default_rng(M * 1000 + N + K)
    A = bf16_bits(rng.normal(size=(M, K)))
    B = bf16_bits(rng.normal(size=(K, N)))
    C = dev.mfma_gemm_test(A, B)
    ref = bits_to_f32(A).astype(np.float64) @ bits_to_f32(B).astype(np.float64)
    # products of bf16 values are exact in fp32; only the summation order differs
    np.testing.assert_allclose(C, ref, rtol=1e-5, atol=1e-5 * np.sqrt(K))


def test_initial_params_match_oracle(dev):
    m = dev.mlp(seed=5)
    assert m.n_params == mlp.N_PARAMS
    assert np.array_equal(m.params(), mlp.init_params(5))
    m.close()


def random_set(seed, n):
    rng = np.random.default_rng(seed)
    ctx = rng.integers(0, mlp.N_TOK, size=(n, mlp.CTX)).astype(np.uint16)
    ctx[: n // 3, 4:] = mlp.N_TOK - 1  # padded contexts
    labels = rng.integers(0, mlp.N_OUT, size=n).astype(np.uint16)
    p = rng.normal(0, 0.05, mlp.N_PARAMS).astype(np.float32)
    return p, ctx, labels


def block_rel_err(a, b):
    out, o = {}, 0
    for name, shape in mlp.SHAPES:
        n = int(np.prod(shape))
        x, y = a[o:o + n].astype(np.float64), b[o:o + n].astype(np.float64)
        out[name] = np.linalg.norm(x - y) / max(np.linalg.norm(y), 1e-30)
        o += n
    return out


@pytest.mark.parametrize("n", [1, 64, 300])
def test_loss_grad_matches_oracle(dev, n):
    print(f"loss_grad n={n}", flush=True)
    p, ctx, labels = random_set(n, n)
    m = dev.mlp(seed=0)
    m.set_params(p)
    loss, g = m.loss_grad(ctx, labels)
    lo, go = mlp.loss_grad(p, ctx, labels, bf16_ops=True, dtype=np.float32)
    assert abs(loss - lo) < 1e-4 * max(1.0, abs(lo))
    errs = block_rel_err(g, go)
    # same bf16 operands, fp32 accumulation: only summation order (and the
    # rare bf16 rounding it flips) differs
    assert max(errs.values()) < 2e-3, errs
    # and close to the exact float64 gradient
    _, g64 = mlp.loss_grad(p.astype(np.float64), ctx, labels, bf16_ops=False, dtype=np.float64)
    assert max(block_rel_err(g, g64).values()) < 6e-2  # bf16 operands: ~2^-8 relative per rounding
    m.close()


def test_train_steps_match_oracle_adam(dev, oracle):
    p, ctx, labels = random_set(42, 40)
    steps, batch, lr, seed = 3, 32, 1e-3, 77
    m = dev.mlp(seed=0)
    m.set_params(p)
    losses = m.train(ctx, labels, steps=steps, batch=batch, lr=lr, seed=seed)
    pg = m.params()
    q = p.copy()
    mm = np.zeros_like(q)
    vv = np.zeros_like(q)
    ref_losses = []
    for t in range(steps):
        idx = mlp.batch_indices(t, batch, len(ctx), seed, oracle.philox4x32_10)
        lo, g = mlp.loss_grad(q, ctx[idx], labels[idx], bf16_ops=True, dtype=np.float32)
        ref_losses.append(lo)
        q, mm, vv = mlp.adam(q, g, mm, vv, t + 1, np.float32(lr))
        q, mm, vv = q.astype(np.float32), mm.astype(np.float32), vv.astype(np.float32)
    np.testing.assert_allclose(losses, ref_losses, rtol=1e-4)
    d = np.abs(pg - q)
    # Adam moves each weight by about lr per step whatever the gradient's
    # size, so a weight whose tiny gradient differs in sign may move the
    # other way; such weights must be rare and the moves bounded.
    assert d.max() <= 2 * lr * steps + 1e-6
    assert np.mean(d > 0.05 * lr) < 0.01, np.mean(d > 0.05 * lr)
    m.close()


def test_predict_matches_oracle(dev):
    p, ctx, labels = random_set(3, 1100)  # crosses the 1,024-row predict tile
    m = dev.mlp(seed=0)
    m.set_params(p)
    probs = m.predict(ctx)
    Z, _ = mlp.forward(p, ctx, bf16_ops=True, dtype=np.float32)
    ref = mlp.softmax(Z.astype(np.float64))
    np.testing.assert_allclose(probs.sum(axis=1), 1.0, atol=1e-5)
    np.testing.assert_allclose(probs, ref, atol=2e-5, rtol=2e-3)
    m.close()


def test_proposal_folds_pointers_like_the_oracle(dev):
    p, ctx, _ = random_set(11, 300)
    m = dev.mlp(seed=0)
    m.set_params(p)
    probs = m.predict(ctx)
    for eps in (0.1, 0.02):
        w = m.proposal(ctx, eps)
        assert np.array_equal(w, mlp.proposal_weights(mlp.fold_pointers(probs, ctx), eps))
        assert (w.sum(axis=1, dtype=np.uint64) == 2 ** 32).all()
    m.close()


def test_training_learns_a_copy_task(dev):
    """Label = the context's first token: the loss must fall far below the
    uniform log(2863) and the argmax must recover the token."""
    rng = np.random.default_rng(8)
    ctx = rng.integers(0, mlp.N_CLS, size=(4096, mlp.CTX)).astype(np.uint16)
    labels = ctx[:, 0].copy()
    m = dev.mlp(seed=1)
    losses = m.train(ctx, labels, steps=600, batch=256, lr=3e-3, seed=5)
    assert np.isfinite(losses).all()
    assert losses[:10].mean() > 7.0
    assert losses[-50:].mean() < 2.0, losses[-50:].mean()
    probs = m.predict(ctx[:512])
    assert np.mean(probs.argmax(axis=1) == labels[:512]) > 0.8
    m.close()


def test_bad_tokens_are_rejected(dev):
    from sx6502 import core
    m = dev.mlp(seed=0)
    ctx = np.zeros((2, 8), np.uint16)
    ctx[1, 3] = 2864
    with pytest.raises(core.SxError):
        m.predict(ctx)
    with pytest.raises(core.SxError):
        m.loss_grad(np.zeros((2, 8), np.uint16), np.array([0, 2871], np.uint16))
    m.close()
