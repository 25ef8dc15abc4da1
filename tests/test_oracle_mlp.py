"""Pins the CPU oracle of the learned-proposal network (oracle/mlp.py) without
the GPU: finite differences and torch autograd for the gradient, torch's own
bf16 conversion and Adam for the pieces, and the C library's host-side
proposal-weight conversion against the restatement."""
import numpy as np
import pytest

from oracle import mlp


def small_problem(seed=0, n=6):
    rng = np.random.default_rng(seed)
    p = rng.normal(0, 0.2, mlp.N_PARAMS)
    ctx = rng.integers(0, mlp.N_TOK, size=(n, mlp.CTX))
    ctx[0, 5:] = mlp.N_TOK - 1  # padding
    labels = rng.integers(0, mlp.N_OUT, size=n)  # instructions and pointers
    return p, ctx, labels


def test_param_layout():
    assert mlp.N_PARAMS == 2864 * 64 + 512 * 256 + 256 + 256 * 2871 + 2871 == 1052471
    p = np.arange(mlp.N_PARAMS, dtype=np.float64)
    assert np.array_equal(mlp.pack(mlp.unpack(p)), p)


def test_gradient_matches_central_differences():
    p, ctx, labels = small_problem()
    loss, g = mlp.loss_grad(p, ctx, labels, bf16_ops=False, dtype=np.float64)
    rng = np.random.default_rng(1)
    offs = np.cumsum([0] + [int(np.prod(s)) for _, s in mlp.SHAPES])
    # parameters that the batch touches: embeddings of its tokens, all of W1/b1/W2/b2
    touched_e = [int(t) * mlp.EMB + d for t in ctx.reshape(-1)[:4] for d in (0, 17, 63)]
    idx = touched_e + list(rng.integers(offs[1], offs[-1], size=40))
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp, _ = mlp.loss_grad(q, ctx, labels, bf16_ops=False, dtype=np.float64)
        q[i] -= 2 * h
        lm, _ = mlp.loss_grad(q, ctx, labels, bf16_ops=False, dtype=np.float64)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, g[i])
    # an embedding row no context token uses has zero gradient
    unused = sorted(set(range(mlp.N_TOK)) - set(ctx.reshape(-1).tolist()))[0]
    assert not g[unused * mlp.EMB:(unused + 1) * mlp.EMB].any()


def test_gradient_matches_torch_autograd():
    torch = pytest.importorskip("torch")
    p, ctx, labels = small_problem(seed=3, n=9)
    loss, g = mlp.loss_grad(p, ctx, labels, bf16_ops=False, dtype=np.float64)
    w = {k: torch.tensor(v, dtype=torch.float64, requires_grad=True) for k, v in mlp.unpack(p).items()}
    X = w["E"][torch.tensor(ctx)].reshape(len(ctx), -1)
    H = torch.relu(X @ w["W1"] + w["b1"])
    Z = H @ w["W2"] + w["b2"]
    tl = torch.nn.functional.cross_entropy(Z, torch.tensor(labels))
    tl.backward()
    gt = mlp.pack({k: v.grad.numpy() for k, v in w.items()})
    assert abs(loss - tl.item()) < 1e-12
    np.testing.assert_allclose(g, gt, rtol=1e-9, atol=1e-13)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.normal(0, 1, 10000), rng.normal(0, 1e-30, 100), [0.0, -0.0, 1.0, 65504.0]])
    x = x.astype(np.float32)
    # exact ties: 1 + 2^-8 rounds to even (1.0); 1 + 3 * 2^-8 rounds up to 1 + 2^-6
    x = np.concatenate([x, np.float32([1 + 2 ** -8, 1 + 3 * 2 ** -8])])
    ours = mlp.bf16(x)
    theirs = torch.tensor(x).to(torch.bfloat16).to(torch.float32).numpy()
    assert np.array_equal(ours.view(np.uint32), theirs.view(np.uint32))
    assert ours[-2] == 1.0 and ours[-1] == np.float32(1 + 2 ** -6)


def test_bf16_gradient_close_to_exact():
    p, ctx, labels = small_problem(seed=7, n=16)
    _, g64 = mlp.loss_grad(p, ctx, labels, bf16_ops=False, dtype=np.float64)
    _, g16 = mlp.loss_grad(p.astype(np.float32), ctx, labels, bf16_ops=True, dtype=np.float32)
    rel = np.linalg.norm(g16 - g64) / np.linalg.norm(g64)
    assert rel < 6e-2, rel  # bf16 operands: ~2^-8 relative per rounding


def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(9)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(5)]
    tp = torch.tensor(p0, dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([tp], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for t, g in enumerate(grads, 1):
        opt.zero_grad()
        tp.grad = torch.tensor(g)
        opt.step()
        p, m, v = mlp.adam(p, g, m, v, t, 1e-2)
    np.testing.assert_allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_batch_indices_are_uniform(oracle):
    idx = mlp.batch_indices(0, 4096, 7, 123, oracle.philox4x32_10)
    counts = np.bincount(idx, minlength=7)
    assert counts.min() > 4096 / 7 * 0.85
    assert not np.array_equal(idx[:64], mlp.batch_indices(1, 64, 7, 123, oracle.philox4x32_10))


def test_proposal_weights_library_matches_restatement():
    from sx6502 import core
    rng = np.random.default_rng(11)
    probs = rng.dirichlet(np.full(2863, 0.05), size=8).astype(np.float32)
    probs[0] = 0
    probs[0, 17] = 1.0  # a one-hot row still gives every instruction weight >= 1
    for eps in (1.0, 0.25, 0.01):
        w = core.proposal_weights(probs, eps)
        assert np.array_equal(w, mlp.proposal_weights(probs, eps))
        assert (w.sum(axis=1, dtype=np.uint64) == 2 ** 32).all()
        assert w.min() >= 1
    # eps = 1 is the uniform table up to the remainder placement
    w = core.proposal_weights(probs, 1.0)
    assert w.max() - w.min() <= 2863 + 1
    with pytest.raises(core.SxError):
        core.proposal_weights(-probs, 0.1)
    with pytest.raises(core.SxError):
        core.proposal_weights(probs, 0.0)


def test_fold_pointers():
    probs = np.zeros((2, mlp.N_OUT), np.float32)
    probs[0, 5] = 0.25
    probs[0, mlp.N_CLS + 0] = 0.5   # position 0 holds instruction 7
    probs[0, mlp.N_CLS + 2] = 0.25  # position 2 is padding: dropped
    probs[1, mlp.N_CLS + 1] = 1.0   # position 1 holds instruction 5
    ctx = np.full((2, 8), mlp.N_TOK - 1)
    ctx[0, 0], ctx[0, 1] = 7, 5
    ctx[1, 0], ctx[1, 1] = 7, 5
    q = mlp.fold_pointers(probs, ctx)
    assert q.shape == (2, mlp.N_CLS)
    assert q[0, 7] == 0.5 and q[0, 5] == 0.25 and q[0].sum() == 0.75
    assert q[1, 5] == 1.0 and q[1].sum() == 1.0

