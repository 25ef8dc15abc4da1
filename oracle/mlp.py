"""CPU oracle for the learned STOKE proposal network (test infrastructure only).

Only tests/, the smoke check and the CPU-baseline leg of bench.py use this
module; the product path runs the HIP kernels of lib6502core.

The network has no counterpart in the reference (its mutator draws
instructions uniformly, pkg/stoke/mutator.go:108-121); what this file restates
is the definition in include/sx6502.h:

    X = [E[t0] .. E[t7]]            8 tokens x 64-d embeddings = 512
    H = relu(X W1 + b1)             256
    Z = H W2 + b2                   2,871 logits: 2,863 instructions, then
                                    8 pointers (output 2863 + j = "the
                                    instruction at context position j")
    loss = mean_b -log softmax(Z_b)[y_b]

with bf16 operands for the products (E, W1, W2, H and the back-propagated
dZ, dH rounded to nearest-even bf16) and fp32 accumulation, when ``bf16`` is
set; with ``bf16=False`` and float64 it is the exact function, which the
tests pin by central finite differences and by torch autograd on the CPU.
"""
from __future__ import annotations

import numpy as np

N_TOK = 2864
N_CLS = 2863
CTX = 8
N_OUT = N_CLS + CTX
EMB = 64
IN = CTX * EMB
HID = 256

SHAPES = [("E", (N_TOK, EMB)), ("W1", (IN, HID)), ("b1", (HID,)), ("W2", (HID, N_OUT)), ("b2", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 to bfloat16 (nearest, ties to even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p: np.ndarray) -> dict:
    out, o = {}, 0
    for name, shape in SHAPES:
        n = int(np.prod(shape))
        out[name] = p[o:o + n].reshape(shape)
        o += n
    return out


def pack(d: dict) -> np.ndarray:
    return np.concatenate([np.asarray(d[name]).reshape(-1) for name, _ in SHAPES])


def forward(p: np.ndarray, ctx: np.ndarray, bf16_ops: bool = True, dtype=np.float32):
    """Logits Z (n x 2871) plus the intermediates the backward pass needs."""
    w = unpack(np.asarray(p, dtype=dtype))
    rnd = (lambda a: bf16(a.astype(np.float32)).astype(dtype)) if bf16_ops else (lambda a: a)
    E, W1, W2 = rnd(w["E"]), rnd(w["W1"]), rnd(w["W2"])
    X = E[np.asarray(ctx, dtype=np.int64)].reshape(len(ctx), IN)
    H = np.maximum(X @ W1 + w["b1"], 0)
    Hb = rnd(H)
    Z = Hb @ W2 + w["b2"]
    return Z, dict(X=X, H=H, Hb=Hb, W1=W1, W2=W2, rnd=rnd)


def softmax(Z: np.ndarray) -> np.ndarray:
    e = np.exp(Z - Z.max(axis=1, keepdims=True))
    return e / e.sum(axis=1, keepdims=True)


def loss_grad(p: np.ndarray, ctx: np.ndarray, labels: np.ndarray, bf16_ops: bool = True, dtype=np.float32):
    """Mean cross entropy and its gradient with respect to the packed parameters."""
    ctx = np.asarray(ctx, dtype=np.int64)
    labels = np.asarray(labels, dtype=np.int64)
    B = len(ctx)
    Z, c = forward(p, ctx, bf16_ops, dtype)
    zmax = Z.max(axis=1, keepdims=True)
    lse = np.log(np.exp(Z - zmax).sum(axis=1)) + zmax[:, 0]
    loss = float(np.mean(lse - Z[np.arange(B), labels]))
    P = softmax(Z)
    dZ = P.copy()
    dZ[np.arange(B), labels] -= 1
    dZ /= B
    rnd = c["rnd"]
    dZb = rnd(dZ)
    g = {"b2": dZ.sum(axis=0), "W2": c["Hb"].T @ dZb}
    dH = (dZb @ c["W2"].T) * (c["H"] > 0)
    dHb = rnd(dH)
    g["b1"] = dH.sum(axis=0)
    g["W1"] = c["X"].T @ dHb
    dX = (dHb @ c["W1"].T).reshape(B, CTX, EMB)
    dE = np.zeros((N_TOK, EMB), dtype=dtype)
    np.add.at(dE, ctx.reshape(-1), dX.reshape(-1, EMB))
    g["E"] = dE
    return loss, pack(g).astype(dtype)


def adam(p, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One bias-corrected Adam update at step t (1-based); returns p, m, v."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed: int) -> np.ndarray:
    """The initial weights of sx_mlp_create: xorshift64 from seed * golden + 1."""
    mask = (1 << 64) - 1
    st = (seed * 0x9E3779B97F4A7C15 + 1) & mask
    n_e, n_w1, n_w2 = N_TOK * EMB, IN * HID, HID * N_OUT
    draws = np.empty(n_e + n_w1 + n_w2, dtype=np.float64)
    for i in range(len(draws)):
        st ^= (st << 13) & mask
        st ^= st >> 7
        st ^= (st << 17) & mask
        draws[i] = (st >> 40) / 16777216.0
    r = (draws.astype(np.float32) * np.float32(2) - np.float32(1)).astype(np.float32)
    w = {
        "E": np.float32(0.1) * r[:n_e],
        "W1": r[n_e:n_e + n_w1] / np.float32(np.sqrt(np.float32(IN))),
        "b1": np.zeros(HID, np.float32),
        "W2": r[n_e + n_w1:] / np.float32(np.sqrt(np.float32(HID))),
        "b2": np.zeros(N_OUT, np.float32),
    }
    return pack(w).astype(np.float32)


def batch_indices(step: int, batch: int, n: int, seed: int, philox) -> np.ndarray:
    """Examples of one training step: Philox4x32-10 with counter (step, row,
    0x5354, 0) and key (seed lo, seed hi); index = high word of r0 * n."""
    out = np.empty(batch, dtype=np.int64)
    for b in range(batch):
        r = philox((step, b, 0x5354, 0), (seed & 0xFFFFFFFF, seed >> 32))
        out[b] = (int(r[0]) * n) >> 32
    return out


def fold_pointers(probs: np.ndarray, ctx: np.ndarray) -> np.ndarray:
    """Instruction distribution of sx_mlp_proposal: the 2,863 instruction
    outputs plus, in position order and in fp32, each pointer's output on
    the instruction at its context position (padding adds nothing)."""
    probs = np.asarray(probs, dtype=np.float32)
    q = probs[:, :N_CLS].copy()
    for r in range(len(q)):
        for j in range(CTX):
            t = int(ctx[r][j])
            if t < N_CLS:
                q[r, t] = np.float32(q[r, t] + probs[r, N_CLS + j])
    return q


def proposal_weights(probs: np.ndarray, eps: float) -> np.ndarray:
    """Restatement of sx_proposal_weights: floor(w * (2^32 - 2863)) + 1, the
    remainder to the first largest entry."""
    probs = np.asarray(probs, dtype=np.float32).astype(np.float64)
    tot = np.cumsum(probs, axis=1)[:, -1:]  # sequential sum, as the C loop
    w = (1 - eps) * probs / tot + eps / N_CLS
    u = (np.floor(w * (2.0 ** 32 - N_CLS)) + 1).astype(np.uint64)
    for r in range(len(u)):
        best = int(np.argmax(u[r]))
        u[r, best] += np.uint64(2 ** 32) - u[r].sum()
    return u.astype(np.uint32)
