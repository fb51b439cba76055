"""float64 numpy restatement of the learned-tuning MLP (test infrastructure).

TEST INFRASTRUCTURE ONLY: the checker for the GPU MLP kernels
(csrc/mlp.hip).  No reference counterpart exists (the reference has no
learned component); this oracle is pinned by central finite differences of
its own loss (tests/test_oracle_mlp.py) and by bf16 rounding known answers.

Layout of the parameter vector (fp32 masters, NP entries):
    W1[64][16] b1[64] W2[64][64] b2[64] W3[2][64] b3[2]
Forward/backward use the parameters rounded to bf16 (round to nearest even);
loss L = 1/(2n) sum_i sum_o (y_io - t_io)^2.
"""
import numpy as np

NI, NH, NO = 16, 64, 2
OW1 = 0
OB1 = OW1 + NH * NI
OW2 = OB1 + NH
OB2 = OW2 + NH * NH
OW3 = OB2 + NH
OB3 = OW3 + NO * NH
NP = OB3 + NO


def bf16_round(p):
    """Round float32 values to bf16 precision (nearest, ties to even)."""
    u = np.ascontiguousarray(p, dtype=np.float32).view(np.uint32).astype(np.uint64)
    special = (u & 0x7F800000) == 0x7F800000
    r = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    r = np.where(special, u, r).astype(np.uint32)
    return r.view(np.float32)


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OW1:OB1].reshape(NH, NI), p[OB1:OW2], p[OW2:OB2].reshape(NH, NH), p[OB2:OW3],
            p[OW3:OB3].reshape(NO, NH), p[OB3:NP])


def forward(p, x, round_bf16=True):
    q = bf16_round(p) if round_bf16 else np.asarray(p, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(q)
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1.T + b1
    h1 = np.maximum(z1, 0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0)
    return h2 @ W3.T + b3, (x, z1, h1, z2, h2)


def loss_grad(p, x, t, round_bf16=True):
    """Loss and dL/dparams (w.r.t. the bf16-rounded parameters)."""
    q = bf16_round(p) if round_bf16 else np.asarray(p, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(q)
    y, (x, z1, h1, z2, h2) = forward(q, x, round_bf16=False)
    t = np.asarray(t, dtype=np.float64)
    n = len(x)
    L = float(((y - t) ** 2).sum() / (2 * n))
    dy = (y - t) / n
    g = np.zeros(NP)
    g[OW3:OB3] = (dy.T @ h2).ravel()
    g[OB3:NP] = dy.sum(0)
    dz2 = (dy @ W3) * (z2 > 0)
    g[OW2:OB2] = (dz2.T @ h1).ravel()
    g[OB2:OW3] = dz2.sum(0)
    dz1 = (dz2 @ W2) * (z1 > 0)
    g[OW1:OB1] = (dz1.T @ x).ravel()
    g[OB1:OW2] = dz1.sum(0)
    return L, g
