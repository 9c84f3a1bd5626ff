"""FP64 CPU restatement of the surrogate network (TEST INFRASTRUCTURE).

nin -> H -> H -> nout with tanh hidden units; loss = mean over samples and outputs of
1/2 (y - t)^2.  With `bf16=True` every matrix-product operand is rounded to bfloat16
(round to nearest even of its FP32 value) exactly where the GPU kernels round, so the
GPU's BF16-MFMA gradients can be compared to it; with `bf16=False` it is plain FP64 and
is pinned by central finite differences in tests/test_mlp.py.  No reference code exists
for this part (the reference has no learned component).
"""
import numpy as np


def bf16(a):
    a32 = np.ascontiguousarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _ident(a):
    return np.asarray(a, dtype=np.float64)


def unpack(theta, nin, H, nout):
    """Split the flat parameter vector (order W1, b1, W2, b2, W3, b3)."""
    sizes = [nin * H, H, H * H, H, H * nout, nout]
    out, o = [], 0
    for s in sizes:
        out.append(np.asarray(theta[o:o + s], dtype=np.float64))
        o += s
    W1, b1, W2, b2, W3, b3 = out
    return W1.reshape(nin, H), b1, W2.reshape(H, H), b2, W3.reshape(H, nout), b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def forward(theta, X, nin, H, nout, bf=False):
    r = bf16 if bf else _ident
    W1, b1, W2, b2, W3, b3 = unpack(theta, nin, H, nout)
    H1 = np.tanh(r(X) @ r(W1) + b1)
    H2 = np.tanh(r(H1) @ r(W2) + b2)
    Y = r(H2) @ r(W3) + b3
    return H1, H2, Y


def loss_and_grad(theta, X, T, nin, H, nout, bf=False):
    r = bf16 if bf else _ident
    W1, b1, W2, b2, W3, b3 = unpack(theta, nin, H, nout)
    H1, H2, Y = forward(theta, X, nin, H, nout, bf)
    n = X.shape[0]
    E = Y - T
    loss = 0.5 * np.sum(E * E) / (n * nout)
    dY = E / (n * nout)
    gW3 = r(H2).T @ r(dY)
    gb3 = dY.sum(axis=0)
    dZ2 = (r(dY) @ r(W3).T) * (1.0 - H2 * H2)
    gW2 = r(H1).T @ r(dZ2)
    gb2 = dZ2.sum(axis=0)
    dZ1 = (r(dZ2) @ r(W2).T) * (1.0 - H1 * H1)
    gW1 = r(X).T @ r(dZ1)
    gb1 = dZ1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam_step(theta, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v
