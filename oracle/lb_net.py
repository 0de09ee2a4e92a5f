"""FP64 CPU restatement of the learned-boundary network (test infrastructure).

Only the tests and the smoke check use this module, as the checker of the HIP
training and inference kernels; the product never calls it.  The network has
no counterpart in the reference: it follows the definition in
ouc_amd/csrc/learned.hip (three 3x3 cross-correlations with zero padding,
5 -> 16 -> 16 -> 2 channels, no bias, identity activations; weights rounded to
BF16 before use; masked mean-squared error).

Pins: the gradient is checked against central finite differences of the loss
(tests/test_lb_oracle.py), and the composed 7x7 filter against the layered
evaluation on random inputs (linearity makes them equal away from the window
edges); the BF16 rounding against numpy's own float32 bit arithmetic on the
round-to-nearest-even ties.
"""
import numpy as np

SHAPES = ((16, 5), (16, 16), (2, 16))
N_WEIGHTS = sum(o * i * 9 for o, i in SHAPES)


def bf16_round(w):
    """Round float32 values to bfloat16 (nearest, ties to even), as float32."""
    u = np.ascontiguousarray(w, np.float32).view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32)


def split(w):
    w = np.asarray(w, np.float64).ravel()
    out, o = [], 0
    for co, ci in SHAPES:
        out.append(w[o:o + co * ci * 9].reshape(co, ci, 3, 3))
        o += co * ci * 9
    return out


def conv(x, k):
    """x [M][Ci][H][W], k [Co][Ci][3][3] -> [M][Co][H][W], zero padding."""
    M, Ci, H, W = x.shape
    xp = np.zeros((M, Ci, H + 2, W + 2))
    xp[:, :, 1:H + 1, 1:W + 1] = x
    out = np.zeros((M, k.shape[0], H, W))
    for ky in range(3):
        for kx in range(3):
            out += np.einsum("mchw,oc->mohw", xp[:, :, ky:ky + H, kx:kx + W], k[:, :, ky, kx])
    return out


def conv_t(g, k):
    """Adjoint of conv with respect to its input."""
    M, Co, H, W = g.shape
    out = np.zeros((M, k.shape[1], H + 2, W + 2))
    for ky in range(3):
        for kx in range(3):
            out[:, :, ky:ky + H, kx:kx + W] += np.einsum("mohw,oc->mchw", g, k[:, :, ky, kx])
    return out[:, :, 1:H + 1, 1:W + 1]


def conv_w(x, g):
    """Gradient of sum(g * conv(x, k)) with respect to k."""
    M, Ci, H, W = x.shape
    xp = np.zeros((M, Ci, H + 2, W + 2))
    xp[:, :, 1:H + 1, 1:W + 1] = x
    dk = np.zeros((g.shape[1], Ci, 3, 3))
    for ky in range(3):
        for kx in range(3):
            dk[:, :, ky, kx] = np.einsum("mohw,mchw->oc", g, xp[:, :, ky:ky + H, kx:kx + W])
    return dk


def forward(x, w, rounded=True):
    ks = split(bf16_round(w) if rounded else w)
    a = np.asarray(x, np.float64)
    acts = [a]
    for k in ks:
        a = conv(a, k)
        acts.append(a)
    return acts


def loss_grad(x, y, mask, w, rounded=True):
    """Masked MSE over both output channels and its gradient (FP64)."""
    ks = split(bf16_round(w) if rounded else w)
    acts = forward(x, w, rounded)
    mask = np.asarray(mask, np.float64)[:, None]
    cnt = mask.sum() * 2
    d = acts[-1] - y
    loss = float((mask * d * d).sum() / cnt)
    g = 2 * mask * d / cnt
    grads = [None] * 3
    for l in (2, 1, 0):
        grads[l] = conv_w(acts[l], g)
        if l:
            g = conv_t(g, ks[l])
    return loss, np.concatenate([gg.ravel() for gg in grads])


def composed(w, rounded=True):
    """The 7x7 filter [2][5][7][7] equal to the three layers (cross-correlation
    taps: out[y][x] = sum K[dy][dx] in[y + dy - 3][x + dx - 3])."""
    k1, k2, k3 = split(bf16_round(w) if rounded else w)
    A = np.zeros((16, 5, 5, 5))
    for y2 in range(3):
        for x2 in range(3):
            for y1 in range(3):
                for x1 in range(3):
                    A[:, :, y2 + y1, x2 + x1] += np.einsum("ab,bc->ac", k2[:, :, y2, x2], k1[:, :, y1, x1])
    K = np.zeros((2, 5, 7, 7))
    for y3 in range(3):
        for x3 in range(3):
            K[:, :, y3:y3 + 5, x3:x3 + 5] += np.einsum("ab,bcyx->acyx", k3[:, :, y3, x3], A)
    return K
