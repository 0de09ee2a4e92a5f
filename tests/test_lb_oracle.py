"""CPU checks of the learned-boundary oracle (oracle/lb_net.py) and of the
host-side data pipeline (ouc_amd/learned.py)."""
import numpy as np
import pytest

from oracle import lb_net


@pytest.fixture(scope="module")
def problem():
    rng = np.random.default_rng(3)
    w = (0.3 * rng.standard_normal(lb_net.N_WEIGHTS)).astype(np.float32)
    x = rng.standard_normal((2, 5, 9, 12))
    y = rng.standard_normal((2, 2, 9, 12))
    m = (rng.random((2, 9, 12)) > 0.3).astype(float)
    return w, x, y, m


def test_gradient_matches_finite_differences(problem):
    w, x, y, m = problem
    w64 = w.astype(np.float64)
    _, g = lb_net.loss_grad(x, y, m, w64, rounded=False)
    rng = np.random.default_rng(4)
    for k in rng.choice(lb_net.N_WEIGHTS, 24, replace=False):
        e = np.zeros_like(w64)
        e[k] = 1e-6
        lp, _ = lb_net.loss_grad(x, y, m, w64 + e, rounded=False)
        lm, _ = lb_net.loss_grad(x, y, m, w64 - e, rounded=False)
        fd = (lp - lm) / 2e-6
        assert abs(g[k] - fd) <= 1e-6 * max(1.0, abs(fd)), (k, g[k], fd)


def test_composed_filter_equals_layers_away_from_edges(problem):
    w, x, _, _ = problem
    K = lb_net.composed(w)
    out = lb_net.forward(x, w)[-1]
    xp = np.pad(x, ((0, 0), (0, 0), (3, 3), (3, 3)))
    o2 = np.zeros_like(out)
    for dy in range(7):
        for dx in range(7):
            o2 += np.einsum("mchw,oc->mohw", xp[:, :, dy:dy + 9, dx:dx + 12], K[:, :, dy, dx])
    assert np.abs(o2 - out)[:, :, 2:-2, 2:-2].max() < 1e-10 * np.abs(out).max()
    # one cell from the edge the layered net's zero padding differs: the check has teeth
    assert np.abs(o2 - out)[:, :, 1:-1, 1:-1].max() > 1e-3


def test_bf16_rounding_ties_to_even():
    v = np.array([1.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 1 + 2 ** -7, -(1 + 2 ** -8 + 2 ** -20), 3.0e38], np.float32)
    r = lb_net.bf16_round(v)
    assert list(r[:4]) == [1.0, 1.0, 1.015625, 1.0078125]
    assert r[4] == np.float32(-1.0078125)  # above the tie: away from zero
    assert np.isfinite(r[5])


def mirrored_fields(rng, X, Z):
    """Five channel arrays (D_x, D_z, txx, tzz, txz) that are the x-mirror of
    themselves on the staggered grid."""
    from ouc_amd.learned import SX, XS
    ch = []
    for c in range(5):
        a = rng.standard_normal((X, Z))
        n = X - SX[c]
        b = np.zeros((X, Z))
        # value at i equals XS * value at X-1-SX-i
        b[:n] = 0.5 * (a[:n] + XS[c] * a[X - 1 - SX[c] - np.arange(n)])
        ch.append(b)
    return ch


def test_side_frames_agree_under_mirror_symmetry():
    from ouc_amd import learned as LB
    rng = np.random.default_rng(5)
    X, Z = 20, 14
    ch = mirrored_fields(rng, X, Z)
    left, right = LB.to_side(ch, 0), LB.to_side(ch, 1)
    for c in range(5):
        n = X - LB.SX[c]
        np.testing.assert_allclose(left[c][:n], right[c][:n], atol=1e-12)
    # transposing the fields maps the x-low frame onto the z-low frame
    chT = [ch[LB.TSWAP[c]].T for c in range(5)]
    top = LB.to_side(chT, 2)
    for c in range(5):
        np.testing.assert_allclose(top[c], left[c], atol=1e-12)


def test_windows_mask_only_the_ring_row():
    from ouc_amd import learned as LB
    rng = np.random.default_rng(6)
    X = Z = 40
    V = rng.standard_normal((2, X, Z)).astype(np.float32)
    Vm = rng.standard_normal((2, X, Z)).astype(np.float32)
    Vn = rng.standard_normal((2, X, Z)).astype(np.float32)
    p2 = np.full((X, Z), 3000.0 ** 2)
    q2 = np.full((X, Z), 1700.0 ** 2)
    x, y, m = LB.windows_from_state(Vm, V, Vn, p2, q2, 5.0, 5.0, 4e-4, 8, 0.0, rng)
    assert x.shape[1:] == (5, 7, 70) and y.shape[1:] == (2, 7, 70) and m.shape[1:] == (7, 70)
    assert m[:, :3].sum() == 0 and m[:, 4:].sum() == 0 and m[:, 3].sum() > 0
    assert np.all(x[:, :, :3] == 0)  # rows outside the grid
    # every ring cell of each field is a target exactly once across the four sides
    # (corners by the smallest depth; the ghost lines never)
    nring = 0
    for f in range(2):
        for i in range(X):
            for j in range(Z):
                ghost = (f == 0 and i == X - 1) or (f == 1 and j == Z - 1)
                live = (1 <= i < X - 2 and 1 <= j < Z - 1) if f == 0 else (1 <= i < X - 1 and 1 <= j < Z - 2)
                nring += (not ghost and not live)
    # a window pixel is kept only when both fields' cells belong to that side
    assert m.sum() <= nring / 2 + 4
