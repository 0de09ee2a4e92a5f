"""Pins on the CPU oracle, each by an independent route: defining properties,
closed forms, brute force, numpy's IEEE conversions, and an independent
textbook elastic scheme.  None of these re-types the oracle's own formula."""
import numpy as np
import pytest

from conftest import rel_l2
from ouc_amd import abi, velocity_models as vm


# ---- stencil coefficients and operators ------------------------------------

@pytest.mark.parametrize("N", range(1, 9))
def test_coeff_moment_conditions(oracle, N):
    """Staggered first-derivative coefficients satisfy
    sum_k c_k (2k-1)^(2m-1) = delta_{m,1}, m = 1..N (exactness to order 2N)."""
    c = oracle.coeffs(N)
    k = np.arange(1, N + 1)
    for m in range(1, N + 1):
        s = np.sum(c * (2.0 * k - 1) ** (2 * m - 1))
        target = 1.0 if m == 1 else 0.0
        # the coefficients are float32-rounded; scale tolerance by the term size
        tol = 1e-6 * np.sum(np.abs(c) * (2.0 * k - 1) ** (2 * m - 1))
        assert abs(s - target) <= tol, (N, m, s)


def test_coeff_reference_literals(oracle):
    """N = 6 uses the reference's literals (P-S-seperating.cu:221), which match
    the exact solution of the moment system to float precision."""
    c = oracle.coeffs(6)
    lit = np.array([1.2213364, -9.6931458e-2, 1.7447662e-2, -2.9672895e-3, 3.5900540e-4, -2.1847812e-5],
                   np.float32).astype(np.float64)
    assert np.array_equal(c, lit)
    k = np.arange(1, 7)
    A = np.array([(2.0 * k - 1) ** (2 * m - 1) for m in range(1, 7)])
    exact = np.linalg.solve(A, np.eye(6)[0])
    assert np.max(np.abs(c - exact) / np.abs(exact)) < 3e-7


@pytest.mark.parametrize("N", [5, 6])
@pytest.mark.parametrize("kind", [0, 1, 2, 3])
def test_staggered_operator_polynomial_exact(oracle, N, kind):
    """D- at cell i is f'(x_i - h/2), D+ is f'(x_i + h/2), exactly for
    polynomials of degree <= 2N-1 (up to rounding of the float coefficients)."""
    h = 5.0
    n = 4 * N + 8
    rng = np.random.default_rng(N + 10 * kind)
    a = rng.standard_normal(2 * N)  # degree 2N-1
    xs = (np.arange(n) - n / 2) * h / 40.0
    poly = np.polynomial.Polynomial(a)
    dpoly = poly.deriv()
    along_x = kind in (0, 1)
    shift = -0.5 if kind in (0, 2) else 0.5
    f1 = poly(xs)
    f = np.tile(f1[:, None], (1, n)) if along_x else np.tile(f1[None, :], (n, 1))
    out = oracle.dop(N, kind, f, h / 40.0)
    expect1 = dpoly(xs + shift * h / 40.0)
    for i in range(N, n - N):
        for j in range(N, n - N):
            idx = i if along_x else j
            e = expect1[idx]
            assert abs(out[i, j] - e) <= 2e-6 * (1 + abs(e)) * np.max(np.abs(f1)), (i, j, out[i, j], e)


# ---- Ricker, CPML, padding, mute ---------------------------------------------

def test_ricker_reference_defaults(oracle):
    """fm = 30 Hz, dt = 1 ms: t0 = ceil(1/(fm dt)) = 34, peak 1 at t0,
    symmetric, zero mean, zero crossings at +-1/(pi fm sqrt 2)."""
    s, t0 = oracle.ricker(200, 30.0, np.float32(0.001))
    assert t0 == 34
    assert s[t0] == pytest.approx(1.0, abs=1e-12)
    for k in range(1, 30):
        assert s[t0 + k] == pytest.approx(s[t0 - k], rel=1e-12, abs=1e-15)
    assert abs(s.sum()) < 1e-3  # sampled zero-mean wavelet
    # sign change between the samples around tau = 1/(pi fm sqrt2) = 7.5 ms
    tc = 1.0 / (np.pi * 30.0 * np.sqrt(2.0)) / 1e-3
    lo, hi = int(np.floor(tc)), int(np.ceil(tc))
    assert s[t0 + lo] > 0 > s[t0 + hi]


def test_ricker_is_gaussian_second_derivative(oracle):
    """r(tau) = -(1/(2 pi^2 fm^2)) d^2/dtau^2 exp(-pi^2 fm^2 tau^2)."""
    fm, dt = 20.0, np.float32(0.0004)
    s, t0 = oracle.ricker(400, fm, dt)
    pi = 3.1415926
    tau = (np.arange(400) - t0) * float(dt)
    g = lambda x: np.exp(-(pi * fm * x) ** 2)
    e = 1e-6
    d2 = (g(tau + e) - 2 * g(tau) + g(tau - e)) / e ** 2
    r = -d2 / (2 * pi ** 2 * fm ** 2)
    assert np.max(np.abs(r - s)) < 1e-4


def test_padding_is_edge_replication(oracle):
    """xiangbian() pads every side by the nearest interior cell."""
    nx, nz, L = 17, 13, 6
    rng = np.random.default_rng(1)
    vp = rng.uniform(2000, 3000, (nx, nz)).astype(np.float32)
    vs = rng.uniform(1000, 1500, (nx, nz)).astype(np.float32)
    p = abi.make_params(nx, nz, L, 3, 10, 5, 5, 1e-3)
    vpp, vsp, ax, az = oracle.setup(p, vp, vs)
    assert np.array_equal(vpp, np.pad(vp.astype(np.float64), L, mode="edge"))
    assert np.array_equal(vsp, np.pad(vs.astype(np.float64), L, mode="edge"))


def test_cpml_profile(oracle):
    """4th-power profile scaled by vp, zero on the interior, log10(1/R) scale,
    strict far-side test (i = Xn-L undamped)."""
    nx, nz, L = 30, 20, 10
    vp = np.full((nx, nz), 3000.0, np.float32)
    R = 1e-3
    p = abi.make_params(nx, nz, L, 4, 10, 5, 5, R)
    _, _, ax, az = oracle.setup(p, vp, vp / 2)
    Xn, Zn = nx + 2 * L, nz + 2 * L
    assert np.all(ax[L:Xn - L + 1, :] == 0)
    assert np.all(az[:, L:Zn - L + 1] == 0)
    d0 = 3.0 * 1.5 * 3000.0 / L  # log10(1/1e-3) = 3
    assert ax[0, 5] == pytest.approx(d0, rel=1e-14)
    for i in range(1, L):
        # ratio pins the exponent 4 and the (L - i)/L argument
        assert ax[i, 5] / ax[0, 5] == pytest.approx(((L - i) / L) ** 4, rel=1e-12)
        assert ax[Xn - i, 5] == pytest.approx(ax[i, 5], rel=1e-12)  # mirror image
    assert np.array_equal(ax.T[:Zn, :Xn][:, :].shape, (Zn, Xn))


def test_mute_threshold_properties(oracle):
    """remove(): threshold 2 t0 + dist dh/(dt vp) - symmetric about the source,
    equal to 2 t0 at zero offset, growing by dh/(dt vp) per cell far away."""
    nx, nz = 64, 32
    vp = np.full((nx, nz), 2500.0, np.float32)
    p = abi.make_params(nx, nz, 10, 4, 500, 5, 5, np.float32(0.0004))
    shot = dict(sx=30, sz=0, rec_z=0, fm=20.0)
    thr = oracle.mute_thresholds(p, vp, vp / 2, shot)
    _, t0 = oracle.ricker(10, 20.0, np.float32(0.0004))
    assert thr[30] == pytest.approx(2 * t0)
    assert thr[25] == pytest.approx(thr[35])
    step = 5.0 / (float(np.float32(0.0004)) * 2500.0)
    assert thr[60] - thr[59] == pytest.approx(step, rel=1e-6)


# ---- FP16 strip codec -------------------------------------------------------

def test_f16_round_to_nearest_even_matches_ieee(oracle):
    """The oracle's binary16 conversion equals numpy's IEEE one, including ties,
    subnormals, overflow and signed zero."""
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32) * np.float32(10.0) ** rng.integers(-9, 6, 20000),
        np.array([0.0, -0.0, 65504.0, 65520.0, 65519.996, 1e-8, 5.96e-8, 2.98e-8, 3e-8, -6.1e-5, 6.1035156e-05,
                  1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 2049.0, 2051.0, np.inf, -np.inf], np.float32),
    ])
    h = oracle.f32_to_f16(x)
    with np.errstate(over="ignore"):
        ieee = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(h, ieee)


def test_chunk_codec_scale_and_error(oracle):
    rng = np.random.default_rng(6)
    x = (rng.standard_normal(64 * 50) * 10.0 ** rng.uniform(-12, 3)).astype(np.float32)
    x[64 * 7:64 * 8] = 0  # an all-zero chunk
    h, e = oracle.fp16_encode(x)
    y = oracle.fp16_decode(h, e)
    for c in range(50):
        xc = x[64 * c:64 * c + 64]
        m = np.max(np.abs(xc))
        if m == 0:
            assert np.all(y[64 * c:64 * c + 64] == 0)
            continue
        scaled = m * 2.0 ** (-int(e[c]))
        assert 2 ** 14 <= scaled < 2 ** 15
        err = np.abs(y[64 * c:64 * c + 64] - xc)
        # half-ulp of binary16 at the value, or the subnormal floor, times the scale
        bound = np.maximum(np.abs(xc) * 2.0 ** -11, 2.0 ** -25 * 2.0 ** int(e[c]))
        assert np.all(err <= bound * (1 + 1e-6))


def test_predictive_codec_error_and_running_sum(oracle):
    """Successive slots of smooth signals through the predictive FP16 codec:
    (1) the first slot is the single-slot codec (encoder state starts at
    zero); (2) each decoded value is off by about 2^-12 of the one-step change,
    not of the value: the median relative error is below 3e-5, against about
    1e-4 for rounding each value to binary16; (3) the running sum of the
    decoding errors stays bounded (it telescopes to the last fed-back error),
    at least 10x below the random walk of independent rounding."""
    rng = np.random.default_rng(11)
    T, n = 400, 64 * 6 + 5
    t = np.arange(T)[:, None] * 0.0004
    f = rng.uniform(10, 40, n)[None, :]
    amp = 10.0 ** rng.uniform(-6, 2, n)[None, :]
    x = (amp * np.sin(2 * np.pi * f * t + rng.uniform(0, 6, n)[None, :])).astype(np.float32)
    h, e, y = oracle.fp16_encode(x, decoded=True)
    h1, e1 = oracle.fp16_encode(x[0])
    assert np.array_equal(h[0], h1) and np.array_equal(e[0], e1)
    x64 = x.astype(np.float64)
    err = y.astype(np.float64) - x64
    big = np.abs(x64) > 1e-3 * np.max(np.abs(x64), axis=0)  # away from zero crossings
    ind = np.stack([oracle.fp16_decode(*oracle.fp16_encode(x[k])) for k in range(T)]).astype(np.float64) - x64
    med_p = np.median(np.abs(err[big]) / np.abs(x64[big]))
    med_i = np.median(np.abs(ind[big]) / np.abs(x64[big]))
    print(f"[predictive codec] median relative error {med_p:.2e} (independent rounding {med_i:.2e})")
    assert med_p < 3e-5 and med_i > 5e-5
    scale = np.max(np.abs(x64), axis=0)
    walk_p = np.max(np.abs(np.cumsum(err, axis=0)), axis=0) / scale
    walk_i = np.max(np.abs(np.cumsum(ind, axis=0)), axis=0) / scale
    print(f"[predictive codec] running-sum max / amplitude: median {np.median(walk_p):.2e} "
          f"(independent rounding {np.median(walk_i):.2e})")
    assert np.median(walk_p) < 0.1 * np.median(walk_i)


def test_predictive_fp16_strips_image_error(oracle):
    """FP64 oracle, FP16 strips against FP32 strips, same shot: the
    predictive codec keeps the image within 1e-4 below row 8 (rounding every
    slot to binary16 on its own gives about 3e-2 on this case, error feedback
    alone about 4e-4)."""
    from ouc_amd import velocity_models as vm
    nx, nz, nt = 160, 96, 900
    vp, vs, _ = vm.layered_random(nx, nz, seed=3)
    p = abi.make_params(nx, nz, 24, 5, nt, 5.0, 5.0, 0.0004, R=1e-3, strip_precision=abi.STRIP_FP32)
    shot = dict(sx=80, sz=0, rec_z=0, fm=22.0, mute=1)
    ref = oracle.rtm_shot(p, vp, vs, shot, 64)
    p.strip_precision = abi.STRIP_FP16
    got = oracle.rtm_shot(p, vp, vs, shot, 64)
    for k in ["img_pp", "img_ps"]:
        err = np.linalg.norm(got[k][:, 8:] - ref[k][:, 8:]) / np.linalg.norm(ref[k][:, 8:])
        print(f"[fp16 strips, FP64 oracle] {k} below row 8 rel-L2 {err:.2e}")
        assert err < 1e-4, (k, err)


# ---- strip ring ---------------------------------------------------------------

@pytest.mark.parametrize("N", [3, 5])
def test_ring_is_exactly_the_read_band(oracle, N):
    """Ring cells = the frame [L-N, Xn-L+N)^2 minus [L+N, Xn-L-N)^2, each once."""
    nx, nz, L = 40, 24, 8
    p = abi.make_params(nx, nz, L, N, 10, 5, 5, 1e-3)
    ij = oracle.ring(p)
    Xn, Zn = nx + 2 * L, nz + 2 * L
    got = set(map(tuple, ij.tolist()))
    assert len(got) == len(ij) == 4 * N * (nx + nz)
    want = {(i, j) for i in range(L - N, Xn - L + N) for j in range(L - N, Zn - L + N)
            if not (L + N <= i < Xn - L - N and L + N <= j < Zn - L - N)}
    assert got == want


# ---- physics invariants -----------------------------------------------------

def test_homogeneous_explosion_has_no_s_wave(oracle):
    """Constant vp, vs, explosive source: the S part stays at rounding level
    (curl of a gradient vanishes; C - E and B - A cancel) until the CPML."""
    nx, nz, L, N, nt = 120, 120, 20, 5, 150
    vp = np.full((nx, nz), 3000.0, np.float32)
    vs = np.full((nx, nz), 1700.0, np.float32)
    p = abi.make_params(nx, nz, L, N, nt, 5, 5, np.float32(0.0004), R=1e-3)
    seis = oracle.forward_shot(p, vp, vs, dict(sx=60, sz=60, rec_z=70, fm=25.0), 64)
    assert np.max(np.abs(seis[2])) > 0
    ratio = np.max(np.abs(seis[4:6])) / np.max(np.abs(seis[2:4]))
    assert ratio < 1e-12, ratio


def _textbook_displacement(nx, nz, L, N, nt, h, dt, vp, vs, src, si, sj, rz, c):
    """Independent constant-density elastic scheme on the same staggered grid:
    tau_xx = (l+2m) dux/dx + l duz/dz, tau_zz = l dux/dx + (l+2m) duz/dz,
    tau_xz = m (dux/dz + duz/dx) with l = vp^2 - 2 vs^2, m = vs^2 at the cell,
    u'' = div tau; explosive source on both normal stresses.  No CPML: valid
    while the wavefield has not reached the pad."""
    Xn, Zn = nx + 2 * L, nz + 2 * L
    pv = np.pad(vp.astype(np.float64), L, mode="edge") ** 2
    sv = np.pad(vs.astype(np.float64), L, mode="edge") ** 2
    lam = pv - 2 * sv

    def Dm(f, ax):
        out = np.zeros_like(f)
        for m in range(1, N + 1):
            out += c[m - 1] * (np.roll(f, -(m - 1), ax) - np.roll(f, m, ax))
        return out / h

    def Dp(f, ax):
        out = np.zeros_like(f)
        for m in range(1, N + 1):
            out += c[m - 1] * (np.roll(f, -m, ax) - np.roll(f, m - 1, ax))
        return out / h

    ux = np.zeros((Xn, Zn)); uz = np.zeros((Xn, Zn)); uxp = ux.copy(); uzp = uz.copy()
    mask = np.zeros((Xn, Zn), bool)
    mask[N:Xn - N, N:Zn - N] = True
    rec = np.zeros((nx, nt))
    for t in range(nt):
        exx = Dm(ux, 0); ezz = Dm(uz, 1); exz = Dp(ux, 1); ezx = Dp(uz, 0)
        txx = pv * exx + lam * ezz
        tzz = lam * exx + pv * ezz
        txz = sv * (exz + ezx)
        txx[si, sj] += src[t]
        tzz[si, sj] += src[t]
        for f in (txx, tzz, txz):
            f[~mask] = 0
        ax_ = Dp(txx, 0) + Dm(txz, 1)
        az_ = Dm(txz, 0) + Dp(tzz, 1)
        nxu = 2 * ux - uxp + dt * dt * ax_
        nzu = 2 * uz - uzp + dt * dt * az_
        nxu[~mask] = 0; nzu[~mask] = 0
        uxp, uzp, ux, uz = ux, uz, nxu, nzu
        rec[:, t] = ux[L:L + nx, rz]
    return rec


def test_p_plus_s_is_textbook_elastic(oracle):
    """Vp + Vs of the split scheme equals an independent textbook elastic
    update (P-S-seperating.cu:291-294 summed), before the CPML is reached."""
    nx, nz, L, N, nt = 80, 80, 12, 5, 90
    vp, vs, _ = vm.layered_random(nx, nz, seed=11)
    h, dt = 5.0, float(np.float32(0.0004))
    p = abi.make_params(nx, nz, L, N, nt, h, h, dt, R=1e-3)
    shot = dict(sx=40, sz=40, rec_z=30, fm=30.0)
    seis = oracle.forward_shot(p, vp, vs, shot, 64)
    src, _ = oracle.ricker(nt, 30.0, np.float32(dt))
    # dt enters as the float32 value widened to double, as in the oracle
    ref = _textbook_displacement(nx, nz, L, N, nt, float(np.float32(h)), float(np.float32(dt)), vp, vs, src, 40 + L,
                                 40 + L, 30 + L, oracle.coeffs(N))
    assert np.max(np.abs(ref)) > 0
    assert rel_l2(seis[0], ref) < 1e-10


def test_exact_reconstruction_reverses_forward(oracle):
    """EXACT semantics: the interior-only reverse leapfrog reproduces every
    forward snapshot (P-S-seperating.cu:890-1042 with the strips aligned)."""
    nx, nz, nt = 48, 40, 260
    vp, vs, _ = vm.layered_random(nx, nz, seed=2)
    p = abi.make_params(nx, nz, 12, 5, nt, 5, 5, np.float32(0.0004), R=1e-3, semantics=abi.SEM_EXACT)
    err = oracle.recon_error(p, vp, vs, dict(sx=24, sz=0, rec_z=0, fm=25.0), 64)
    assert err.max() < 1e-10, err.max()
    p.semantics = abi.SEM_FAITHFUL
    err_f = oracle.recon_error(p, vp, vs, dict(sx=24, sz=0, rec_z=0, fm=25.0), 64)
    # the reference's one-step-late velocity strips (P-S-seperating.cu:2272 vs 2337)
    assert err_f.max() > 1e-3
