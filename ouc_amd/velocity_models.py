"""Synthetic elastic models and acquisition used by the tests and the benchmark.

All models are x-major (nx, nz) float32 arrays on the interior grid.
"""
import numpy as np

from . import abi


def reference_three_layer(nx=200, nz=200):
    """The reference's built-in model (P-S-seperating.cu:1999-2040): vp
    3400/3800/4500 on z-cells 0-49/50-99/100+, vs 2900/3800 split at 100."""
    z = np.arange(nz)
    vp = np.where(z < 50, 3400.0, np.where(z < 100, 3800.0, 4500.0))
    vs = np.where(z < 100, 2900.0, 3800.0)
    return (np.tile(vp, (nx, 1)).astype(np.float32), np.tile(vs, (nx, 1)).astype(np.float32))


def two_layer(nx=200, nz=200, interface=100, vp_top=2000.0, vp_bot=3000.0, ratio=1.732):
    """Forward-modelling case: vp 2000/3000 split at z = interface, vs = vp/1.732."""
    z = np.arange(nz)
    vp = np.tile(np.where(z < interface, vp_top, vp_bot), (nx, 1)).astype(np.float32)
    vs = (vp / np.float32(ratio)).astype(np.float32)
    return vp, vs


def _gaussian_smooth(field, sigma_cells):
    """Separable Gaussian filter with reflecting edges, sigma in cells."""
    r = int(np.ceil(4 * sigma_cells))
    x = np.arange(-r, r + 1)
    k = np.exp(-0.5 * (x / sigma_cells) ** 2)
    k /= k.sum()
    out = field
    for axis in (0, 1):
        pad = [(0, 0), (0, 0)]
        pad[axis] = (r, r)
        f = np.pad(out, pad, mode="reflect")
        out = np.apply_along_axis(lambda v: np.convolve(v, k, mode="valid"), axis, f)
    return out


def layered_random(nx, nz, seed, nlayers=6, vp_min=1800.0, vp_max=4000.0, vs_ratio=1.8, sigma=0.03,
                   corr_len_m=50.0, dx=5.0):
    """Flat layers vp_min..vp_max, vs = vp/vs_ratio, multiplied by (1 + delta),
    delta = sigma * unit-RMS(Gaussian-smoothed white noise, correlation length
    corr_len_m).  rho (Gardner, 310 vp^0.25) is returned but unused by the physics."""
    rng = np.random.default_rng(seed)
    z = np.arange(nz)
    layer = np.minimum(z * nlayers // nz, nlayers - 1)
    vp1 = vp_min + (vp_max - vp_min) * layer / max(nlayers - 1, 1)
    vp = np.tile(vp1, (nx, 1)).astype(np.float64)
    noise = rng.standard_normal((nx, nz))
    sm = _gaussian_smooth(noise, corr_len_m / dx / np.sqrt(2.0))
    sm /= np.sqrt(np.mean(sm ** 2))
    vp = vp * (1.0 + sigma * sm)
    vp = vp.astype(np.float32)
    vs = (vp / np.float32(vs_ratio)).astype(np.float32)
    rho = (310.0 * vp.astype(np.float64) ** 0.25).astype(np.float32)
    return vp, vs, rho


def random_layered(nx, nz, seed, vp_lo=1500.0, vp_hi=4500.0, nlayers=None, vs_ratio=1.8):
    """Random flat layers (thickness and vp drawn from seed), for learned-boundary data."""
    rng = np.random.default_rng(seed)
    if nlayers is None:
        nlayers = int(rng.integers(2, 7))
    cuts = np.sort(rng.choice(np.arange(1, nz), size=nlayers - 1, replace=False))
    vps = np.sort(rng.uniform(vp_lo, vp_hi, size=nlayers))
    z = np.arange(nz)
    idx = np.searchsorted(cuts, z, side="right")
    vp = np.tile(vps[idx], (nx, 1)).astype(np.float32)
    vs = (vp / np.float32(vs_ratio)).astype(np.float32)
    return vp, vs


# ---- standard configurations ------------------------------------------------

def cfg_forward_200():
    """Single-shot elastic forward modelling: 200 x 200 at 5 m, N = 5, 20-cell
    CPML at R = 1e-4, two-layer model, Ricker 25 Hz at (100, 5), dt 0.5 ms,
    2000 steps, 200 receivers at z = 5."""
    p = abi.make_params(200, 200, 20, 5, 2000, 5.0, 5.0, 0.0005, R=1e-4)
    vp, vs = two_layer(200, 200)
    shot = dict(sx=100, sz=5, rec_z=5, fm=25.0, mute=0)
    return p, vp, vs, shot


def cfg_rtm_1024(nt=4000, seed=1234, **kw):
    """Plain single-shot RTM: 1024 x 512 at 5 m, N = 5, 30-cell CPML, six layers
    1800-4000 m/s with a 3 % random perturbation, Ricker 20 Hz, dt 0.4 ms."""
    p = abi.make_params(1024, 512, 30, 5, nt, 5.0, 5.0, 0.0004, R=1e-3, **kw)
    vp, vs, rho = layered_random(1024, 512, seed)
    shot = dict(sx=512, sz=0, rec_z=0, fm=20.0, mute=1)
    return p, vp, vs, shot


def cfg_rtm_2048(nt=6000, seed=2048, nshots=64, **kw):
    """Batched RTM: 64 shots on 2048 x 1024 at 5 m, sources x = 16 + 32 s at z = 0,
    N = 5, 30-cell CPML, same model family, Ricker 20 Hz, dt 0.4 ms."""
    p = abi.make_params(2048, 1024, 30, 5, nt, 5.0, 5.0, 0.0004, R=1e-3, **kw)
    vp, vs, rho = layered_random(2048, 1024, seed)
    shots = [dict(sx=16 + 32 * s, sz=0, rec_z=0, fm=20.0, mute=1) for s in range(nshots)]
    return p, vp, vs, shots
