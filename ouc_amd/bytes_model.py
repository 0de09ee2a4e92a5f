"""Algorithmic (compulsory) HBM bytes of the engines' kernels.

Each FP32 operand a kernel must read or write is counted once per cell and
launch; stencil neighbours are reuse, not traffic.  The reference's own
two-pass form (forward_s + forward_u, P-S-seperating.cu:205-384) is the
canonical model: 120 B per interior and 232 B per CPML cell-update.
"""
import os

from . import abi

CANONICAL_INTERIOR = 120.0
CANONICAL_CPML = 232.0


def _cells(p):
    Xn, Zn = p.nx + 2 * p.npml, p.nz + 2 * p.npml
    upd = (Xn - 2 * p.half_order) * (Zn - 2 * p.half_order)
    interior = p.nx * p.nz
    return Xn, Zn, upd, interior


def plain_velocity_bytes(p):
    """k_velocity over the update region: reads theta, omega, 4 g, q, ax, az,
    4 F, and the 10 state arrays (Phi, Q, S with their increments); writes the
    10 state arrays; CPML cells also read and write the 8 O arrays."""
    Xn, Zn, upd, interior = _cells(p)
    base = 4.0 * (6 + 3 + 4 + 10 + 10)
    cpml = upd - interior
    return upd * base + cpml * 4.0 * 16


def fused_core_bytes(p, nshots):
    """k_core in forward mode, per launch: every core cell
    [L+N, Xn-L-N) x [L+N, Zn-L-N) of every shot reads the 8 state arrays
    (V, DV, S, DS; x and z) and vp^2, vs^2, and writes the 8 state arrays:
    18 floats = 72 B.  Halo re-reads are cache reuse, not counted."""
    Xn, Zn = p.nx + 2 * p.npml, p.nz + 2 * p.npml
    core = max(0, Xn - 2 * (p.npml + p.half_order)) * max(0, Zn - 2 * (p.npml + p.half_order))
    return nshots * core * 72.0


def dominant_kernel_bytes(p, engine, nshots):
    if engine == abi.ENGINE_FUSED:
        return fused_core_bytes(p, nshots), "k_core"
    return plain_velocity_bytes(p), "k_velocity"


def measured_traffic(kernel_name, nshots=1):
    """HBM bytes per launch of `kernel_name` over `nshots` shots, from the
    committed rocprofv3 PMC summary (profiles/pmc_traffic.md: lines
    'kernel: <bytes> bytes per shot', FETCH_SIZE and WRITE_SIZE already scaled
    by a plain copy's readings), or None."""
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "profiles", "pmc_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            line = line.strip().lstrip("-* ")
            if line.startswith(kernel_name + ":"):
                try:
                    return float(line.split(":", 1)[1].split()[0]) * nshots
                except ValueError:
                    return None
    return None
