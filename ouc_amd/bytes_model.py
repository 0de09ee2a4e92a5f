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


def fused_regions(p):
    """Cell counts per shot of the fused engine's kernel regions (CPML
    contexts): the update region, its core, the CPML band around the core, the
    interior, the reconstruction core [L+2N, Xn-L-2N)^2 and band, and the
    strip ring (four strips 2N cells wide straddling the interior's edge)."""
    N = p.half_order
    Xn, Zn, upd, interior = _cells(p)
    core = max(0, p.nx - 2 * N) * max(0, p.nz - 2 * N)
    rcore = max(0, p.nx - 4 * N) * max(0, p.nz - 4 * N)
    ring = 2 * (2 * N * (p.nx + 2 * N)) + 2 * (2 * N * max(0, p.nz - 2 * N))
    return {"update": upd, "core": core, "band": upd - core, "interior": interior, "recon_core": rcore,
            "recon_band": interior - rcore, "ring": ring}


# Algorithmic bytes per cell of each fused kernel (FP32 operands once):
#   k_core step: V, DV, S, DS (x, z) read and written, vp^2, vs^2 read  = 72
#   k_core imaging: + source P level (2) read, 3 accumulators RMW     = 104
#   k_edge (CPML band): the core's 72 + ax, az + 4 F RMW + 8 O RMW    = 176
#   reconstruction (core mode and band kernels): V, DV, S, DS RMW,
#     vp^2, vs^2, + source P level (2) written                        = 80
#   strip save, per ring cell and field: the value read, FP16 residual
#     written, encoder state (prediction, fed-back error) RMW        = 22 (FP32 strips: 8)
#   strip decode, per ring cell and field: residual read, previous
#     decoded value read, decoded value written                       = 10 (FP32 strips: 0)
CORE_STEP_B, CORE_IMAGE_B, EDGE_B, RECON_B = 72.0, 104.0, 176.0, 80.0


def fused_loop_bytes(p, nshots, fp16):
    """Algorithmic bytes of one batch's forward loop and backward loop (all
    steps, all kernels), for the loop-level rates next to the roofline."""
    r = fused_regions(p)
    nf = 8
    save = r["ring"] * nf * (22.0 if fp16 else 8.0)
    dec = r["ring"] * nf * (10.0 if fp16 else 0.0)
    fwd = r["core"] * CORE_STEP_B + r["band"] * EDGE_B + save
    bwd = (r["core"] * CORE_IMAGE_B + r["band"] * EDGE_B + r["interior"] * RECON_B + dec)
    return nshots * p.nt * fwd, nshots * (p.nt - 1) * bwd
