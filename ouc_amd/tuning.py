"""Tile autotuning with a results cache.

The fused engine's interior kernel has two launch knobs: the x columns each
256-lane block streams and the columns of loads kept in flight.  The best
choice depends on the grid and the batch size (how many blocks fill the 256
CUs).  `tuned(engine, nshots)` returns the cached choice for this device and
shape, or runs the engine's own autotuner (ouc_autotune) and records it.

Cache: JSON at $OUC_TILE_CACHE (default ~/.cache/ouc_amd/tiles.json), keyed by
"<gfx arch>|<Xn>x<Zn>|N<half order>|B<shots>|<boundary>".
"""
import json
import os

from .engine import device_arch

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "ouc_amd", "tiles.json")


def cache_path():
    return os.environ.get("OUC_TILE_CACHE", DEFAULT_CACHE)


def shape_key(arch, params, nshots):
    if params.boundary:
        pad = params.band if params.band > 0 else 8
    else:
        pad = params.npml
    Xn, Zn = params.nx + 2 * pad, params.nz + 2 * pad
    return f"{arch}|{Xn}x{Zn}|N{params.half_order}|B{int(nshots)}|{'learned' if params.boundary else 'cpml'}"


def load_cache(path=None):
    path = path or cache_path()
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def store(key, entry, path=None):
    path = path or cache_path()
    data = load_cache(path)
    data[key] = entry
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tuned(engine, nshots, reps=5, path=None, device=0):
    """Set the engine's tile from the cache, tuning (and caching) on a miss.
    Returns (core_cols, prefetch, from_cache)."""
    key = shape_key(device_arch(device), engine.p, nshots)
    hit = load_cache(path).get(key)
    if hit:
        engine.set_tile(hit["core_cols"], hit["prefetch"])
        return hit["core_cols"], hit["prefetch"], True
    (cols, pf), table = engine.autotune(nshots, reps)
    store(key, {"core_cols": cols, "prefetch": pf,
                "ms_per_launch": {f"{c}x{p}": v for (c, p), v in table.items()}}, path)
    return cols, pf, False
