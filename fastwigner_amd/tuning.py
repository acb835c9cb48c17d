"""FFT tile schedules picked by the autotuner (``fastwigner_amd.autotune``).

The tuned entries live in ``tuning.json`` next to this file, keyed by
propagator kind, grid and precision, the way FFTW wisdom persists plans across
runs (strong_field_physics_fftw_1d.py:21-28,249-255). A missing entry falls back
to the library's built-in defaults.
"""
import json
import os

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
_cache = None


def _load():
    global _cache
    if _cache is None:
        try:
            with open(PATH) as f:
                _cache = json.load(f)
        except (OSError, ValueError):
            _cache = {}
    return _cache


def key(kind, nx, npp, precision):
    return "%s:%dx%d:%s" % (kind, nx, npp, "fp32" if precision else "fp64")


def entry(kind, nx, npp, precision):
    """The tuned entry (a dict with "tiles" and, for Wigner grids, "v_table") or None."""
    return _load().get(key(kind, nx, npp, precision))


def lookup(kind, nx, npp, precision):
    """(col_e, col_p, row_e, row_p) or None."""
    e = entry(kind, nx, npp, precision)
    return tuple(e["tiles"]) if e else None


def batch_entry(precision):
    """The tuned batch schedule ({"e_for_n": [E for N = 64 .. 1024], ...}) or None."""
    return _load().get("batch:%s" % ("fp32" if precision else "fp64"))


def store_entry(name, value, path=PATH):
    """Add or replace the entry `name` of the tuning file at ``path`` (merging with what it holds)."""
    global _cache
    try:
        with open(path) as f:
            data = json.load(f)
    except (OSError, ValueError):
        data = {}
    data[name] = value
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    if path == PATH:
        _cache = None


def store(kind, nx, npp, precision, tiles, info=None, path=PATH):
    """Set the tiles (and timings) of one entry of the tuning file at ``path``; the
    entry's other keys (a tuned V-factor policy, say) are kept."""
    name = key(kind, nx, npp, precision)
    try:
        with open(path) as f:
            old = json.load(f).get(name, {})
    except (OSError, ValueError):
        old = {}
    store_entry(name, {**old, "tiles": list(tiles), **(info or {})}, path=path)
