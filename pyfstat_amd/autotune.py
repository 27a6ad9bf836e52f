"""Autotuner for the batched map kernels: rows (t0) per work item per kernel
mode, chosen by timing the step's kernels on the caller's own batch.

A work item is up to ``rows`` t0 rows x 256 tau columns of one template.  More
rows amortise the item's set-up (diagonal staging for the rect kernel, the
chunk walk for the exp recurrence) and the per-item reductions; fewer rows give
more items to spread over the 256 CUs and a smaller LDS footprint.  The best
value depends on the batch's shapes, so it is measured: each candidate value
is planned (``BatchedTransientMaps.set_rows`` + ``replan``, which pass it to
``tcw_batch_plan_rows``) and the step's kernels timed with HIP events
(``run_timed``); the fastest is kept and cached per (device, mode, shape
class).  The setting belongs to the tuned batch's plan: other batches and
single-map calls keep the library's own choice.

The reference has no counterpart (its kernels use fixed 1-D launches,
pyfstat/tcw_fstat_map_funcs.py:862-870, 944-952).
"""

import json
import math
import os

import numpy as np

from . import _hiplib

CANDIDATES = {
    _hiplib.MODE_RECT_FAST: (16, 32, 64, 128, 192, 256, 384),
    _hiplib.MODE_RECT_GENERIC: (16, 32, 64),
    _hiplib.MODE_EXP_DIRECT: (16, 32, 64),
    _hiplib.MODE_EXP_FAST: (64, 128, 256, 512, 1024),
}

def cache_path():
    env = os.environ.get("PYFSTAT_AMD_TUNE_CACHE")
    if env:
        return env
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "pyfstat_amd", "autotune.json")


def _load_cache(path):
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _save_cache(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only costs re-tuning


def _bucket(x):
    return int(round(math.log2(max(float(x), 1.0))))


def shape_class(b, mode, device_name=""):
    """Cache key: device, mode, and log2 buckets of the template count and of
    the templates' median t0 rows and tau columns (none of which depends on
    the rows-per-item setting being tuned)."""
    shapes = np.array([b._shape(t) for t in range(b.T)], dtype=np.int64)
    version = _hiplib.load().tcw_version().decode(errors="replace")
    return (f"{version}|{device_name}|mode{mode}|T{_bucket(b.T)}"
            f"|rows{_bucket(np.median(shapes[:, 0]))}|cols{_bucket(np.median(shapes[:, 1]))}")


def set_rows(b, mode, rows):
    """Plan batch ``b`` with ``rows`` per work item for ``mode`` (0: the
    library's choice); raises ValueError for a value the library rejects,
    and then restores the previous value and plan."""
    prev = int(b.rows_by_mode[int(mode)])
    b.set_rows(mode, rows)
    try:
        b.replan()
    except ValueError:
        b.set_rows(mode, prev)
        b.replan()
        raise


def reset(b):
    """Back to the library's choice for every mode of batch ``b``."""
    for mode in range(_hiplib.MODE_COUNT):
        b.set_rows(mode, 0)
    b.replan()


def _time_step(b, reps):
    """Best of reps: stage (merge, prefix or exp chunk scan) + maps +
    finalize, each timed with HIP events; rows per item moves work between
    the exp chunk scan and the map kernel, so the whole step is what counts."""
    b.run_timed()  # warm: first launch after a replan
    return min(sum(b.run_timed()) for _ in range(reps))


def tune(b, reps=3, candidates=None, use_cache=True, device_name=None, log=None):
    """Pick rows per item for every kernel mode the batch uses; leaves the
    winners set on ``b`` (re-planned with them).  Returns
    {mode: {"rows": r, "step_ms": t, "tried": {r: t}, "cached": bool}}."""
    from .tcw_fstat_map_funcs import _hip_device_names, _HipState

    cands = dict(CANDIDATES if candidates is None else candidates)
    if device_name is None:
        names = _hip_device_names()
        dev = _HipState.device or 0
        device_name = names[dev] if dev < len(names) else ""
    path = cache_path()
    cache = _load_cache(path) if use_cache else {}
    items = b.items_per_mode()
    result = {}
    changed = False
    for mode in range(_hiplib.MODE_COUNT):
        if items[mode] == 0 or mode not in cands:
            continue
        key = shape_class(b, mode, device_name)
        if use_cache and key in cache:
            try:
                set_rows(b, mode, cache[key]["rows"])
            except (ValueError, KeyError, TypeError):
                # a value this library no longer accepts (its limits changed
                # under the same version string): tune again
                del cache[key]
            else:
                result[mode] = dict(cache[key], cached=True)
                changed = True
                continue
        tried = {}
        for rows in cands[mode]:
            set_rows(b, mode, rows)
            tried[rows] = _time_step(b, reps)
            if log:
                log(f"autotune mode {mode}: rows {rows} -> step {tried[rows]:.4f} ms")
        best = min(tried, key=tried.get)
        b.set_rows(mode, best)
        changed = True
        result[mode] = {"rows": int(best), "step_ms": float(tried[best]),
                        "tried": {str(k): float(v) for k, v in tried.items()}, "cached": False}
        cache[key] = {k: v for k, v in result[mode].items() if k != "cached"}
    if changed:
        b.replan()
    if use_cache:
        _save_cache(path, cache)
    return result
