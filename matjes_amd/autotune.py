"""Tile-shape autotuner for batched runs.

Every workgroup of a batch holds 256 sites; what varies per system is the
tile's shape (bx x by), which sets the halo re-read ratio (and, for the fused
step, the ring of stage-1 sites recomputed around the tile) and how rows map
to cache lines.  Systems are grouped by size class (side length); for each
launch form (the fused step, one launch per stage) coordinate descent over
the classes (largest share of sites first) times each candidate shape for
that class with the others fixed, by HIP events around one captured segment,
and the fastest form with its shapes is kept.  Results are cached in a JSON
file keyed by the batch's composition, so a workload is tuned once.

The timed steps advance the batch's states: tune a batch that is thrown away
afterwards (mj_batch_add copies the contexts' states, so a fresh batch built
from the same contexts starts from the original states).
"""
import hashlib
import json
import os

import numpy as np

from . import batch as _batch

SHAPES = [(32, 8), (64, 4), (16, 16), (128, 2)]
CLASSES = [48, 96, 192, 384]   # upper bounds of the side-length classes; beyond: last class


def size_class(nx):
    for k, lim in enumerate(CLASSES):
        if nx <= lim:
            return k
    return len(CLASSES)


def cache_path():
    return os.environ.get("MJ_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "matjes_amd", "tiles.json"))


def _key(dims, reach):
    h = hashlib.sha1(json.dumps([sorted(map(list, dims)), reach, "forms"]).encode()).hexdigest()[:16]
    return "batch-" + h


def _load(path):
    try:
        with open(path) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(path, data):
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass


def tiles_for(dims, choice):
    """Per-system (bx, by) from a {class: shape} choice."""
    return [tuple(choice.get(str(size_class(d[0])), SHAPES[0])) for d in dims]


FORMS = {"fused": _batch.FUSED, "stages": _batch.STAGES}


def tune(batch, dims, timestep, steps=None, reach=1, use_cache=True, log=None, seg=50, forms=("fused", "stages")):
    """Pick a launch form and a tile shape per size class for `batch` (a
    batch.Batch whose systems have lattice dims `dims`, integration 2,
    captured segments of `seg` steps: the exchange interval, or 50); leaves
    the batch set to them and returns ({class: shape}, report) with the
    form in report["launch"]."""
    path = cache_path()
    key = _key(dims, reach)
    cache = _load(path) if use_cache else {}
    if key in cache:
        choice = cache[key]["choice"]
        batch.set_launch(FORMS[cache[key].get("launch", "fused")])
        batch.set_tiles(tiles_for(dims, choice))
        return choice, {"cached": True, "ms_per_step": cache[key]["ms_per_step"],
                        "launch": cache[key].get("launch", "fused")}
    steps = steps or seg
    sites = {}
    for d in dims:
        c = str(size_class(d[0]))
        sites[c] = sites.get(c, 0) + int(np.prod(d))
    order = sorted(sites, key=lambda c: -sites[c])

    def measure(choice):
        batch.set_tiles(tiles_for(dims, choice))
        # align to a segment, capture and replay one, then time whole segments
        batch.steps((-batch.nstep) % seg + seg, timestep)
        return batch.profile(steps, timestep) / steps

    results = {}
    ntrials = 0
    for form in forms:
        batch.set_launch(FORMS[form])
        choice = {c: SHAPES[0] for c in order}
        best = measure(choice)
        ntrials += 1
        for c in order:
            for shp in SHAPES:
                if tuple(choice[c]) == shp:
                    continue
                old = choice[c]
                choice[c] = shp
                t = measure(choice)
                ntrials += 1
                if log:
                    log("tune %s class %s shape %s: %.3f ms/step" % (form, c, shp, t))
                if t < best:
                    best = t
                else:
                    choice[c] = old
        results[form] = (best, dict(choice))
    form = min(results, key=lambda f: results[f][0])
    best, choice = results[form]
    batch.set_launch(FORMS[form])
    batch.set_tiles(tiles_for(dims, choice))
    if use_cache:
        cache[key] = {"choice": {c: list(s) for c, s in choice.items()}, "ms_per_step": best, "launch": form}
        _save(path, cache)
    return choice, {"cached": False, "ms_per_step": best, "trials": ntrials, "launch": form,
                    "ms_per_step_by_form": {f: r[0] for f, r in results.items()}}
