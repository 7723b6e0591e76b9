"""Autotuner for the term-chunk sizes of the sieve kernels.

A request's 2^(h-1) (or prod(e_i + 1)) terms are cut into chunks, one workgroup per chunk.  The
chunk size trades tail balance and latency-hiding across workgroups (many small chunks) against
the per-workgroup cost of staging the matrix and folding partial sums (few large chunks).  The
library aims at a number of workgroups per request per kernel class (`wh_tune_set`); this
module measures candidate values on the device and keeps, per class, the fastest when it beats
the built-in default by more than 1% (`choose`):

* kind 0: the all-ones kernels (every multiplicity 1, even N), by tile side class T = 1..6
  (n_pad = 8T), plain and loop;
* kind 1: the general kernel (odd N, and repeated rows beyond the repeated-row kernel's
  limits), plain and loop;
* kind 2: the repeated-row kernel (even N, 2E <= 16, h <= 16), by tile side class T = 1
  (2E <= 8), 3 (2E <= 12, four terms per wave) and 2 (2E <= 16), plain and loop.

For the all-ones and repeated-row classes the tuner then times the waves per workgroup, 1 to 4
against the built-in choice (the count that fits the most waves per CU in the LDS), and keeps a
count only when it beats that choice by more than 1% (`wh_tune_set_waves`; a count whose
workgroup does not fit the LDS is ignored by the library).  Last it times the next larger tile
side for those classes (all-ones T = 1..5, repeated-row sides 1 and 3: the same requests with
more zero padding) and keeps it only when it wins by more than 1% (`wh_tune_set_tile`).

The table lives in a JSON file (`tuning.json` next to this module by default, or the path in
the environment variable ``WH_TUNING``; ``WH_TUNING=none`` keeps the built-in defaults) that the
library applies when it is loaded.  Tuning changes only the chunking: each result changes in
its last bits at most (summation order), and stays deterministic for a given table.
"""
import json
import os
import time

import numpy as np

from . import _native

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")
CANDIDATES = (256, 512, 1024, 2048, 4096)
# per kind, when the caller names no candidates: the repeated-row kernel's requests have hundreds
# of terms, the general kernel's chunks at most 256 terms
KIND_CANDIDATES = {0: CANDIDATES, 1: (64, 128, 256, 512, 1024), 2: (4, 8, 16, 32, 64)}


def table_path():
    """The tuning file the library applies, or None when tuning is switched off."""
    p = os.environ.get("WH_TUNING")
    if p is None:
        return DEFAULT_PATH
    if p.strip().lower() in ("", "none", "off", "0"):
        return None
    return p


def apply(entries, lib=None):
    """Sets the library's table from a list of {"kind", "T", "loop", "blocks_per_request"} (and
    "waves" for the all-ones and repeated-row classes; absent = 0, the built-in choice)."""
    L = lib or _native.load()
    for e in entries:
        rc = L.wh_tune_set(int(e["kind"]), int(e["T"]), int(e["loop"]), int(e["blocks_per_request"]))
        if rc == 0 and int(e["kind"]) != 1:
            rc = L.wh_tune_set_waves(int(e["kind"]), int(e["T"]), int(e["loop"]), int(e.get("waves", 0)))
        if rc == 0 and int(e["kind"]) != 1:
            rc = L.wh_tune_set_tile(int(e["kind"]), int(e["T"]), int(e["loop"]), int(e.get("tile", 0)))
        if rc != 0:
            raise ValueError(L.wh_last_error().decode())


def current(lib=None):
    """The library's table: {(kind, T, loop): blocks_per_request}, 0 = built-in default."""
    L = lib or _native.load()
    out = {}
    for kind, Ts in ((0, range(1, 7)), (1, (0,)), (2, (1, 2, 3))):
        for T in Ts:
            for loop in (0, 1):
                out[(kind, T, loop)] = int(L.wh_tune_get(kind, T, loop))
    return out


def current_waves(lib=None):
    """The library's waves table: {(kind, T, loop): waves}, 0 = built-in choice."""
    L = lib or _native.load()
    return {k: int(L.wh_tune_get_waves(*k)) for k in current(L) if k[0] != 1}


def current_tiles(lib=None):
    """The library's tile table: {(kind, T, loop): tile side}, 0 = the class's own side."""
    L = lib or _native.load()
    return {k: int(L.wh_tune_get_tile(*k)) for k in current(L) if k[0] != 1}


def load(path=None, lib=None):
    """Applies a tuning file (default: ``table_path()``); returns its entries, or [] if none."""
    p = path if path is not None else table_path()
    if p is None or not os.path.exists(p):
        return []
    with open(p, encoding="utf-8") as f:
        doc = json.load(f)
    entries = doc.get("entries", [])
    apply(entries, lib)
    return entries


def rep_side(two_e):
    """The repeated-row kernel's tile side for a term matrix of side 2E: 1 (2E <= 8), 3 (<= 12,
    four terms per wave) or 2 (<= 16)."""
    return 1 if two_e <= 8 else 3 if two_e <= 12 else 2


# the next larger tile side the tuner tries per class (repeated-row sides by the 2E they hold)
NEXT_TILE = dict([((0, T), T + 1) for T in range(1, 6)] + [((2, 1), 3), ((2, 3), 2)])


def _class_workload(kind, T, loop, rng):
    """Representative requests of one kernel class: about 2^18 terms in total (at most 256
    requests), complex CN(0, 1/n) symmetric matrices."""
    from . import batch as B

    def sym(n):
        X = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) / np.sqrt(2 * n)
        return (X + X.T) / 2

    if kind == 0:
        n = 8 * T if T < 6 else 42  # n = 48 runs seconds per call; 42 is in the same class
        h = n // 2
        count = int(min(256, max(1, 2 ** 18 // 2 ** (h - 1))))
        return [B.prepare(sym(n), loop=bool(loop)) for _ in range(count)]
    if kind == 2:
        # repeated-row kernel, 32 photons: 4 modes (2E <= 8, side 1), 8 modes (2E in 10..12,
        # side 3: four terms per wave; 1024 requests, the repeated-row benchmark's batch, whose
        # hundreds of workgroups per CU change the best waves count) or 16 modes (2E in 14..16,
        # side 2; fewer requests, as each has tens of thousands of terms)
        modes, count = {1: (4, 256), 3: (8, 1024), 2: (16, 32)}[T]
        preps = []
        while len(preps) < count:
            rpt = rng.multinomial(32, [1 / modes] * modes)
            mu = (rng.normal(size=modes) + 1j * rng.normal(size=modes)) if loop else None
            p = B.prepare_repeated(sym(modes), rpt, mu=mu, loop=bool(loop))
            if p.on_device and rep_side(2 * len(p.job.edge_reps)) == T:
                preps.append(p)
        return preps
    # general kernel: repeated rows past the repeated-row kernel's h <= 16, 8 modes, 40 photons
    preps = []
    for _ in range(256):
        rpt = rng.multinomial(40, [1 / 8] * 8)
        mu = (rng.normal(size=8) + 1j * rng.normal(size=8)) if loop else None
        preps.append(B.prepare_repeated(sym(8), rpt, mu=mu, loop=bool(loop)))
    return preps


def _time(preps, reps, torch, stream):
    from . import batch as B

    db = B.DeviceBatch(preps)
    try:
        db.run(stream)  # warm-up (kernel attributes, caches)
        torch.cuda.synchronize()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            db.run(stream)
            stream.synchronize()
            best = min(best, time.perf_counter() - t0)
        return best
    finally:
        db.close()


DEFAULT_BLOCKS = {0: 1024, 1: 256, 2: 64}  # the library's built-in targets (all-ones, general, repeated-row)


def choose(kind, times, margin=0.01):
    """The candidate to keep: the fastest, if it beats the built-in default by more than
    ``margin`` (relative), else the default (timing noise does not move the table)."""
    best = min(times, key=times.get)
    d = DEFAULT_BLOCKS[kind]
    if d in times and times[best] >= (1.0 - margin) * times[d]:
        return d
    return best


def choose_waves(times, margin=0.01):
    """The waves count (or tile side) to keep: the fastest, if it beats the built-in choice (0)
    by more than ``margin``, else 0."""
    best = min(times, key=times.get)
    if times[best] >= (1.0 - margin) * times[0]:
        return 0
    return best


def autotune(path=DEFAULT_PATH, candidates=None, reps=3, classes=None, seed=1234, log=print):
    """Times every candidate per kernel class on the current device, applies the choice
    (``choose``) and writes the table to ``path`` (None: do not write).  Returns the entries."""
    import torch

    L = _native.load()
    rng = np.random.default_rng(seed)
    if classes is None:
        classes = ([(0, T, loop) for T in range(1, 7) for loop in (0, 1)] + [(1, 0, 0), (1, 0, 1)]
                   + [(2, T, loop) for T in (1, 3, 2) for loop in (0, 1)])
    stream = torch.cuda.Stream()
    entries = []
    for kind, T, loop in classes:
        preps = _class_workload(kind, T, loop, rng)
        times = {}
        for c in (candidates or KIND_CANDIDATES[kind]):
            _native.check(L.wh_tune_set(kind, T, loop, int(c)))
            times[c] = _time(preps, reps, torch, stream)
        best = choose(kind, times)
        _native.check(L.wh_tune_set(kind, T, loop, int(best)))
        e = {"kind": kind, "T": T, "loop": loop, "blocks_per_request": int(best),
             "ms": {str(c): round(1e3 * v, 4) for c, v in times.items()}}
        if kind != 1:
            wtimes = {}
            for w in (0, 1, 2, 3, 4):
                _native.check(L.wh_tune_set_waves(kind, T, loop, w))
                wtimes[w] = _time(preps, reps, torch, stream)
            wbest = choose_waves(wtimes)
            _native.check(L.wh_tune_set_waves(kind, T, loop, int(wbest)))
            e["waves"] = int(wbest)
            e["waves_ms"] = {str(w): round(1e3 * v, 4) for w, v in wtimes.items()}
            if (kind, T) in NEXT_TILE:
                ttimes = {}
                for tl in (0, NEXT_TILE[(kind, T)]):
                    _native.check(L.wh_tune_set_tile(kind, T, loop, tl))
                    ttimes[tl] = _time(preps, reps, torch, stream)
                tbest = choose_waves(ttimes)
                _native.check(L.wh_tune_set_tile(kind, T, loop, int(tbest)))
                e["tile"] = int(tbest)
                e["tile_ms"] = {str(tl): round(1e3 * v, 4) for tl, v in ttimes.items()}
        entries.append(e)
        if log:
            log("class kind=%d T=%d loop=%d: keep %d workgroups/request %s, waves %s %s, tile %s %s"
                % (kind, T, loop, best, e["ms"], e.get("waves", "-"), e.get("waves_ms", ""), e.get("tile", "-"),
                   e.get("tile_ms", "")))
    if path is not None:
        doc = {"device": torch.cuda.get_device_name(),
               "candidates": {str(k): list(candidates or v) for k, v in KIND_CANDIDATES.items()}, "entries": entries}
        with open(path, "w", encoding="utf-8") as f:
            json.dump(doc, f, indent=1)
    return entries


if __name__ == "__main__":
    import sys

    autotune(path=sys.argv[1] if len(sys.argv) > 1 else DEFAULT_PATH)
