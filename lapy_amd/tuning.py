"""A cache of the batched eigensolver's tuned launch shapes, kept across
processes (no counterpart in the reference, whose ShapeDNA loop solves one
mesh at a time, lapy/shapedna.py:96-166).

``lh_eigs_autotune`` picks the worker count of ``lh_eigs_batched`` and, for
each matrix shape, the ELL SpMM's workgroup size (64, 128 or 256 threads).
This module stores what a context holds in a JSON file, keyed by the device
architecture (``lh_device_arch``) and, per entry, by the matrix shape
``(n, nnz)`` with the ``k`` it was tuned for::

    {"format": 1,
     "archs": {"gfx950:sramecc+:xnack-": {
         "workers": 16,
         "spmm": [{"n": 2562, "nnz": 17922, "k": 50, "ncol": 64, "block": 128}, ...]}}}

Entries of other architectures are kept when a file is updated and ignored
when it is loaded.
"""

from __future__ import annotations

import json
import os
import tempfile

from . import _lib

FORMAT = 1


def _read(path) -> dict:
    if not os.path.exists(path):
        return {"format": FORMAT, "archs": {}}
    with open(path) as f:
        doc = json.load(f)
    if doc.get("format") != FORMAT or not isinstance(doc.get("archs"), dict):
        raise ValueError(f"{path}: not a lapy_amd tuning cache of format {FORMAT}")
    return doc


def merge(doc: dict, arch: str, workers: int, entries) -> dict:
    """doc with arch's workers replaced (when > 0) and its entries added or
    replaced by shape (n, nnz)."""
    a = doc["archs"].setdefault(arch, {"workers": 0, "spmm": []})
    if workers > 0:
        a["workers"] = int(workers)
    by_shape = {(e["n"], e["nnz"]): dict(e) for e in a["spmm"]}
    for e in entries:
        by_shape[(int(e["n"]), int(e["nnz"]))] = {k: int(e[k]) for k in ("n", "nnz", "k", "ncol", "block")}
    a["spmm"] = sorted(by_shape.values(), key=lambda e: (e["n"], e["nnz"]))
    return doc


def entries_for(doc: dict, arch: str):
    """(workers, entries) stored for arch ((0, []) when none)."""
    a = doc["archs"].get(arch)
    return (0, []) if a is None else (int(a.get("workers", 0)), list(a.get("spmm", [])))


def save(path, ctx=None) -> None:
    """Merge the context's tuned shapes into the cache file (atomic write)."""
    ctx = ctx or _lib.default_context()
    workers, entries = ctx.tuning()
    doc = merge(_read(path), ctx.arch(), workers, entries)
    d = os.path.dirname(os.path.abspath(path))
    fd, tmp = tempfile.mkstemp(dir=d, prefix=".tuning-", suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(doc, f, indent=1)
    os.replace(tmp, path)


def load(path, ctx=None) -> int:
    """Install the cache's shapes for this device's architecture in the
    context; returns how many SpMM shapes it held."""
    ctx = ctx or _lib.default_context()
    workers, entries = entries_for(_read(path), ctx.arch())
    ctx.set_tuning(workers, entries)
    return len(entries)


def autotune_cached(As, Bs, ks, path, ctx=None, sigma: float = -0.01, **opts) -> bool:
    """Load the cache; run lh_eigs_autotune only when it lacks the worker
    count or one of these problems' shapes (n, nnz, k), then save.  Returns
    True when the tuner ran."""
    ctx = ctx or _lib.default_context()
    if os.path.exists(path):
        load(path, ctx)
    workers, entries = ctx.tuning()
    have = {(e["n"], e["nnz"], e["k"]) for e in entries}
    need = {(a.n, a.nnz, int(k)) for a, k in zip(As, ks)}
    if workers > 0 and need <= have:
        return False
    _lib.eigs_autotune(ctx, As, Bs, ks, sigma, **opts)
    save(path, ctx)
    return True
