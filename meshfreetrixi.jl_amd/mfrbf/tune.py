"""Launch-shape autotuner for the rhs! kernels.

The stencil kernel and the transposed gather each come in variants that
differ only in how many stencil entries a thread loads before it uses them
(the gathers in flight per thread against the registers they hold).  The
per-node operation order is the same in every variant, so the choice
changes time, never results.

``autotune(ctx, t, precisions)`` times every variant of a configured
context with ``mf_time_rhs`` (graph-replayed chains of the single kernel),
keeps the fastest, sets it with ``mf_set_tuning`` and returns the table.
Tables are cached per (nodes, stencil size, equations, viscosity, device)
in a JSON file (``$MFRBF_TUNE_CACHE``, default
``~/.cache/mfrbf/tune.json``) so later contexts of the same shape skip the
sweep.
"""
import json
import os

from . import _lib as L

FLUX_BATCHES = {15: (3, 5), 20: (5, 10), 30: (3, 5, 6, 10)}
TRANSPOSE_BATCHES = (2, 4, 8)
PREC_NAMES = {L.MF_PREC_F64: "f64", L.MF_PREC_SPLIT32: "split32", L.MF_PREC_F64_REF: "f64_ref"}


def cache_path():
    return os.environ.get("MFRBF_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "mfrbf", "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(table):
    path = cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass  # a read-only home only loses the cache


def shape_key(n, k, eq, visc, prec, device_name=""):
    # node counts are bucketed to powers of two: the best batch depends on
    # occupancy and cache residency, which change with the order of N
    bucket = max(int(n), 1).bit_length()
    return "%s|n2^%d|k%d|eq%d|visc%d|%s" % (device_name, bucket, k, eq, visc, PREC_NAMES[prec])


def sweep(ctx, t, prec, k, calls=50, rounds=3):
    """Times every variant of one precision; returns {"flux": {b: ms},
    "transpose": {b: ms}} (ms per launch of the kernel alone).  Each
    variant is timed in `rounds` rounds, the pairing and order rotated
    between rounds, and keeps its minimum: a slow outlier in one round
    does not decide the pick."""
    res = {"flux": {}, "transpose": {}}
    fb = FLUX_BATCHES.get(k, (0,))
    tb = TRANSPOSE_BATCHES
    n = max(len(fb), len(tb))
    for r in range(rounds):
        for i in range(n):
            f = fb[(i + r) % len(fb)]
            b = tb[(i + 2 * r) % len(tb)]
            ctx.set_tuning(L.MF_TUNE_FLUX_BATCH, f, prec)
            ctx.set_tuning(L.MF_TUNE_TRANSPOSE_BATCH, b, prec)
            ctx.time_rhs(t, prec, 3)  # warm the new variant
            _, parts = ctx.time_rhs(t, prec, calls)
            if f not in res["flux"] or parts[0] < res["flux"][f]:
                res["flux"][f] = float(parts[0])
            if parts[1] > 0 and (b not in res["transpose"] or parts[1] < res["transpose"][b]):
                res["transpose"][b] = float(parts[1])
    return res


def autotune(ctx, t, n, k, eq, visc, precisions=(L.MF_PREC_F64,), calls=50, use_cache=True,
             device_name=""):
    """Picks and sets the fastest batch sizes for each precision.  Returns
    {precision name: {"flux": b, "transpose": b, "sweep": {...} or None}}."""
    table = _load() if use_cache else {}
    out = {}
    dirty = False
    for prec in precisions:
        key = shape_key(n, k, eq, visc, prec, device_name)
        hit = table.get(key)
        if hit is None:
            sw = sweep(ctx, t, prec, k, calls)
            best_f = min(sw["flux"], key=sw["flux"].get)
            best_t = min(sw["transpose"], key=sw["transpose"].get) if sw["transpose"] else 0
            hit = {"flux": int(best_f), "transpose": int(best_t),
                   "sweep": {"flux": {str(a): b for a, b in sw["flux"].items()},
                             "transpose": {str(a): b for a, b in sw["transpose"].items()}}}
            table[key] = hit
            dirty = True
        ctx.set_tuning(L.MF_TUNE_FLUX_BATCH, hit["flux"], prec)
        ctx.set_tuning(L.MF_TUNE_TRANSPOSE_BATCH, hit["transpose"], prec)
        out[PREC_NAMES[prec]] = hit
    if dirty and use_cache:
        _save(table)
    return out
