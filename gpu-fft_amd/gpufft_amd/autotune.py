"""Autotuner for the product kernels.

Times candidate kernel configurations on the current device and keeps the
fastest per (device, workload signature) in a JSON tuning database, so later
plans for the same workload reuse the choice without timing again.

Tuned knobs (each changes speed only; results are the same up to rounding,
and rounded products are identical):
  polymul plans  radix bits of the LDS steps (gpufft_polymul_plan_set_radix_bits)
  batched plans  radix bits and persistent grid size (gpufft_batched_plan_set_tuning)

The database is $GPUFFT_AMD_TUNING_DB, else ~/.cache/gpufft_amd/tuning.json.
"""
import hashlib
import json
import os
import tempfile

DB_ENV = "GPUFFT_AMD_TUNING_DB"


def db_path():
    return os.environ.get(DB_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "gpufft_amd",
                                                  "tuning.json")


def load_db(path=None):
    path = path or db_path()
    try:
        with open(path) as f:
            db = json.load(f)
        return db if isinstance(db, dict) else {}
    except (OSError, ValueError):
        return {}


def save_db(db, path=None):
    """Atomic write (temp file + rename)."""
    path = path or db_path()
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=d or ".", prefix=".tuning.", suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(db, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def signature(*parts):
    """Stable key of a workload description (any JSON-serialisable parts)."""
    blob = json.dumps(parts, sort_keys=True, separators=(",", ":"))
    return hashlib.sha1(blob.encode()).hexdigest()[:16]


def device_key():
    import torch
    p = torch.cuda.get_device_properties(torch.cuda.current_device())
    return "%s|%s|%d" % (p.name, getattr(p, "gcnArchName", ""), p.multi_processor_count)


def best_of(timings):
    """Candidate with the smallest time; ties keep the first listed."""
    best = None
    for cand, t in timings:
        if best is None or t < best[1]:
            best = (cand, t)
    return best


def time_ms(fn, stream, reps=5, warmup=1):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def tune(key, candidates, apply, run, stream, reps=5, use_db=True, db_file=None):
    """Pick the fastest candidate for `key`.

    apply(cand) configures the plan, run() executes one timed unit on
    `stream`.  With use_db, a stored choice for (device, key) is applied
    without timing; otherwise every candidate is timed and the winner stored.
    Returns (candidate, {candidate: ms})."""
    full = device_key() + "|" + key
    db = load_db(db_file) if use_db else {}
    if use_db and full in db:
        cand = tuple(db[full]["choice"])
        apply(cand)
        return cand, {}
    timings = []
    for cand in candidates:
        apply(cand)
        timings.append((tuple(cand), time_ms(run, stream, reps)))
    cand, t = best_of(timings)
    apply(cand)
    if use_db:
        db = load_db(db_file)
        db[full] = {"choice": list(cand), "ms": t,
                    "timings": {",".join(map(str, c)): ms for c, ms in timings}}
        save_db(db, db_file)
    return cand, {c: ms for c, ms in timings}


def tune_polymul(plan, run, stream, key=None, reps=5, use_db=True, db_file=None):
    rbs = (3, 4) if plan.precision == 0 else (3, 4, 5)
    key = key or signature("polymul", plan.ring_power, plan.precision)
    return tune(key, [(r,) for r in rbs], lambda c: setattr(plan, "radix_bits", c[0]), run, stream,
                reps, use_db, db_file)


def tune_batched(plan, run, stream, key, grids=None, reps=3, use_db=True, db_file=None,
                 tiles=(14, 13)):
    """Radix of the LDS steps, persistent grid size, where split products are
    combined (0: separate kernel, 1: inside the persistent kernel) and the
    LDS tile (log2 points per workgroup: 14, one workgroup of 8 waves per
    CU, or 13, two per CU).  Tiles the plan refuses (a product would need
    more than 2^4 sub-problems) are skipped.  FP64 plans have the one tile
    13 (one workgroup per CU), radix 2^3 or 2^4, and grids of one or two
    workgroups per CU."""
    import torch

    from . import PREC_F64
    cus = torch.cuda.get_device_properties(torch.cuda.current_device()).multi_processor_count
    f64 = getattr(plan, "precision", None) == PREC_F64
    ok_tiles = []
    for t in tiles if not f64 else (13,):
        try:
            plan.tile = t
            ok_tiles.append(t)
        except ValueError:
            pass
    cands = []
    for t in ok_tiles:
        if f64:
            gs = grids or (cus, 2 * cus)
            radices = (3, 4)
        else:
            per_cu = 1 if t == 14 else 2
            gs = grids or (cus * per_cu // 2, cus * per_cu)
            radices = (3, 4, 5)
        cands += [(rb, g, cmb, t) for cmb in (0, 1) for rb in radices for g in gs]

    def apply(c):
        if len(c) > 3 and plan.tile != c[3]:
            plan.tile = c[3]
        plan.set_tuning(c[0], c[1])
        plan.combine_in_kernel = bool(c[2]) if len(c) > 2 else False
    return tune(key, cands, apply, run, stream, reps, use_db, db_file)
