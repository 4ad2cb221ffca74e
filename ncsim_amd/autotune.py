"""Autotuner for the batched form.

The batched kernel's speed depends on three planning choices that trade
against each other differently per workload and device:

- tile width (qubits per fused sweep, at most 13 for complex64 and 12 for
  complex128): wider tiles cover more gates per pass over the state, and
  tiles below the largest run in workgroups of 256 threads (four resident
  per CU instead of two);
- fixed low tile qubits (5 / 4 by default): more give longer contiguous
  runs per load, fewer leave the planner more freedom (fewer sweeps);
- tiles per queue item: longer items need fewer device-scope releases,
  shorter ones spread a step over more workgroups.

`tune_batch` times each candidate on a stratified sample of the batch's own
entries (same circuits, same Paulis, seeded draws) and takes the fastest
(near-ties to the larger queue item).  Longer queue items pay off only when
the whole batch keeps every workgroup busy, which a sample does not (the
4096-circuit complex128 mix: 851 ms per run with 32 tiles per item, 797 ms
with 64, 783 ms with 128, while a quarter of it runs fastest with 64 and an
eighth with 32), so when the whole batch runs in about `full_budget_s` or
less, the item sizes are then timed on the batch itself with the chosen tile
shape (the results do not depend on the item size).  The winner is applied
to the batch and recorded in a JSON cache keyed by device, precision and the
batch's shape (qubit-count histogram, draws or none), so later batches of
the same kind skip the search.
"""
import ctypes
import json
import os
import time

from . import _lib as L

DEFAULT_CACHE = os.path.join(os.path.expanduser("~"), ".cache", "ncsim_amd", "batch_tune.json")


def device_arch():
    buf = ctypes.create_string_buffer(256)
    st = L.lib().ncsv_device_arch(buf, 256)
    if st != 0:
        from . import _check
        _check(st)
    return buf.value.decode()


def _signature(batch):
    ns = [c.num_qubits() for c in batch._circuits]
    buckets = {}
    for n in ns:
        b = f"{(n // 4) * 4}-{(n // 4) * 4 + 3}"
        buckets[b] = buckets.get(b, 0) + 1
    total = len(ns)
    hist = ",".join(f"{k}:{round(v / total, 1)}" for k, v in sorted(buckets.items()))
    noisy = "draws" if any(batch._ndraws) else "nodraws"
    return f"{'c64' if batch.precision == 1 else 'c128'}|{hist}|{noisy}"


def candidates(precision):
    top = 13 if precision == 1 else 12
    low = 5 if precision == 1 else 4
    out = []
    for bits in (top, top - 1):
        for lo in (low, 3):
            for items in (32, 128):
                out.append({"tile_bits": bits, "low_bits": lo, "tiles_per_item": items})
    return out


def _apply(batch, cfg):
    batch.set_tile_bits(cfg["tile_bits"])
    batch.set_shape(cfg["low_bits"], cfg["tiles_per_item"])


def choose(timings):
    """The candidate to apply from [(config, kernel_ms), ...]: the fastest,
    with near-ties (within 3 %, about the samples' run-to-run spread) going
    to the larger queue item.  A sample holds a small share of the full
    batch's items, and the device-scope release at each item end weighs
    more at full size (2000 noisy 24-qubit trajectories: 12.4 s with 128
    tiles per item against 13.2 s with 32, within 2 % of each other on a
    48-trajectory sample)."""
    fastest = min(t[1] for t in timings)
    near = [t for t in timings if t[1] <= fastest * 1.03]
    return max(near, key=lambda t: (t[0]["tiles_per_item"], -t[1]))[0]


ITEMS = (32, 64, 128, 256)  # queue item sizes timed on the whole batch


def tune_batch(batch, sample=96, reps=2, cache=None, force=False, log=None, full_budget_s=2.0):
    """Chooses and applies the fastest candidate; returns
    {"config": ..., "timings": [(config, kernel_ms), ...], "cached": bool}.
    The item-size stage runs the batch itself (seeded draws), so results
    read before a later run are those of a tuning run."""
    from . import HipBatch

    cache = cache or os.environ.get("NCSV_TUNE_CACHE", DEFAULT_CACHE)
    key = f"{device_arch()}|{_signature(batch)}"
    db = {}
    if os.path.exists(cache):
        try:
            with open(cache) as f:
                db = json.load(f)
        except (OSError, ValueError):
            db = {}
    if not force and key in db:
        _apply(batch, db[key]["config"])
        return {"config": db[key]["config"], "timings": db[key].get("timings", []), "cached": True}

    n = len(batch)
    order = sorted(range(n), key=lambda i: batch._circuits[i].num_qubits())
    step = max(1, n // sample)
    idx = order[::step][:sample]
    seeds = list(range(len(idx)))
    timings = []
    for cfg in candidates(batch.precision):
        b = HipBatch(batch.precision)
        _apply(b, cfg)
        for i in idx:
            b.add(batch._circuits[i], batch._paulis[i])
        b.run(seeds=seeds)  # plans and uploads
        best = None
        for _ in range(reps):
            b.run(seeds=seeds)
            ms = b.stats()["kernel_ms"]
            best = ms if best is None else min(best, ms)
        timings.append((cfg, best))
        if log:
            log(f"autotune {cfg}: {best:.2f} ms")
        del b
    cfg = choose(timings)
    est_s = min(t[1] for t in timings) * n / len(idx) / 1e3  # the whole batch at the sample's rate
    if n > len(idx) and est_s <= full_budget_s:
        whole = []
        all_seeds = list(range(n))
        for items in ITEMS:
            c2 = dict(cfg, tiles_per_item=items)
            _apply(batch, c2)
            batch.run(seeds=all_seeds)  # plans once per tile shape, captures the launch
            batch.run(seeds=all_seeds)
            ms = batch.stats()["kernel_ms"]
            whole.append((c2, ms))
            if log:
                log(f"autotune {c2} on the whole batch: {ms:.2f} ms")
        cfg = min(whole, key=lambda t: t[1])[0]
        timings += whole
    _apply(batch, cfg)
    db[key] = {"config": cfg, "timings": timings, "sample": len(idx), "when": time.strftime("%Y-%m-%d %H:%M:%S")}
    try:
        os.makedirs(os.path.dirname(cache), exist_ok=True)
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    except OSError:
        pass
    return {"config": cfg, "timings": timings, "cached": False}
