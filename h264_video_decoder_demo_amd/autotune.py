"""Band-height autotuner for the two sequential stages.

Intra prediction and deblocking run as bands of MB rows, one workgroup per band (one
row per wave for intra, two for deblocking).  The band height sets the waves per
workgroup and so the waves per CU, and trades dependency-chain steps inside a band
against band crossings through device memory; the best heights depend on the device
and on the mix of picture sizes that share a launch.  `autotune` measures candidate
heights on the caller's own streams -- their pictures (all, or the first n_frames),
staged once and batched by runs of independent frames (engine.independent_runs) -- and caches the
winner per device (architecture, name, CU count, clock), library build, engine
settings and stream mix (a JSON file under the user's cache folder).

Every height gives the same bytes; only the time differs.
"""
from __future__ import annotations

import collections
import hashlib
import json
import os
import time

DEBLOCK_ROWS = (4, 6, 8, 10, 12, 16)
_clock = time.perf_counter
INTRA_ROWS = (4, 6, 8)


def cache_path() -> str:
    root = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(root, "h264g", "autotune.json")


def _library_id() -> str:
    """Digest of the built library: a rebuilt kernel invalidates earlier measurements."""
    from ._lib import LIB_PATH
    h = hashlib.sha1()
    with open(LIB_PATH, "rb") as f:
        for chunk in iter(lambda: f.read(1 << 20), b""):
            h.update(chunk)
    return h.hexdigest()[:16]


def _device_id(device: int) -> str:
    """Architecture, name, CU count and clock of the HIP device."""
    import ctypes
    from ._lib import check, lib
    buf = ctypes.create_string_buffer(256)
    check(lib().h264g_device_describe(device, buf, len(buf)))
    return buf.value.decode(errors="replace")


def _key(gens, n_frames: int, deblock_rows, intra_rows, device: int = 0, engine_kwargs: dict | None = None) -> str:
    """Cache key: the device, the library build, the engine settings that change the
    schedules, the stream mix and the candidates."""
    mix = sorted(collections.Counter((g.width_mbs, g.height_mbs) for g in gens).items())
    kw = sorted((k, v) for k, v in (engine_kwargs or {}).items() if k not in ("deblock_waves", "intra_waves"))
    blob = json.dumps([_device_id(device), _library_id(), kw, mix, n_frames, list(deblock_rows), list(intra_rows)])
    return hashlib.sha1(blob.encode()).hexdigest()[:16]


def autotune(gens, n_frames: int = 0, reps: int = 2, deblock_rows=DEBLOCK_ROWS, intra_rows=INTRA_ROWS,
             device: int = 0, use_cache: bool = True, measure: bool = True,
             engine_kwargs: dict | None = None) -> dict:
    """Best (deblock_rows, intra_rows) for streams `gens` (Generator-like: width_mbs,
    height_mbs, dpb_slots, num_pictures, picture_raw(i)); n_frames 0 = every picture.
    Coordinate search: the deblocking height with the intra height at 8, then the
    intra height.  measure=False only looks the choice up (defaults 8 / 8 when it is
    not cached), so that nothing is launched.  Returns {"deblock_waves",
    "intra_waves", "ms", "cached", "seconds"}."""
    from .engine import Engine, independent_runs

    t_start = _clock()
    n = min(g.num_pictures for g in gens)
    if n_frames > 0:
        n = min(n, n_frames)
    key = _key(gens, n, deblock_rows, intra_rows, device, engine_kwargs)
    path = cache_path()
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                hit = json.load(f).get(key)
            if hit:
                return {**hit, "cached": True, "seconds": _clock() - t_start}
        except (OSError, ValueError):
            pass
    if not measure:
        return {"deblock_waves": 8, "intra_waves": 8, "ms": {}, "cached": False, "seconds": _clock() - t_start}
    kw = dict(engine_kwargs or {})
    kw.update(deblock_waves=min(deblock_rows), intra_waves=min(intra_rows))  # strips fit every candidate
    eng = Engine(device=device, **kw)
    try:
        rows = []
        for g in gens:
            sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
            rows.append([eng.begin_raw(sid, *g.picture_raw(i)) for i in range(n)])

        def measure(d: int, i: int) -> float:
            eng.set_bands(d, i)
            batches = [eng.batch_create([r[f] for f in run for r in rows]) for run in independent_runs(rows)]
            try:
                for b in batches:  # graph capture
                    eng.batch_run(b)
                eng.sync()
                t0 = _clock()
                for _ in range(reps):
                    for b in batches:
                        eng.batch_run(b)
                eng.sync()
                return (_clock() - t0) / reps * 1e3
            finally:
                for b in batches:
                    eng.batch_destroy(b)

        ms = {}
        i0 = 8 if 8 in intra_rows else max(intra_rows)
        for d in deblock_rows:
            ms[(d, i0)] = measure(d, i0)
        best_d = min(deblock_rows, key=lambda d: ms[(d, i0)])
        for i in intra_rows:
            if (best_d, i) not in ms:
                ms[(best_d, i)] = measure(best_d, i)
        best = min(((best_d, i) for i in intra_rows), key=lambda k: ms[k])
    finally:
        eng.close()
    out = {"deblock_waves": best[0], "intra_waves": best[1],
           "ms": {f"{d}/{i}": round(v, 3) for (d, i), v in sorted(ms.items())}}
    if use_cache:
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            data = {}
            if os.path.exists(path):
                with open(path) as f:
                    data = json.load(f)
            data[key] = out
            with open(path + ".tmp", "w") as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(path + ".tmp", path)
        except (OSError, ValueError):
            pass
    return {**out, "cached": False, "seconds": _clock() - t_start}
