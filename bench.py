#!/usr/bin/env python3
"""Throughput of bit-exact H.264 picture reconstruction on one MI355X.

Workload (default): 64 independent streams of uneven size -- 16 x 352x288,
24 x 1280x720, 16 x 1920x1088, 8 x 3840x2160 -- each 16 frames of High-profile
IPBB syntax (8x8 transform on half the eligible MBs, explicit weighted
bi-prediction, partitions down to 4x4, 4 references), drawn per stream from
seed 0x5EED0004 + stream index by the library's generator.  All syntax is staged
in HBM before timing.

One step = the whole job: the 16 frame indices in decoding order, each as one
batched launch per stage (inter, intra, deblocking) over the 64 streams' pictures.
The same pictures are also run through the plain per-picture path (one launch per
stage per picture), and the batched output is checked byte-for-byte against it.

Prints one JSON line (metric, value = batched Mpix/s, plain Mpix/s, p50/p99 frame
latency, the roofline of the dominant kernel, the reference decoder's CPU
reconstruction timed on this host's cores, and end-to-end decoding of the reference's
own sample-stream prefixes, byte stream to host planes, beside the reference decoder).
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "bit-exact recon Mpix/s, batched vs plain on 1 GPU; p50/p99 frame latency"
HBM_PEAK_GBS = 8000.0
SEED4 = 0x5EED0004
MIX = [(16, 22, 18), (24, 80, 45), (16, 120, 68), (8, 240, 135)]  # (streams, width MBs, height MBs)
STAGES = ("inter", "intra", "bs", "deblock")
KERNELS = ("k_inter", "k_intra", "k_bs", "k_deblock")
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_config4.json")


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def stream_specs(frames):
    specs, sid = [], 0
    for n, W, H in MIX:
        for _ in range(n):
            specs.append({"width_mbs": W, "height_mbs": H, "n_frames": frames, "seed": SEED4 + sid, "spare_slots": 2})
            sid += 1
    return specs


def pct(vals, q):
    v = sorted(vals)
    return v[min(len(v) - 1, max(0, int(round(q / 100.0 * (len(v) - 1)))))]


def dpb_hash(eng, sids, slots):
    h = hashlib.sha256()
    for sid in sids:
        for s in range(slots[sid]):
            for p in eng.readback(sid, s):
                h.update(p.tobytes())
    return h.hexdigest()


def cpu_baseline(n_proc, frames):
    """The reference decoder's reconstruction (oracle/_ref/ref_harness: its own
    sources, timed around reconstruction + deblocking only) on a bounded sample of
    the workload, one process per core; falls back to the C oracle (port)."""
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    W, H = 80, 45
    seeds = [SEED4 + 16 + i for i in range(n_proc)]  # the first 720p streams of the mix
    sample = f"{n_proc} x 1280x720 streams of the mix (seeds 0x5EED0004+16..), first {frames} frames each"
    if os.path.exists(harness):
        with tempfile.TemporaryDirectory() as tmp:
            procs = []
            for i, sd in enumerate(seeds):
                d = os.path.join(tmp, str(i))
                os.makedirs(d)
                procs.append(subprocess.Popen([harness, "--config", "2", "--width", str(W), "--height", str(H),
                                               "--frames", str(frames), "--seed", str(sd), "--time", "--out", d],
                                              stdout=subprocess.PIPE, text=True))
            rates, mpix, secs = [], 0.0, 0.0
            for p in procs:
                out, _ = p.communicate(timeout=600)
                if p.returncode != 0:
                    raise RuntimeError("reference harness failed")
                r = json.loads(out.strip().splitlines()[-1])
                rates.append(r["mpix"] / r["recon_seconds"])
                mpix += r["mpix"]
                secs += r["recon_seconds"]
        return {"value": round(sum(rates), 3), "unit": "Mpix/s", "cores": n_proc, "kind": "reference",
                "sample": sample + "; reference sources compiled by oracle/Makefile, one process per core",
                "per_core_mpix_s": round(mpix / secs, 3), "cpu_seconds": round(secs, 2)}
    from h264_video_decoder_demo_amd import Generator
    from oracle.refcpu import Dpb

    def one(sd):
        g = Generator(2, width_mbs=W, height_mbs=H, n_frames=frames, seed=sd)
        dpb = Dpb(W, H, g.dpb_slots)
        pics = [g.picture(i) for i in range(g.num_pictures)]
        t = time.perf_counter()
        for desc, mbs, lv in pics:
            dpb.decode(desc, mbs, lv, 2)
        return W * H * 256 * len(pics) / 1e6, time.perf_counter() - t

    with cf.ThreadPoolExecutor(n_proc) as ex:
        res = list(ex.map(one, seeds))
    return {"value": round(sum(m / s for m, s in res), 3), "unit": "Mpix/s", "cores": n_proc, "kind": "port",
            "sample": sample + "; C oracle oracle/refcpu.c, one thread per core",
            "per_core_mpix_s": round(sum(m for m, _ in res) / sum(s for _, s in res), 3)}


def serving_check(specs, gens, args):
    """Bit-exactness of the serving leg, outside its timing: the same streams served again
    by the loop (continuous batching, same windows), each stream's whole DPB hashed just
    before the loop closes it, against the same stream decoded by the plain per-picture
    path (one picture at a time, one HIP stream) in a context of its own."""
    from h264_video_decoder_demo_amd import Engine
    from h264_video_decoder_demo_amd.serve import ServeLoop
    kw = dict(waves=args.waves, deblock_waves=args.deblock_waves, intra_waves=args.intra_waves)
    served = {}

    def on_close(i, eng, sid):
        served[i] = dpb_hash(eng, [sid], {sid: gens[i].dpb_slots})

    with Engine(device=0, dpb_budget_bytes=200 * 10**9, **kw) as eng:
        ServeLoop(eng, specs, window_ms=2.0, gens=gens, on_close=on_close).run()
    equal = 0
    with Engine(device=0, lane_split_mbs=-1, **kw) as eng:
        for i, g in enumerate(gens):
            sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
            for k in range(g.num_pictures):
                p = eng.begin_raw(sid, *g.picture_raw(k))
                eng.submit(p)
                eng.release(p)
            equal += dpb_hash(eng, [sid], {sid: g.dpb_slots}) == served.get(i)
            eng.close_stream(sid)
    if equal != len(gens):
        raise SystemExit(f"bench.py: {len(gens) - equal} served streams differ from their plain decode")
    return {"streams": len(gens), "equal_dpb": equal, "frames": sum(g.num_pictures for g in gens)}


BYTE_STREAMS = ("heavyhand_cavlc_15", "heavyhand_cabac_10", "gop121_12", "gop121_idc0_12", "heavyhand_cavlc_sl_8")


def byte_streams(reps):
    """End-to-end decoding of the reference's own sample streams (the prefixes committed
    under tests/golden): H264VideoDecoder.open on the Annex B file -- host parsing (NAL,
    slice headers, CAVLC / CABAC, POC, lists, MV derivation), GPU reconstruction, read-back
    of every output picture to host planes as the reference's callback receives them --
    best of `reps` runs after one warm-up, each output picture checked against the
    reference decoder's digests outside the timing; beside it the host parse alone and the
    reference decoder (oracle/_ref/ref_decode, every reference translation unit but
    main.cpp, compiled unmodified) on the same file, one process."""
    import warnings
    from h264_video_decoder_demo_amd import Engine
    from h264_video_decoder_demo_amd.bitstream import Bitstream
    from h264_video_decoder_demo_amd.decoder import H264VideoDecoder
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "ref_decode")
    out = {}
    with Engine(device=0) as eng, warnings.catch_warnings():
        warnings.simplefilter("ignore", RuntimeWarning)  # gop121's damaged slice, decoded as the reference does
        for name in BYTE_STREAMS:
            path = os.path.join(ROOT, "tests", "golden", name + ".h264")
            with open(os.path.join(ROOT, "tests", "golden", name + ".json")) as f:
                ref = json.load(f)["pictures"]
            got = []

            def keep(pic, user, err):
                if pic is not None:
                    got.append(pic)
                return 0

            def decode():
                got.clear()
                dec = H264VideoDecoder(eng)
                dec.set_output_frame_callback_functuin(keep, None)
                t = time.perf_counter()
                rc = dec.open(path)
                return time.perf_counter() - t, rc

            decode()
            best = min(decode()[0] for _ in range(reps))
            same = len(got) == len(ref) and all(
                (p.PicOrderCnt, hashlib.sha256(p.m_pic_buff_luma.tobytes()).hexdigest(),
                 hashlib.sha256(p.m_pic_buff_cb.tobytes()).hexdigest(),
                 hashlib.sha256(p.m_pic_buff_cr.tobytes()).hexdigest()) == (r["poc"], r["y"], r["cb"], r["cr"])
                for p, r in zip(got, ref))
            if not same:
                raise SystemExit(f"bench.py: {name} decoded differently from the reference decoder")
            t = time.perf_counter()
            n = sum(1 for _ in Bitstream(path).pictures())
            parse = time.perf_counter() - t
            mpix = sum(r["width"] * r["height"] for r in ref) / 1e6
            row = {"pictures": n, "mpix": round(mpix, 2), "decode_ms": round(1e3 * best, 1),
                   "fps": round(n / best, 1), "mpix_s": round(mpix / best, 1), "host_parse_ms": round(1e3 * parse, 1),
                   "bitexact_vs_reference_decoder": same}
            if os.path.exists(ref_bin):
                t = time.perf_counter()
                subprocess.run([ref_bin, path], check=True, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
                tr = time.perf_counter() - t
                row.update({"reference_decoder_ms": round(1e3 * tr, 1), "reference_decoder_fps": round(n / tr, 2),
                            "speedup": round(tr / best, 1)})
            out[name] = row
            log(f"byte stream {name}: {row}")
    return out


def serving_capacity(args, rates, seconds, bound_ms):
    """Serving capacity: the serving loop (128 streams of the serving mix, GOP 32, 2 ms
    windows, 200 GB DPB budget) at rising offered rates, each for `seconds`; capacity is
    the highest offered rate whose frame p99 latency stays at or under bound_ms.  Stops
    at the first rate over the bound."""
    from h264_video_decoder_demo_amd import Engine, Generator
    from h264_video_decoder_demo_amd.serve import ServeLoop, make_specs
    kw = dict(waves=args.waves, deblock_waves=args.deblock_waves, intra_waves=args.intra_waves)
    runs = []
    for fps in rates:
        specs = make_specs(128, total_fps=float(fps), duration_s=seconds, seed=0x5EED0005 + int(fps))
        with cf.ThreadPoolExecutor(16) as ex:
            gens = list(ex.map(lambda sp: Generator(2, width_mbs=sp.width_mbs, height_mbs=sp.height_mbs,
                                                    n_frames=sp.n_frames, seed=sp.seed, gop=32, spare_slots=2), specs))
        with Engine(device=0, dpb_budget_bytes=200 * 10**9, **kw) as eng:
            res = ServeLoop(eng, specs, window_ms=2.0, gens=gens).run()
        r = {"target_fps": fps, "offered_fps": round(res.offered_fps, 1), "served_fps": round(res.served_fps, 1),
             "frames": res.frames, "offered_mpix_s": round(res.mpix / seconds, 1), "p50_ms": round(res.pct(50), 3),
             "p99_ms": round(res.pct(99), 3), "max_batch": res.max_batch}
        r["ok"] = r["p99_ms"] <= bound_ms
        runs.append(r)
        log(f"capacity: {fps} fps offered -> p50 {r['p50_ms']} ms, p99 {r['p99_ms']} ms, max batch {r['max_batch']}")
        if not r["ok"]:
            break
    ok = [r for r in runs if r["ok"]]
    return {"p99_bound_ms": bound_ms, "seconds_per_rate": seconds,
            "capacity_fps": ok[-1]["offered_fps"] if ok else None,
            "capacity_mpix_s": ok[-1]["offered_mpix_s"] if ok else None, "runs": runs}


def summarize_pmc(fetch_csv: str, write_csv: str, copy_bytes: int, out_json: str) -> dict:
    """HBM traffic per launch from two rocprofv3 --pmc runs of `bench.py --profile-run`
    (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch), each scaled by what the counter
    reads for the calibration copy of copy_bytes (read once, written once)."""
    import csv
    import collections

    def per_kernel(path, counter):
        acc = collections.defaultdict(list)
        with open(path) as f:
            for row in csv.DictReader(f):
                if row.get("Counter_Name") != counter:
                    continue
                name = row["Kernel_Name"]
                for k in KERNELS + ("k_copy", "k_bs", "k_deblock_band", "k_intra_band", "k_deblock_rows"):
                    # the frame-picture instance (k_inter<false, false>: no MBAFF, flat lists)
                    if f"::{k}(" in name or f"::{k}<false>(" in name or f"::{k}<false, false>(" in name:
                        acc[k].append(float(row["Counter_Value"]))
        return {k: sum(v) / len(v) for k, v in acc.items()}

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    f_scale = fetch["k_copy"] * 1024.0 / copy_bytes   # counter bytes per true byte
    w_scale = write["k_copy"] * 1024.0 / copy_bytes
    out = {"fetch_scale": f_scale, "write_scale": w_scale, "copy_bytes": copy_bytes}
    for k in fetch:
        if k == "k_copy":
            continue
        out[k] = round(fetch[k] * 1024.0 / f_scale + write.get(k, 0.0) * 1024.0 / w_scale)
    with open(out_json, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed passes over the whole 64-stream job")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=16)
    ap.add_argument("--plain-steps", type=int, default=1)
    ap.add_argument("--cpu-procs", type=int, default=16)
    ap.add_argument("--cpu-frames", type=int, default=16)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--no-graphs", action="store_true", help="launch batch stages directly instead of hipGraph replay")
    ap.add_argument("--deblock", choices=["bands", "mb_wavefront", "mb_rows"], default="bands",
                    help="deblocking schedule")
    ap.add_argument("--deblock-waves", type=int, default=0, help="MB rows per band of the bands schedule (0 = 8)")
    ap.add_argument("--intra", choices=["bands", "mb_wavefront"], default="bands", help="intra schedule")
    ap.add_argument("--intra-waves", type=int, default=0, help="MB rows per band of the bands intra schedule (0 = 8)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default band heights instead of autotuning them on the workload's first pictures")
    ap.add_argument("--serve-seconds", type=float, default=3.0,
                    help="length of the serving run (Poisson arrivals, 128 streams, 400 frames/s); 0 = skip")
    ap.add_argument("--capacity-rates", default="800,1600,3200,6400",
                    help="offered rates (frames/s) of the serving capacity sweep, rising; empty = skip")
    ap.add_argument("--capacity-seconds", type=float, default=1.0)
    ap.add_argument("--capacity-p99-ms", type=float, default=50.0)
    ap.add_argument("--profile-run", action="store_true",
                    help="short run for rocprofv3: 1 timed step, no warmup, no latency / plain / CPU legs, "
                         "then the library's calibration copy of --copy-mib MiB")
    ap.add_argument("--one-lane", action="store_true",
                    help="batched run in a one-lane context (every stage, deblocking included, on one HIP "
                         "stream): the isolated shape whose kernel durations add up to the step time")
    ap.add_argument("--no-isolated", action="store_true", help="skip the isolated (one-lane) batched leg")
    ap.add_argument("--no-byte-streams", action="store_true",
                    help="skip the end-to-end decoding of the reference's sample-stream prefixes")
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--summarize-pmc", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="write profiles/traffic_config4.json from two rocprofv3 --pmc runs of --profile-run")
    args = ap.parse_args()
    if args.summarize_pmc:
        print(json.dumps(summarize_pmc(*args.summarize_pmc, args.copy_mib << 20, TRAFFIC_FILE)))
        return
    if args.profile_run:
        args.steps, args.warmup, args.no_plain, args.no_cpu_baseline, args.no_isolated = 1, 0, True, True, True
        args.no_byte_streams = True
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this engine runs on one GPU; streams would be spread over GPUs by an outer "
              "scheduler (no collectives involved). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    from h264_video_decoder_demo_amd import Engine, Generator
    from h264_video_decoder_demo_amd.engine import independent_runs

    specs = stream_specs(args.frames)
    t = time.perf_counter()
    with cf.ThreadPoolExecutor(16) as ex:
        gens = list(ex.map(lambda s: Generator(2, **s), specs))
    log(f"generated {len(gens)} streams in {time.perf_counter() - t:.1f} s")

    # band heights: explicit, or picked by the autotuner on this workload (outside the
    # timed region; cached per machine).  --profile-run only looks the choice up, so
    # that the profiled kernels are the measured run's alone: run bench.py once first.
    tune = None
    if (args.deblock == "bands" and args.intra == "bands" and not args.deblock_waves and not args.intra_waves
            and not args.no_autotune):
        from h264_video_decoder_demo_amd.autotune import autotune
        tune = autotune(gens, measure=not args.profile_run, engine_kwargs=dict(waves=args.waves, graphs=not args.no_graphs))
        args.deblock_waves, args.intra_waves = tune["deblock_waves"], tune["intra_waves"]
        log(f"autotune: bands of {args.deblock_waves} / {args.intra_waves} rows (deblocking / intra) in "
            f"{tune['seconds']:.1f} s{' (cached)' if tune['cached'] else ''}: {tune.get('ms')}")

    schedule = {"bands": 0, "mb_wavefront": 1, "mb_rows": 2}[args.deblock]
    eng = Engine(device=0, profile=False, waves=args.waves, graphs=not args.no_graphs, deblock_schedule=schedule,
                 deblock_waves=args.deblock_waves, intra_schedule={"bands": 0, "mb_wavefront": 1}[args.intra],
                 intra_waves=args.intra_waves, lane_split_mbs=-1 if args.one_lane else 0)
    sids, slots, pics = [], {}, []
    alg = [0, 0, 0, 0]
    for g in gens:
        sid = eng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
        sids.append(sid)
        slots[sid] = g.dpb_slots
        row = []
        for i in range(g.num_pictures):
            p = eng.begin_raw(sid, *g.picture_raw(i))
            ab = eng.alg_bytes(p)
            alg[0] += ab[0]
            alg[1] += ab[1]
            alg[3] += ab[2]
            alg[2] += g.width_mbs * g.height_mbs * (28 + 172 + 32)  # bS pass: aux + records in, 32 B out per MB
            row.append(p)
        pics.append(row)
    nframes = min(len(r) for r in pics)
    mpix_step = sum(g.width_mbs * g.height_mbs * 256 * nframes for g in gens) / 1e6
    # one batch per run of mutually independent frame indices (I0 | P3 | B1 B2 P6 | ...):
    # each stream's pictures are launched by dependency level, not one at a time
    runs_f = independent_runs(pics)
    batches = [eng.batch_create([row[f] for f in run for row in pics]) for run in runs_f]
    n_batches = len(batches)
    log(f"batches per step: {n_batches} (frame index runs {runs_f})")
    log(f"staged {sum(len(r) for r in pics)} pictures ({mpix_step:.1f} Mpix per step) in "
        f"{time.perf_counter() - t:.1f} s")

    # ---- batched: warmup (captures each batch's stage graph), then K timed passes
    for _ in range(max(1, args.warmup)):
        for b in batches:
            eng.batch_run(b)
    eng.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for b in batches:
            eng.batch_run(b)
    eng.sync()
    t_batched = time.perf_counter() - t0
    batched = mpix_step * args.steps / t_batched
    h_batched = dpb_hash(eng, sids, slots)
    log(f"batched: {batched:.1f} Mpix/s ({1e3 * t_batched / args.steps:.2f} ms per step)")

    # ---- the same K passes with HIP events around every stage launch (direct launches):
    # per-stage kernel durations for the roofline
    eng.set_profile(True)
    eng.stage_times()  # reset the accumulators
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for b in batches:
            eng.batch_run(b)
    eng.sync()
    t_events = time.perf_counter() - t0
    runs, stage_ms = eng.stage_times()  # one run per lane per batch
    eng.set_profile(False)
    log(f"batched with events: {mpix_step * args.steps / t_events:.1f} Mpix/s; "
        f"stage ms per step {[round(x / args.steps, 3) for x in stage_ms]}")

    if args.profile_run:
        from h264_video_decoder_demo_amd._lib import check, lib
        check(lib().h264g_calibration_copy(eng._h, args.copy_mib << 20, 1))
        log(f"calibration copy of {args.copy_mib} MiB done")
        print(json.dumps({"profile_run": True, "mpix_per_step": mpix_step, "seconds": t_batched,
                          "alg_bytes_per_step": alg, "copy_bytes": args.copy_mib << 20}), flush=True)
        return

    # ---- batch latency: each batch submitted and waited for alone
    lat_b = []
    for _ in range(max(1, args.steps)):
        for b in batches:
            t1 = time.perf_counter()
            eng.batch_run(b)
            eng.sync()
            lat_b.append(1e3 * (time.perf_counter() - t1))

    plain = None
    lat_p = []
    isolated = None
    peng = None
    if not args.no_plain or not args.no_isolated:
        # a one-lane context (every stage on one HIP stream, deblocking included) holding
        # the same pictures with DPBs of its own, which must end up identical to the
        # batched context's
        peng = Engine(device=0, waves=args.waves, deblock_schedule=schedule, deblock_waves=args.deblock_waves,
                      intra_schedule={"bands": 0, "mb_wavefront": 1}[args.intra], intra_waves=args.intra_waves,
                      lane_split_mbs=-1)
        psids, pslots, ppics = [], {}, []
        for g in gens:
            sid = peng.open_stream(g.width_mbs, g.height_mbs, g.dpb_slots)
            psids.append(sid)
            pslots[sid] = g.dpb_slots
            ppics.append([peng.begin_raw(sid, *g.picture_raw(i)) for i in range(nframes)])
    if not args.no_isolated:
        # ---- isolated: the batched launches of the same job in the one-lane context, with
        # HIP events around every stage.  Nothing runs beside a kernel, so each kernel's
        # average duration is its own and the stage durations add up to at most the step.
        pbatches = [peng.batch_create([row[f] for f in run for row in ppics]) for run in runs_f]
        for b in pbatches:
            peng.batch_run(b)
        peng.sync()
        peng.set_profile(True)
        peng.stage_times()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            for b in pbatches:
                peng.batch_run(b)
        peng.sync()
        t_iso = time.perf_counter() - t0
        iso_runs, iso_ms = peng.stage_times()
        peng.set_profile(False)
        if dpb_hash(peng, psids, pslots) != h_batched:
            raise SystemExit("bench.py: batched runs in the one-lane and two-lane contexts differ")
        for b in pbatches:
            peng.batch_destroy(b)
        isolated = {"ms_per_step": 1e3 * t_iso / args.steps, "runs": iso_runs, "stage_ms": iso_ms}
        log(f"isolated (one lane): {mpix_step * args.steps / t_iso:.1f} Mpix/s, {isolated['ms_per_step']:.2f} ms "
            f"per step; stage ms per step {[round(x / args.steps, 3) for x in iso_ms]} "
            f"(sum {sum(iso_ms) / args.steps:.2f})")
    if not args.no_plain:
        # ---- plain: the reference's shape -- one picture at a time, one launch per stage
        # per picture, all on one HIP stream
        t0 = time.perf_counter()
        for _ in range(args.plain_steps):
            for f in range(nframes):
                for row in ppics:
                    peng.submit(row[f])
        peng.sync()
        t_plain = time.perf_counter() - t0
        plain = mpix_step * args.plain_steps / t_plain
        h_plain = dpb_hash(peng, psids, pslots)
        if h_plain != h_batched:
            raise SystemExit("bench.py: batched and plain outputs differ")
        for f in range(nframes):
            for row in ppics:
                t1 = time.perf_counter()
                peng.submit(row[f])
                peng.sync()
                lat_p.append(1e3 * (time.perf_counter() - t1))
        log(f"plain: {plain:.1f} Mpix/s; batched/plain = {batched / plain:.1f}x")
    if peng is not None:
        peng.close()

    # ---- roofline of the dominant kernel (algorithmic bytes / event-timed duration)
    # Each batch launches one stage sequence per lane with pictures (runs counts
    # them); a kernel's algorithmic bytes over its summed launch durations is its rate.
    # Durations come from the isolated leg when it ran (one lane: nothing beside the
    # kernel), else from the two-lane run (durations include the other lane's kernels).
    def stage_table(runs_, ms_):
        launches_ = max(1, runs_)
        out_ = {}
        for k in range(4):
            ms = ms_[k] / launches_
            out_[STAGES[k]] = {"avg_launch_ms": round(ms, 4),
                               "alg_bytes_per_launch": int(alg[k] * args.steps / launches_),
                               "alg_GBs": round(alg[k] * args.steps / (ms_[k] * 1e-3) / 1e9, 1) if ms > 0 else None}
        return launches_, out_

    src_runs, src_ms = (isolated["runs"], isolated["stage_ms"]) if isolated else (runs, stage_ms)
    dom = max(range(4), key=lambda k: src_ms[k])
    launches, stages = stage_table(src_runs, src_ms)
    per_launch_bytes = alg[dom] * args.steps / launches
    avg_ms = src_ms[dom] / launches
    achieved = per_launch_bytes / (avg_ms * 1e-3) / 1e9
    kname = {0: "k_inter", 1: "k_intra_band" if args.intra == "bands" else "k_intra", 2: "k_bs",
             3: {"bands": "k_deblock_band", "mb_wavefront": "k_deblock", "mb_rows": "k_deblock_rows"}[args.deblock]}[dom]
    traffic = None
    if os.path.exists(TRAFFIC_FILE):
        with open(TRAFFIC_FILE) as f:
            traffic = json.load(f).get(kname)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": kname,
                "alg_bytes_per_launch": int(per_launch_bytes), "avg_launch_ms": round(avg_ms, 4),
                "timing": "isolated: one-lane context, one HIP stream" if isolated else "two lanes (concurrent)"}
    mc_dbk = (alg[0] + alg[3]) * args.steps / ((src_ms[0] + src_ms[3]) * 1e-3) / 1e9
    _, stages_mix = stage_table(runs, stage_ms)

    serve = None
    if args.serve_seconds > 0:
        # ---- serving: 128 streams of the serving mix, GOP 32, Poisson arrivals of
        # 400 frames/s in total, 2 ms batch windows, DPB budget 200 GB
        from h264_video_decoder_demo_amd.serve import ServeLoop, make_specs
        for b in batches:
            eng.batch_destroy(b)
        batches = []
        for row in pics:
            for p in row:
                eng.release(p)
        for sid in sids:
            eng.close_stream(sid)
        specs = make_specs(128, total_fps=400.0, duration_s=args.serve_seconds)
        t = time.perf_counter()
        with cf.ThreadPoolExecutor(16) as ex:
            sgens = list(ex.map(lambda sp: Generator(2, width_mbs=sp.width_mbs, height_mbs=sp.height_mbs,
                                                     n_frames=sp.n_frames, seed=sp.seed, gop=32, spare_slots=2),
                                specs))
        log(f"serving: generated {sum(g.num_pictures for g in sgens)} frames of {len(specs)} streams in "
            f"{time.perf_counter() - t:.1f} s")
        seng = Engine(device=0, dpb_budget_bytes=200 * 10**9, waves=args.waves, deblock_schedule=schedule,
                      deblock_waves=args.deblock_waves, intra_schedule={"bands": 0, "mb_wavefront": 1}[args.intra],
                      intra_waves=args.intra_waves)
        res = ServeLoop(seng, specs, window_ms=2.0, gens=sgens).run()
        seng.close()
        serve = {"streams": len(specs), "target_fps": 400.0, "offered_fps": round(res.offered_fps, 1),
                 "served_fps": round(res.served_fps, 1), "frames": res.frames, "seconds": round(res.wall_s, 3),
                 "mpix_s": round(res.mpix_per_s, 1), "p50_ms": round(res.pct(50), 3), "p99_ms": round(res.pct(99), 3),
                 "max_batch": res.max_batch, "window_ms": 2.0, "dpb_budget_gb": 200,
                 "note": "mpix_s is the offered load served (the loop keeps up), not capacity"}
        log(f"serving: {res.frames} frames at {serve['offered_fps']} fps offered, p50 {serve['p50_ms']} ms, "
            f"p99 {serve['p99_ms']} ms")
        serve["bitexact_vs_plain"] = serving_check(specs, sgens, args)
        log(f"serving check: {serve['bitexact_vs_plain']}")
        rates = [int(x) for x in args.capacity_rates.split(",") if x.strip()]
        if rates:
            serve["capacity"] = serving_capacity(args, rates, args.capacity_seconds, args.capacity_p99_ms)
            log(f"serving capacity: {serve['capacity']['capacity_fps']} fps at p99 <= {args.capacity_p99_ms} ms")

    streams_e2e = None if args.no_byte_streams else byte_streams(3)

    cpu = None
    if not args.no_cpu_baseline:
        t = time.perf_counter()
        cpu = cpu_baseline(args.cpu_procs, args.cpu_frames)
        log(f"cpu baseline: {cpu['value']} Mpix/s on {cpu['cores']} cores ({time.perf_counter() - t:.1f} s)")

    out = {
        "metric": METRIC, "value": round(batched, 1), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * t_batched / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8", "data": "synthetic",
        "config": {"workload": "64 streams x 16 frames, mixed 352x288/1280x720/1920x1088/3840x2160 "
                               "High-profile IPBB syntax, batched launch per stage per run of independent frames",
                   "streams": len(gens), "frames_per_stream": nframes, "pictures_per_step": len(gens) * nframes,
                   "batches_per_step": n_batches,
                   "mpix_per_step": round(mpix_step, 3)},
        "graphs": not args.no_graphs,
        "deblock_schedule": args.deblock,
        "intra_schedule": args.intra,
        "band_rows": {"deblock": args.deblock_waves or 8, "intra": args.intra_waves or 8,
                      "autotune": ({"ms_per_sample": tune["ms"], "seconds": round(tune["seconds"], 1),
                                    "cached": tune["cached"]} if tune and not tune["cached"] else
                                   ({"cached": True} if tune else None))},
        "batched_event_timed_value": round(mpix_step * args.steps / t_events, 1),
        "plain_value": round(plain, 1) if plain else None,
        "batched_over_plain": round(batched / plain, 2) if plain else None,
        "latency_ms": {"serving_frame_p50": serve["p50_ms"] if serve else None,
                       "serving_frame_p99": serve["p99_ms"] if serve else None,
                       "batched_frame_p50": round(pct(lat_b, 50), 3), "batched_frame_p99": round(pct(lat_b, 99), 3),
                       "plain_picture_p50": round(pct(lat_p, 50), 3) if lat_p else None,
                       "plain_picture_p99": round(pct(lat_p, 99), 3) if lat_p else None},
        "serving": serve,
        "lanes": "streams of >= 16384 MBs per picture on a second HIP stream (plain: one stream)",
        "stage_launches": launches,
        "stages": stages,
        "mc_deblock_alg_GBs": round(mc_dbk, 1),
        "isolated": ({"ms_per_step": round(isolated["ms_per_step"], 3),
                      "stage_ms_per_step": [round(x / args.steps, 3) for x in isolated["stage_ms"]],
                      "sum_stage_ms_per_step": round(sum(isolated["stage_ms"]) / args.steps, 3),
                      "value": round(mpix_step / (isolated["ms_per_step"] * 1e-3), 1),
                      "mc_deblock_alg_GBs_per_step": round((alg[0] + alg[3]) / (isolated["ms_per_step"] * 1e-3) / 1e9, 1)}
                     if isolated else None),
        "stages_two_lanes": stages_mix,
        # both GPU paths against each other; the reference's pictures pin them in the GPU tests
        # (tests/test_gpu_parity.py: every golden case, including full 1280x720, 1920x1088 and
        # 3840x2160 pictures, on the plain and the batched path)
        "bitexact": {"batched_dpb_equals_plain_dpb": (plain is not None), "dpb_sha256": h_batched[:16],
                     "serving_dpb_equals_plain_dpb": serve["bitexact_vs_plain"] if serve else None},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "byte_streams": streams_e2e,
    }
    for b in batches:
        eng.batch_destroy(b)
    print(json.dumps(out), flush=True)
    eng.close()


if __name__ == "__main__":
    main()
