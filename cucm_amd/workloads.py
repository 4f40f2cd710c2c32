"""Benchmark workloads beyond the default stream (used by bench.py).

batch_workload (bench.py's default): 10,000 independent files, lognormal
sizes (median 64 KiB, sigma 1.5, clipped to [512 B, 64 MiB]), file i
generated by the order-2 Markov source seeded 1000 + i, with the settings of
cucm_amd/batch_settings.json (chunks of at most 128 KiB, 2^18 slots, the
launch settings measured on the full batch) or freshly autotuned.  All files
are encoded, then decoded, in one persistent launch each (device-resident
buffers); the plain path issues one launch per file, as the reference does
per batch, and is timed on a seeded random sample of the same files because
each single-file launch runs alone.

serve_workload: Poisson arrivals of mixed compress/decompress requests
through the continuous-batching serving loop (cucm_serve_*), p50/p99 per
load level beside the one-request-at-a-time plain path.
"""
import json
import os
import sys
import time

import numpy as np

from . import datagen
from .codec import Context, make_params, output_cap

MiB = 1 << 20


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def pick_chunk_size(n, s_max):
    """Per-file chunk size: one chunk for small files (no tail padding), S_max
    otherwise.  The chunking is part of the stream; the oracle uses the same."""
    return int(max(1, min(n, s_max)))


METRIC = "compress/decompress MB/s on 1 GPU at identical ratio; batched vs plain; p99"
HBM_PEAK_GBS = 8000.0
ALG_BYTES_PER_INPUT_BYTE = 128.0
BATCH_TRAFFIC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                             "profiles", "batch_traffic.json")


def batch_files(count=10000):
    """The uneven-size batch: count files, sizes lognormal (median 64 KiB,
    sigma 1.5) clipped to [512 B, 64 MiB] (seed 3), file i from the order-2
    Markov source seeded 1000 + i.  Returns (sizes, arena, offsets)."""
    sizes = datagen.lognormal_sizes(count, 65536, 1.5, 512, 64 * MiB, 3)
    arena, offs = datagen.markov2_files(sizes, seed0=1000)
    return sizes, arena, offs


def batch_kernel_name(launch, direction):
    if direction == "decoder":
        return "decode_team_kernel" if launch["team"] else "codec_kernel<decode>"
    return "encode_team_kernel" if launch["team"] else "codec_kernel<encode>"


def batch_workload(args):
    """The uneven-size batch, every file compressed in one persistent launch
    and decompressed in another per step (device-resident buffers), with the
    stream and launch settings of cucm_amd/batch_settings.json (or freshly
    autotuned with --autotune).  Beside it: the plain path (one launch per
    file, as the reference issues one launch sequence per batch) on a seeded
    sample of the same files, the CPU baseline (the oracle, files as tasks,
    all host cores and one), the roofline of the dominant kernel, and the
    serving mix's latencies (serve_measure)."""
    from .autotune import apply_launch, autotune_batch, load_settings, save_settings
    count = getattr(args, "files", 10000) or 10000
    t0 = time.time()
    sizes, arena, offs = batch_files(count)
    total = int(sizes.sum())
    log("generated %d files, %.1f MiB in %.1fs" % (count, total / MiB, time.time() - t0))
    ctx = Context()
    if getattr(args, "autotune", False):
        tuned = autotune_batch(ctx, arena, offs, sizes, log=log)
        if getattr(args, "save_settings", False):
            save_settings(tuned, note="autotuned on %d of %d files" % (tuned["sample_files"], count))
        cfg = tuned
        source = "autotuned (%d-file sample)" % tuned["sample_files"]
    else:
        cfg = load_settings()
        source = "cucm_amd/batch_settings.json"
    s_max, tb = cfg["chunk_max"], cfg["table_bits"]
    enc, dec = cfg["encoder"], cfg["decoder"]
    spec = dec["spec_bits"]
    params = [make_params(pick_chunk_size(int(n), s_max), 128, 100, 0, tb) for n in sizes]
    d_in = ctx.alloc(total)
    d_in.upload(arena)
    caps = [output_cap(int(n), p) for n, p in zip(sizes, params)]
    coffs = np.concatenate([[0], np.cumsum(np.array(caps, dtype=np.uint64) + 255 & ~np.uint64(255))])
    d_c = ctx.alloc(int(coffs[-1]))
    d_o = ctx.alloc(total)
    ins = [d_in.addr + int(o) for o in offs]
    cs = [d_c.addr + int(o) for o in coffs[:-1]]
    outs = [d_o.addr + int(o) for o in offs]
    lens = [int(n) for n in sizes]

    def step():
        apply_launch(ctx, enc)
        a = time.perf_counter()
        cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
        b = time.perf_counter()
        sc = ctx.stats()
        apply_launch(ctx, dec)
        dl, dst = ctx.decompress_many_device(cs, cl, outs, lens, spec)
        c = time.perf_counter()
        sd = ctx.stats()
        bad = [(i, x, y) for i, (x, y) in enumerate(zip(st, dst)) if x or y]
        if bad:
            raise RuntimeError("batch: %d files failed, first (file, compress, decompress) "
                               "status: %s" % (len(bad), bad[:5]))
        return cl, b - a, c - b, sc, sd

    for i in range(args.warmup):
        step()
        log("warmup %d done" % i)
    ctx.synchronize()
    tc, td, kc, kd = [], [], [], []
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        cl, x, y, sc, sd = step()
        tc.append(x)
        td.append(y)
        kc.append(sc["codec_ms"])
        kd.append(sd["codec_ms"])
        log("step %d: compress %.3fs (kernel %.1f ms)  decompress %.3fs (kernel %.1f ms)"
            % (i, x, sc["codec_ms"], y, sd["codec_ms"]))
    ctx.synchronize()
    t_all = time.perf_counter() - t_all0
    back = np.frombuffer(d_o.download(total), dtype=np.uint8)
    lossless = bool(np.array_equal(back, arena))
    comp_total = int(sum(cl))
    res = {
        "metric": METRIC, "value": total / 1e6 / (t_all / args.steps), "unit": "MB/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1000.0 * t_all / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8/int32 fixed-point",
        "data": "synthetic (seeded generator, resident in HBM)",
        "config": {"workload": "batch: %d independent files, lognormal sizes (median 64 KiB, "
                               "sigma 1.5, clipped to [512 B, 64 MiB]), order-2 Markov content "
                               "(file i seeded 1000 + i); chunk = min(file, %d KiB), 2^%d "
                               "slots/order, U=128, P=0; one persistent launch per direction"
                               % (count, s_max // 1024, tb),
                   "files": count, "bytes": total, "chunks": int(sc["tasks"]),
                   "chunk_max": s_max, "table_bits": tb, "unit_bits": 128, "batch_chunks": 100,
                   "encoder_launch": enc, "decoder_launch": dec, "settings": source},
        "compress_MBps": total / 1e6 / float(np.median(tc)),
        "decompress_MBps": total / 1e6 / float(np.median(td)),
        "compress_kernel_ms": float(np.median(kc)),
        "decompress_kernel_ms": float(np.median(kd)),
        "ratio": total / comp_total, "lossless": lossless, "container_bytes": comp_total,
        "compress_waves": int(sc["waves"]), "compress_teams": int(sc["teams"]),
        "decompress_waves": int(sd["waves"]), "decompress_teams": int(sd["teams"]),
        "model_bytes": int(max(sc["model_bytes"], sd["model_bytes"])),
    }
    res["compress_kernel_MBps"] = total / 1e3 / res["compress_kernel_ms"]
    res["decompress_kernel_MBps"] = total / 1e3 / res["decompress_kernel_ms"]
    # dominant kernel: the slower launch
    if res["decompress_kernel_ms"] >= res["compress_kernel_ms"]:
        kname, kms = batch_kernel_name(dec, "decoder"), res["decompress_kernel_ms"]
    else:
        kname, kms = batch_kernel_name(enc, "encoder"), res["compress_kernel_ms"]
    achieved = ALG_BYTES_PER_INPUT_BYTE * total / (kms / 1e3) / 1e9
    res["roofline"] = dict({"bound": "hbm", "kernel": kname, "achieved": achieved,
                            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS},
                           **batch_traffic(kname, total, kms / 1e3))
    if getattr(args, "phases", False):
        # one more decode in the profiling instantiation (per-phase shader
        # clocks of every wave, summed; a few % slower than the timed kernel):
        # where a decoded bit's time goes
        apply_launch(ctx, dec)
        ctx.set_phase_profile(True)
        ctx.decompress_many_device(cs, cl, outs, lens, spec)
        ph, sp = ctx.phase_cycles(), ctx.stats()
        ctx.set_phase_profile(False)
        per_bit = {k: v / (8.0 * total) for k, v in ph.items() if v}
        res["decoder_phases"] = {"kernel_ms": sp["codec_ms"], "waves": int(sp["waves"]),
                                 "spec_bits": int(sp["spec_bits"]),
                                 "wave_clocks_per_decoded_bit": per_bit}
        log("decoder phases (%.1f ms): " % sp["codec_ms"]
            + ", ".join("%s %.0f" % (k, v) for k, v in per_bit.items()))
    # plain path: one launch per file on a seeded sample of the same files
    plain_files = getattr(args, "plain_files", 40)
    if plain_files:
        res.update(plain_sample(ctx, ins, lens, params, cs, caps, outs, cl, enc, dec, spec,
                                plain_files, count))
        res["batched_vs_plain_compress"] = res["compress_MBps"] / res["plain_compress_MBps"]
        res["batched_vs_plain_decompress"] = res["decompress_MBps"] / res["plain_decompress_MBps"]
    res["cpu_baseline"] = None
    res["oracle_sample_identical"] = None
    if not getattr(args, "no_cpu", False):
        res["cpu_baseline"], res["oracle_sample_identical"] = batch_cpu_baseline(
            arena, offs, sizes, d_c, coffs, cl, s_max, tb, args)
    # the serving mix (Poisson arrivals, lognormal sizes): the serving kernel
    # against static-batch lanes at the same loads
    nreq = getattr(args, "serve_requests", 240)
    res["serve"] = None
    if nreq:
        for b in (d_in, d_c, d_o):
            b.free()
        apply_launch(ctx, {"team": None, "waves_per_block": 0})  # automatic, as for any caller
        res["serve"] = serve_measure(ctx, nreq, loads=(0.5, 1.0))
    return res


def batch_traffic(kernel, n, seconds):
    """HBM bytes per launch of the dominant kernel on this workload, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE / TCC_HIT / TCC_MISS runs of
    the same command (FETCH_SIZE scaled by the copy-kernel calibration), with
    the physical rate and the L2 hit rate."""
    if not os.path.exists(BATCH_TRAFFIC):
        return {"traffic": None}
    d = json.load(open(BATCH_TRAFFIC))
    t = d.get("kernels", {}).get(kernel)
    if t is None or t.get("input_bytes") != n:
        return {"traffic": None}
    traffic = t["fetch_bytes"] + t["write_bytes"]
    out = {"traffic": traffic, "hbm_GBps": traffic / seconds / 1e9,
           "hbm_frac_of_peak": traffic / seconds / 1e9 / HBM_PEAK_GBS,
           "l2_hit_rate": t.get("l2_hit_rate"), "traffic_source": d.get("source")}
    ceil = d.get("ceiling")
    if ceil:
        # the same access stream replayed without the model arithmetic
        out["replay_ceiling_MBps"] = ceil["MBps"]
        out["replay_ceiling_frac"] = n / 1e6 / seconds / ceil["MBps"]
    return out


def plain_sample(ctx, ins, lens, params, cs, caps, outs, cl, enc, dec, spec, k, count):
    from .autotune import apply_launch
    rng = np.random.default_rng(7)
    sample = rng.choice(count, size=min(k, count), replace=False)
    sin = [ins[i] for i in sample]
    slen = [lens[i] for i in sample]
    sp = [params[i] for i in sample]
    scs = [cs[i] for i in sample]
    scap = [caps[i] for i in sample]
    souts = [outs[i] for i in sample]
    apply_launch(ctx, enc)
    d = time.perf_counter()
    pl, _ = ctx.compress_many_device(sin, slen, sp, scs, scap, plain=True)
    e = time.perf_counter()
    apply_launch(ctx, dec)
    ctx.decompress_many_device(scs, pl, souts, slen, spec, plain=True)
    f = time.perf_counter()
    sample_bytes = sum(slen)
    return {"plain_sample_files": len(sample), "plain_sample_bytes": int(sample_bytes),
            "plain_compress_MBps": sample_bytes / 1e6 / (e - d),
            "plain_decompress_MBps": sample_bytes / 1e6 / (f - e),
            "plain_identical_to_batched": bool(all(pl[j] == cl[i] for j, i in enumerate(sample)))}


def _median_time(fn, runs=3):
    fn()  # warm-up
    ts = []
    for _ in range(runs):
        t = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t)
    return float(np.median(ts))


def batch_cpu_baseline(arena, offs, sizes, d_c, coffs, cl, s_max, tb, args):
    """The oracle (the CPU restatement of the reference model) on size-stratified
    samples of the same files, one file per task (OpenMP over files, each file
    single-threaded): on every host thread the run may use, and on one; one
    warm-up, then the median of 3 runs per direction.  Its containers are
    compared with the GPU's for the same files."""
    from oracle import oracle as O
    from .autotune import stratified_sample
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or \
        (os.cpu_count() or 1)
    # (on the GPU box OMP_NUM_THREADS is the host share of one GPU, 16 threads;
    # os.cpu_count() reports the whole host)
    out = {"unit": "MB/s (compress)", "kind": "port", "cores": threads,
           "host_cpu_count": os.cpu_count(),
           "threads_source": "--cpu-threads" if args.cpu_threads else
           ("OMP_NUM_THREADS" if os.environ.get("OMP_NUM_THREADS") else "os.cpu_count()")}
    same = True
    for label, nthr, budget in (("all", threads, args.cpu_mib * MiB),
                                ("one", 1, max(args.cpu_mib * MiB // 12, 1 << 20))):
        # files small enough that one file per task balances over nthr
        # threads, taken across their size distribution, largest first
        cap_file = max(budget // (2 * nthr), 64 * 1024)
        pool = np.nonzero(sizes <= cap_file)[0]
        sub_idx = stratified_sample(sizes[pool], budget, seed=11)
        idx = sorted((int(pool[j]) for j in sub_idx), key=lambda i: -int(sizes[i]))
        sl = np.array([int(sizes[i]) for i in idx], dtype=np.uint64)
        so = np.concatenate([[0], np.cumsum(sl)])[:-1].astype(np.uint64)
        sub = np.concatenate([arena[int(offs[i]):int(offs[i]) + int(sizes[i])] for i in idx])
        res = {}

        def comp():
            res["c"] = O.compress_files(sub, so, sl, s_max, 128, 100, tb, threads=nthr)

        tcomp = _median_time(comp)
        blobs, bo, bl = res["c"]

        def decomp():
            res["d"] = O.decompress_files(blobs, bo, bl, sl, threads=nthr)

        tdec = _median_time(decomp)
        ok = bool(np.array_equal(res["d"], sub))
        for j, i in enumerate(idx):
            gpu = d_c.download(int(cl[i]), offset=int(coffs[i]))
            if gpu != blobs[int(bo[j]):int(bo[j] + bl[j])].tobytes():
                same = False
        mb = float(sl.sum()) / 1e6
        if label == "all":
            out.update({"value": mb / tcomp, "decompress_MBps": mb / tdec, "lossless": ok,
                        "sample": "%d of the files up to %d KiB (%.1f MB, size-stratified), "
                                  "oracle compress_files/decompress_files, one file per task "
                                  "on %d threads, largest first; warm-up + median of 3"
                                  % (len(idx), cap_file // 1024, mb, nthr)})
        else:
            out.update({"one_thread_compress_MBps": mb / tcomp,
                        "one_thread_decompress_MBps": mb / tdec,
                        "one_thread_sample": "%d files up to %d KiB (%.1f MB), 1 thread; "
                                             "warm-up + median of 3"
                                             % (len(idx), cap_file // 1024, mb)})
            out["lossless"] = out["lossless"] and ok
        log("cpu baseline (%s, %d threads): %.2f / %.2f MB/s" % (label, nthr, mb / tcomp,
                                                                mb / tdec))
    return out, same


def _pct(xs, q):
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def _fifo_latencies(arrivals, service):
    """Latencies of a single FIFO server (one launch per request, one at a
    time) for the given arrival times and measured service times (Lindley)."""
    free_at = 0.0
    lat = []
    for a, s in zip(arrivals, service):
        start = max(a, free_at)
        free_at = start + s
        lat.append(free_at - a)
    return lat


def serve_requests(count, seed=11):
    """The serving mix: `count` requests, 50% compress of fresh payloads and
    50% decompress of containers made beforehand; payload sizes
    min(round(256 KiB * exp(1.2 Z)), 128 MiB), content from the order-2
    Markov source (request i seeded 5000 + i); 16 KiB chunks, 2^18 slots,
    U = 128.  Returns (requests, params, containers' inputs)."""
    S, tb = 16384, 18
    params = make_params(S, 128, 100, 0, tb)
    rng = np.random.default_rng(seed)
    sizes = datagen.lognormal_sizes(count, 262144, 1.2, 1, 128 * MiB, 13)
    pool_sizes = datagen.lognormal_sizes(count // 2 + 1, 262144, 1.2, 1, 128 * MiB, 17)
    arena, offs = datagen.markov2_files(sizes, seed0=5000)
    parena, poffs = datagen.markov2_files(pool_sizes, seed0=9000)
    fresh = [arena[int(o):int(o) + int(n)] for o, n in zip(offs, sizes)]
    pool = [parena[int(o):int(o) + int(n)] for o, n in zip(poffs, pool_sizes)]
    ops = rng.random(count) < 0.5
    picks = rng.integers(len(pool), size=count)
    return fresh, pool, ops, picks, params


def serve_measure(ctx, count=240, loads=(0.25, 0.5, 0.75, 0.9, 1.0, 1.2), lanes=0):
    """Poisson arrivals of the serving mix through both serving designs: the
    serving kernel (requests' tasks admitted into one running launch) and
    static batches on lanes (each lane runs the oldest pending requests of
    one kind as one batch to the end).  First each design's saturated rate is
    measured by submitting every request at once; then the same arrival
    times, at lambda = f x (the lanes' saturated MB/s) / mean size for each
    load factor f, are replayed through both.  The plain path (one launch per
    request, one request at a time) is timed per request and its FIFO
    latencies computed for the same arrival times.  Every output is compared
    with the same request run alone."""
    import threading
    fresh, pool, ops, picks, params = serve_requests(count)
    blobs = ctx.compress_many(pool, params)
    reqs = []
    for i in range(count):
        if ops[i]:
            reqs.append(("c", fresh[i], len(fresh[i])))
        else:
            k = int(picks[i])
            reqs.append(("d", blobs[k], len(pool[k])))
    mean_bytes = float(np.mean([r[2] for r in reqs]))
    total_bytes = float(sum(r[2] for r in reqs))
    # results of each request run alone, and the plain path's service times
    alone, service = [], []
    for kind, x, n in reqs:
        t0 = time.perf_counter()
        out = ctx.compress(x, params) if kind == "c" else ctx.decompress(x)
        service.append(time.perf_counter() - t0)
        alone.append(out)
    plain_MBps = total_bytes / 1e6 / sum(service)
    log("serve: %d requests, %.1f MB, plain path %.1f MB/s" % (count, total_bytes / 1e6, plain_MBps))

    def run(arrivals):
        launches0 = ctx.serve_launches()
        ids = [None] * count
        results = [None] * count

        def poller():
            for i in range(count):
                while ids[i] is None:
                    time.sleep(0.0002)
                results[i] = ctx.poll(ids[i], wait=True)

        th = threading.Thread(target=poller)
        th.start()
        t0 = time.perf_counter()
        for i, (kind, x, n) in enumerate(reqs):
            if arrivals is not None:
                d = t0 + arrivals[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
            ids[i] = ctx.submit_compress(x, params) if kind == "c" else ctx.submit_decompress(x)
        th.join()
        makespan = time.perf_counter() - t0
        launches = ctx.serve_launches() - launches0
        same = all(results[i][0] == alone[i] for i in range(count))
        lat = [r[1]["latency_ms"] for r in results]
        return {"MBps": total_bytes / 1e6 / makespan, "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                "launches": launches, "mean_batch": float(np.mean([r[1]["batch_jobs"] for r in results])),
                "identical_to_alone": bool(same), "makespan_s": makespan}

    designs = (("kernel", False), ("lanes", True))
    sat = {}
    for name, static in designs:
        ctx.serve_start(lanes, 0, static_lanes=static)
        run(None)  # grows the lanes' buffers; not reported
        sat[name] = run(None)
        ctx.serve_stop()
        log("serve %s saturated: %.1f MB/s p50 %.1f ms p99 %.1f ms, %d launches"
            % (name, sat[name]["MBps"], sat[name]["p50_ms"], sat[name]["p99_ms"], sat[name]["launches"]))
    base = sat["lanes"]["MBps"]
    sweeps = []
    for f in loads:
        lam = f * base * 1e6 / mean_bytes
        arr = datagen.poisson_arrivals(count, lam, 19)
        plain = [1000.0 * x for x in _fifo_latencies(arr, service)]
        row = {"load": f, "lambda_per_s": lam, "offered_MBps": lam * mean_bytes / 1e6,
               "plain_p50_ms": _pct(plain, 50), "plain_p99_ms": _pct(plain, 99)}
        sweeps.append(row)
    for name, static in designs:
        ctx.serve_start(lanes, 0, static_lanes=static)
        run(None)
        for row in sweeps:
            arr = datagen.poisson_arrivals(count, row["lambda_per_s"], 19)
            row[name] = run(arr)
            log("serve %s load %.2f: %.1f MB/s p50 %.1f ms p99 %.1f ms"
                % (name, row["load"], row[name]["MBps"], row[name]["p50_ms"], row[name]["p99_ms"]))
        ctx.serve_stop()
    ok = all(sat[n]["identical_to_alone"] for n, _ in designs) and \
        all(r[n]["identical_to_alone"] for r in sweeps for n, _ in designs)
    return {"requests": count, "bytes": int(total_bytes), "mean_request_bytes": mean_bytes,
            "saturated": sat, "plain_MBps": plain_MBps,
            "plain_p50_ms": _pct([1000 * x for x in service], 50),
            "sweep": sweeps, "identical_to_alone": bool(ok),
            "load_basis": "load f = arrivals at f x the lanes' saturated MB/s"}


def serve_workload(args):
    """The serving mix (serve_measure) as its own bench line: value is the
    serving kernel's saturated MB/s."""
    count = getattr(args, "requests", 240) or 240
    ctx = Context()
    r = serve_measure(ctx, count, lanes=args.lanes)
    sat = r["saturated"]["kernel"]
    return {
        "metric": "compress/decompress MB/s on 1 GPU at identical ratio; batched vs plain; p99",
        "value": sat["MBps"], "unit": "MB/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1000 * sat["makespan_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/int32 fixed-point",
        "data": "synthetic (seeded generator, host buffers, PCIe included)",
        "config": {"workload": "serve: %d Poisson requests, 50%% compress / 50%% decompress, "
                               "lognormal sizes (median 256 KiB, sigma 1.2), 16 KiB chunks, "
                               "2^18 slots, U=128" % count},
        "serve": r, "identical_to_alone": r["identical_to_alone"],
    }


def gather_workload(args):
    """Random one-byte read-modify-write touch rate over a footprint equal to
    the stream workload's live model (resident waves x 7 order tables of
    2^table_bits bytes): the ceiling for the codec's state gathers."""
    import ctypes
    from ._native import probe_lib
    tb = getattr(args, "gather_tb", 22)
    waves = getattr(args, "gather_waves", 256)
    footprint = waves * 7 * (1 << tb)
    bytes_ = 1 << (footprint - 1).bit_length()
    ctx = Context()
    buf = ctx.alloc(bytes_)
    res = {}
    for name, blocks, thr, steps, width, wb in [("latency", 1, 64, 2000, 1, 0),
                                                ("rmw_throughput", 4096, 256, 200, 8, 1),
                                                ("read_throughput", 4096, 256, 200, 8, 0)]:
        ms = ctypes.c_double()
        clk = ctypes.c_uint64()
        rc = probe_lib().cucm_probe_gather(buf.addr, bytes_, blocks, thr, steps, width, wb,
                                     ctypes.byref(ms), ctypes.byref(clk))
        assert rc == 0
        nw = blocks * thr // 64
        res[name] = {"touches_per_s": blocks * thr * steps * width / (ms.value / 1e3),
                     "clocks_per_step_per_wave": clk.value / nw / steps, "ms": ms.value}
    return {"metric": "random one-byte touch rate", "footprint_bytes": bytes_,
            "model_footprint_bytes": footprint, "results": res}


def replay_workload(args):
    """The decoder's and the encoder's state traffic on the uneven-size batch
    (chunks of at most 128 KiB, 2^18 slots, U = 128) replayed by
    libcucm_probe.so without the model arithmetic, at several numbers of
    resident chunks and speculation widths: the memory-system ceiling each
    direction's access pattern allows."""
    import ctypes
    from ._native import probe_lib
    count = getattr(args, "files", 10000) or 10000
    sizes, arena, offs = batch_files(count)
    total = int(sizes.sum())
    ctx = Context()
    d_in = ctx.alloc(total)
    d_in.upload(arena)
    chunkings = {}

    def chunk_list(kib):
        # the batch's files cut into chunks of at most kib KiB
        if kib not in chunkings:
            src, lens = [], []
            for o, n in zip(offs, sizes):
                o, n = int(o), int(n)
                for c in range(0, n, kib * 1024):
                    src.append(d_in.addr + o + c)
                    lens.append(min(kib * 1024, n - c))
            chunkings[kib] = ((ctypes.c_uint64 * len(src))(*src),
                              (ctypes.c_uint32 * len(lens))(*lens), len(src))
        return chunkings[kib]

    out = []
    # (speculation bits, resident waves, flags[, chunk KiB (default 128)]); flags: 1 zero
    # the slot per chunk, 2 no write-backs, 4 pack each round's gathers into one
    # instruction, 8 re-gather before the write-backs, 16 write right after each round's
    # gathers, 32 agent-scope release + acquire fences after every chunk; speculation
    # bits 0: the encoder's stream (every gather of a unit at once; flags 1, 2, 4, and 8: no
    # wait for a unit's write-backs before the next unit's gathers)
    grid = [tuple(int(v) for v in e.split(":")) for e in args.replay_grid.split(",")] \
        if getattr(args, "replay_grid", "") else [
            (1, 1536, 1), (1, 3072, 1), (2, 1280, 1), (2, 1536, 1), (2, 2048, 1), (2, 3072, 1),
            (4, 1280, 1), (2, 1536, 0), (2, 1536, 3), (1, 3072, 3), (2, 1536, 5), (2, 1536, 9),
            (2, 1536, 17), (0, 768, 1), (0, 1536, 1), (0, 3072, 1), (0, 1536, 3),
        ]
    for entry in grid:
        D, waves, flags = entry[:3]
        kib = entry[3] if len(entry) > 3 else 128
        srca, lena, nsrc = chunk_list(kib)
        ms = ctypes.c_double()
        if D == 0:
            rc = probe_lib().cucm_probe_replay_encode(srca, lena, nsrc, 18, 128, waves, flags,
                                                      ctypes.byref(ms))
        else:
            rc = probe_lib().cucm_probe_replay_decode(srca, lena, nsrc, 18, D, 128, waves,
                                                      flags, ctypes.byref(ms))
        if rc:
            raise RuntimeError("replay probe failed: %d" % rc)
        out.append({"stream": "encoder" if D == 0 else "decoder",
                    "spec_bits": D, "waves": waves, "zero_slot": bool(flags & 1),
                    "write_back": not (flags & 18), "packed_lanes": bool(flags & 4),
                    "regather_before_write": bool(D and flags & 8),
                    "no_drain_before_next_unit": bool(not D and flags & 8),
                    "write_per_round": bool(flags & 16),
                    "fences_per_chunk": bool(flags & 32), "chunk_kib": kib, "chunks": nsrc,
                    "ms": ms.value, "MBps": total / 1e3 / ms.value})
        log("replay D=%d waves=%d flags=%d chunks=%d KiB: %.0f ms, %.1f MB/s"
            % (D, waves, flags, kib, ms.value, out[-1]["MBps"]))
    return {"metric": "decoder gather-stream replay (no model arithmetic)", "bytes": total,
            "table_bits": 18, "unit_bits": 128, "results": out}


def copy_workload(args):
    """A plain 16-byte streaming copy of --mib MiB (HBM counter calibration)."""
    import ctypes
    from ._native import probe_lib
    n = (args.mib or 256) * MiB
    ctx = Context()
    a = ctx.alloc(n)
    b = ctx.alloc(n)
    out = []
    for _ in range(3):
        ms = ctypes.c_double()
        assert probe_lib().cucm_probe_copy(b.addr, a.addr, n, ctypes.byref(ms)) == 0
        out.append(ms.value)
    return {"metric": "copy kernel", "bytes": n, "ms": out, "GBps": 2 * n / (min(out) / 1e3) / 1e9}
