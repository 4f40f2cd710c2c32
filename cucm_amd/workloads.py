"""Benchmark workloads beyond the default stream (used by bench.py).

batch_workload: 10,000 independent files, lognormal sizes (median 64 KiB,
sigma 1.5, clipped to [512 B, 64 MiB]), file i generated by the order-2
Markov source seeded 1000 + i; 2^18 slots per order table and chunks of at
most 256 KiB, or the settings the autotuner picks on a sample (--autotune).  All files are
encoded, then decoded, in one persistent launch each (device-resident
buffers); the plain path issues one launch per file, as the reference does
per batch, and is timed on a seeded random sample of the same files because
each single-file launch runs alone on one wave.

serve_workload: Poisson arrivals of mixed compress/decompress requests
through the continuous-batching serving loop (cucm_serve_*), p50/p99 per
load level beside the one-request-at-a-time plain path.
"""
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


def batch_workload(args):
    count = getattr(args, "files", 10000) or 10000
    s_max = getattr(args, "chunk_kib", 256) * 1024
    tb = 18
    spec = args.spec
    t0 = time.time()
    sizes = datagen.lognormal_sizes(count, 65536, 1.5, 512, 64 * MiB, 3)
    arena, offs = datagen.markov2_files(sizes, seed0=1000)
    total = int(sizes.sum())
    log("generated %d files, %.1f MiB in %.1fs" % (count, total / MiB, time.time() - t0))
    ctx = Context()
    tuned = None
    if getattr(args, "autotune", False):
        from .autotune import autotune_batch
        tuned = autotune_batch(ctx, arena, offs, sizes, log=log)
        s_max, tb, spec = tuned["chunk_max"], tuned["table_bits"], tuned["spec_bits"]
        ctx.set_waves_per_block(tuned["waves_per_block"])
        ctx.set_team(2 if tuned.get("team") else 1)
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
    # batched: one persistent launch each way
    for _ in range(max(args.warmup, 0)):
        cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
        ctx.decompress_many_device(cs, cl, outs, lens, spec)
    ctx.synchronize()
    a = time.perf_counter()
    cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
    b = time.perf_counter()
    sc = ctx.stats()
    dl, dst = ctx.decompress_many_device(cs, cl, outs, lens, spec)
    c = time.perf_counter()
    sd = ctx.stats()
    bad = [(i, a, b) for i, (a, b) in enumerate(zip(st, dst)) if a or b]
    if bad:
        raise RuntimeError("batch: %d files failed, first (file, compress, decompress) status: %s"
                           % (len(bad), bad[:5]))
    back = np.frombuffer(d_o.download(total), dtype=np.uint8)
    lossless = bool(np.array_equal(back, arena))
    comp_total = int(sum(cl))
    # plain: one launch per file on a seeded sample of the same files
    rng = np.random.default_rng(7)
    sample = rng.choice(count, size=min(args.plain_files, count), replace=False)
    sin = [ins[i] for i in sample]
    slen = [lens[i] for i in sample]
    sp = [params[i] for i in sample]
    scs = [cs[i] for i in sample]
    scap = [caps[i] for i in sample]
    souts = [outs[i] for i in sample]
    d = time.perf_counter()
    pl, _ = ctx.compress_many_device(sin, slen, sp, scs, scap, plain=True)
    e = time.perf_counter()
    pdl, _ = ctx.decompress_many_device(scs, pl, souts, slen, spec, plain=True)
    f = time.perf_counter()
    plain_identical = all(pl[k] == cl[i] for k, i in enumerate(sample))
    sample_bytes = sum(slen)
    res = {
        "metric": "compress/decompress MB/s on 1 GPU at identical ratio; batched vs plain; p99",
        "value": total / 1e6 / ((b - a) + (c - b)),
        "unit": "MB/s", "n_gpus": 1, "steps": 1, "warmup": args.warmup,
        "ms_per_step": 1000 * ((b - a) + (c - b)), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/int32 fixed-point",
        "data": "synthetic (seeded generator, resident in HBM)",
        "config": {"workload": "batch: %d lognormal files (median 64 KiB, sigma 1.5), 2^%d slots, "
                               "chunk = min(file, %d KiB), U=128%s" % (
                                   count, tb, s_max // 1024, ", autotuned" if tuned else ""),
                   "files": count, "bytes": total, "table_bits": tb, "chunk_max": s_max,
                   "spec_bits": spec,
                   "waves_per_block": tuned["waves_per_block"] if tuned else 0,
                   "team": bool(tuned.get("team")) if tuned else False},
        "batched_compress_MBps": total / 1e6 / (b - a),
        "batched_decompress_MBps": total / 1e6 / (c - b),
        "batched_compress_kernel_ms": sc["codec_ms"],
        "batched_decompress_kernel_ms": sd["codec_ms"],
        "plain_sample_files": len(sample), "plain_sample_bytes": sample_bytes,
        "plain_compress_MBps": sample_bytes / 1e6 / (e - d),
        "plain_decompress_MBps": sample_bytes / 1e6 / (f - e),
        "plain_identical_to_batched": bool(plain_identical),
        "ratio": total / comp_total, "lossless": lossless, "tasks": sc["tasks"],
        "waves": sc["waves"],
    }
    if tuned:
        res["autotune"] = {k: tuned[k] for k in ("chunk_max", "table_bits", "spec_bits",
                                                 "waves_per_block", "team", "sample_files",
                                                 "sample_bytes", "ratio_tol", "seconds")}
        res["autotune"]["trials"] = [{k: t.get(k) for k in ("chunk_max", "table_bits",
                                                            "spec_bits", "waves_per_block",
                                                            "team", "ratio", "score")}
                                     for t in tuned["stream_trials"] + tuned["launch_trials"]]
    res["batched_vs_plain_compress"] = res["batched_compress_MBps"] / res["plain_compress_MBps"]
    res["batched_vs_plain_decompress"] = res["batched_decompress_MBps"] / res["plain_decompress_MBps"]
    return res


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


def serve_workload(args):
    """Poisson arrivals of mixed requests through the serving loop.

    Requests: 50% compress of fresh payloads, 50% decompress of containers
    made in a warm-up; payload sizes min(round(256 KiB * exp(1.2 Z)), 128 MiB),
    content from the order-2 Markov source (request i seeded 5000 + i); the
    small-config model (16 KiB chunks, 2^18 slots, U = 128).  First the
    saturated rate is measured by submitting every request at once; then
    arrivals are replayed at lambda = f x saturated MB/s / mean size for each
    load factor f.  The plain path (one launch per request, one request at a
    time) is timed per request and its FIFO latencies computed for the same
    arrival times.  Every output is compared with the same request run alone.
    """
    import threading
    count = getattr(args, "requests", 240) or 240
    S, tb = 16384, 18
    params = make_params(S, 128, 100, 0, tb)
    rng = np.random.default_rng(11)
    sizes = datagen.lognormal_sizes(count, 262144, 1.2, 1, 128 * MiB, 13)
    pool_sizes = datagen.lognormal_sizes(count // 2 + 1, 262144, 1.2, 1, 128 * MiB, 17)
    arena, offs = datagen.markov2_files(sizes, seed0=5000)
    parena, poffs = datagen.markov2_files(pool_sizes, seed0=9000)
    fresh = [arena[int(o):int(o) + int(n)] for o, n in zip(offs, sizes)]
    pool = [parena[int(o):int(o) + int(n)] for o, n in zip(poffs, pool_sizes)]
    ctx = Context()
    # warm-up: containers that the decompress requests replay
    blobs = ctx.compress_many(pool, params)
    ops = rng.random(count) < 0.5
    reqs = []
    for i in range(count):
        if ops[i]:
            reqs.append(("c", fresh[i], len(fresh[i])))
        else:
            k = int(rng.integers(len(pool)))
            reqs.append(("d", (blobs[k], k), len(pool[k])))
    mean_bytes = float(np.mean([r[2] for r in reqs]))
    total_bytes = float(sum(r[2] for r in reqs))
    # results of each request run alone, and the plain path's service times
    alone, service = [], []
    for kind, x, n in reqs:
        t0 = time.perf_counter()
        if kind == "c":
            out = ctx.compress(x, params)
        else:
            out = ctx.decompress(x[0])
        service.append(time.perf_counter() - t0)
        alone.append(out)
    plain_MBps = total_bytes / 1e6 / sum(service)
    log("serve: %d requests, %.1f MB, plain path %.1f MB/s" % (count, total_bytes / 1e6, plain_MBps))

    # one long-running server for every load level; its lanes' buffers are
    # grown by a first saturated pass that is not reported
    ctx.serve_start(args.lanes, 0)

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
            ids[i] = ctx.submit_compress(x, params) if kind == "c" else ctx.submit_decompress(x[0])
        th.join()
        makespan = time.perf_counter() - t0
        launches = ctx.serve_launches() - launches0
        same = all(results[i][0] == alone[i] for i in range(count))
        lat = [r[1]["latency_ms"] for r in results]
        return {"MBps": total_bytes / 1e6 / makespan, "p50_ms": _pct(lat, 50), "p99_ms": _pct(lat, 99),
                "launches": launches, "mean_batch": float(np.mean([r[1]["batch_jobs"] for r in results])),
                "identical_to_alone": bool(same), "makespan_s": makespan}

    run(None)
    sat = run(None)
    log("serve saturated: %.1f MB/s, %d launches" % (sat["MBps"], sat["launches"]))
    sweeps = []
    for f in [0.25, 0.5, 0.75, 0.9, 1.0, 1.2]:
        lam = f * sat["MBps"] * 1e6 / mean_bytes
        arr = datagen.poisson_arrivals(count, lam, 19)
        r = run(arr)
        plain = [1000.0 * x for x in _fifo_latencies(arr, service)]
        r.update({"load": f, "lambda_per_s": lam, "offered_MBps": lam * mean_bytes / 1e6,
                  "plain_p50_ms": _pct(plain, 50), "plain_p99_ms": _pct(plain, 99)})
        log("serve load %.2f: %.1f MB/s p50 %.1f ms p99 %.1f ms (plain p50 %.1f p99 %.1f)"
            % (f, r["MBps"], r["p50_ms"], r["p99_ms"], r["plain_p50_ms"], r["plain_p99_ms"]))
        sweeps.append(r)
    ctx.serve_stop()
    ok = sat["identical_to_alone"] and all(r["identical_to_alone"] for r in sweeps)
    return {
        "metric": "compress/decompress MB/s on 1 GPU at identical ratio; batched vs plain; p99",
        "value": sat["MBps"], "unit": "MB/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": 1000 * sat["makespan_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u8/int32 fixed-point",
        "data": "synthetic (seeded generator, host buffers, PCIe included)",
        "config": {"workload": "serve: %d Poisson requests, 50%% compress / 50%% decompress, "
                               "lognormal sizes (median 256 KiB, sigma 1.2), 16 KiB chunks, "
                               "2^18 slots, U=128" % count, "lanes": args.lanes or 3},
        "saturated": sat, "plain_MBps": plain_MBps, "plain_p50_ms": _pct([1000 * s for s in service], 50),
        "sweep": sweeps, "identical_to_alone": bool(ok),
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
