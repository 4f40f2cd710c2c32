#!/usr/bin/env python3
"""Benchmark of the MI355X CuCM codec (one JSON line on stdout).

Default workload (the metric's batched configuration): 10,000 independent
files with lognormal sizes (median 64 KiB, sigma 1.5, clipped to [512 B,
64 MiB]), file i from the order-2 Markov source over 64 symbols seeded
1000 + i; per file, chunks of at most 128 KiB and 2^18 slots per order table
(the autotuner's ratio-constrained choice, kept in cucm_amd/batch_settings.json
with the launch settings measured on the full batch), U = 128, P = 0.

One step = compress every file (device-resident inputs and containers) in one
persistent launch, then decompress every file in another, through the C ABI.
`value` is input MB (1e6 B) per second of step time; the compress and
decompress rates, the plain path (one launch per file, as the reference
launches per batch, on a seeded 40-file sample) and the ratio are reported
beside it.  After the timed region the round trip must be lossless; the CPU
baseline (the oracle, files as tasks on every host thread and on one, warm-up
then the median of 3) is timed on size-stratified samples of the same files,
whose containers must equal the GPU's.

Other workloads: --workload stream (256 MiB mixed stream, 1 MiB chunks, 2^22
slots), structured (1 GiB structured binary, 256 KiB chunks, 2^22 slots),
small (1 MiB order-2 Markov, 16 KiB chunks, 2^18 slots), serve (Poisson
request mix, p50/p99), gather (random-touch ceiling), copy (counter
calibration).  --frozen runs a single-stream workload with the frozen model.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MiB = 1 << 20
METRIC = "compress/decompress MB/s on 1 GPU at identical ratio; batched vs plain; p99"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per input byte: 8 bits x 8 orders one-byte state
# gathers + the same number of one-byte state write-backs, 1 B of input and
# ~1/ratio B of output (order-0 and statep/mixer live in LDS).
ALG_BYTES_PER_INPUT_BYTE = 128.0
# The frozen model reads the same 64 states per input byte and writes none.
ALG_BYTES_PER_INPUT_BYTE_FROZEN = 64.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Single-stream workloads: (generator, default MiB, chunk size, table bits, description).
STREAMS = {
    "stream": ("mixed", 256, 1 * MiB, 22,
               "mixed stream %d MiB (60%% order-3 text, 25%% LZ copies, 15%% random 4 KiB "
               "blocks; seed 2), 1 MiB chunks, 2^22 slots/order, U=128, P=0"),
    "structured": ("structured", 1024, 256 * 1024, 22,
                   "structured binary %d MiB (int32 walks, float32 fields, 16-byte records; "
                   "seed 4), 256 KiB chunks, 2^22 slots/order, U=128, P=0"),
    "small": ("markov2", 1, 16 * 1024, 18,
              "order-2 Markov %d MiB (64 symbols, Zipf 1.1; seed 1), 16 KiB chunks, "
              "2^18 slots/order, U=128, P=0"),
}


def stream_workload(args):
    import cucm_amd
    from cucm_amd import datagen

    gen, mib_default, S, tb, desc = STREAMS[args.workload]
    mib = args.mib or mib_default
    n = mib * MiB
    U, batch = 128, 100
    P = args.prelearn_kib * 1024 if args.prelearn_kib >= 0 else (MiB if args.frozen else 0)
    if P >= n:
        raise SystemExit("prelearn must be smaller than the input")
    t0 = time.time()
    seed = {"mixed": 2, "structured": 4, "markov2": 1}[gen]
    data = getattr(datagen, gen)(n, seed)
    log("generated %d MiB %s input in %.1fs" % (mib, gen, time.time() - t0))
    ctx = cucm_amd.Context()
    params = cucm_amd.make_params(S, U, batch, P, tb, frozen=args.frozen)
    C = (n - P + S - 1) // S
    cap = n + n // 2 + 4096 * (C + 1)
    d_in = ctx.alloc(n)
    d_in.upload(data)
    d_comp = ctx.alloc(cap)
    d_out = ctx.alloc(n)
    ctx.synchronize()

    def compress():
        lens, st = ctx.compress_many_device([d_in.addr], [n], params, [d_comp.addr], [cap])
        if st[0] != 0:
            raise RuntimeError("compress status %d" % st[0])
        return lens[0], ctx.stats()

    def decompress(clen):
        lens, st = ctx.decompress_many_device([d_comp.addr], [clen], [d_out.addr], [n], args.spec)
        if st[0] != 0 or lens[0] != n:
            raise RuntimeError("decompress status %d len %d" % (st[0], lens[0]))
        return ctx.stats()

    for i in range(args.warmup):
        clen, _ = compress()
        decompress(clen)
        log("warmup %d done" % i)
    ctx.synchronize()
    tc, td, kc, kd = [], [], [], []
    t_all0 = time.perf_counter()
    for i in range(args.steps):
        a = time.perf_counter()
        clen, sc = compress()
        b = time.perf_counter()
        sd = decompress(clen)
        c = time.perf_counter()
        tc.append(b - a)
        td.append(c - b)
        kc.append(sc["codec_ms"])
        kd.append(sd["codec_ms"])
        log("step %d: compress %.3fs (kernel %.1f ms)  decompress %.3fs (kernel %.1f ms)"
            % (i, b - a, sc["codec_ms"], c - b, sd["codec_ms"]))
    ctx.synchronize()
    t_all = time.perf_counter() - t_all0
    # checks outside the timed region
    back = np.frombuffer(d_out.download(n), dtype=np.uint8)
    lossless = bool(np.array_equal(back, data))
    container = d_comp.download(clen)
    ratio = n / clen
    res = {
        "compress_MBps": n / 1e6 / float(np.median(tc)),
        "decompress_MBps": n / 1e6 / float(np.median(td)),
        "compress_kernel_ms": float(np.median(kc)),
        "decompress_kernel_ms": float(np.median(kd)),
        "ratio": ratio,
        "lossless": lossless,
        "container_bytes": clen,
        "waves": sc["waves"],
        "compress_teams": sc["teams"],
        "decompress_teams": sd["teams"],
        "decompress_spec_bits": sd["spec_bits"],
        "model_bytes": sc["model_bytes"],
    }
    ms_per_step = 1000.0 * t_all / args.steps
    value = n / 1e6 / (t_all / args.steps)
    # dominant kernel: the slower of the two launches (a team kernel when the
    # launch ran as teams of four waves, else the one-wave-per-chunk kernel)
    dom = "decode" if res["decompress_kernel_ms"] >= res["compress_kernel_ms"] else "encode"
    teams = res["decompress_teams"] if dom == "decode" else res["compress_teams"]
    if args.frozen:
        kname = "frozen_kernel<%s>" % dom
    else:
        kname = ("%s_team_kernel" % dom) if teams else ("codec_kernel<%s>" % dom)
    kms = max(res["compress_kernel_ms"], res["decompress_kernel_ms"])
    alg = ALG_BYTES_PER_INPUT_BYTE_FROZEN if args.frozen else ALG_BYTES_PER_INPUT_BYTE
    achieved = alg * (n - P) / (kms / 1000.0) / 1e9
    roofline = {"bound": "hbm", "kernel": kname, "achieved": achieved,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                "traffic": measured_traffic(kname, n - P)}
    roofline.update(physical_rates(kname, n - P, kms / 1000.0))
    res["compress_kernel_MBps"] = (n - P) / 1e6 / (res["compress_kernel_ms"] / 1000.0)
    res["decompress_kernel_MBps"] = (n - P) / 1e6 / (res["decompress_kernel_ms"] / 1000.0)
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(data, container, S, U, tb, P, args)
    line = {
        "metric": METRIC, "value": value, "unit": "MB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_per_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u8/int32 fixed-point",
        "data": "synthetic (seeded generator, resident in HBM)",
        "config": {"workload": (desc % mib) if not (args.frozen or P) else
                   (desc % mib).replace("P=0", "P=%d KiB%s" % (P // 1024,
                                                               ", frozen model" if args.frozen
                                                               else "")),
                   "chunks": C, "chunk_size": S, "unit_bits": U, "table_bits": tb,
                   "batch_chunks": batch, "spec_bits": args.spec or "automatic", "prelearn_bytes": P,
                   "frozen": bool(args.frozen)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    line.update(res)
    line["oracle_sample_identical"] = parity
    return line


def measured_traffic(kernel, n):
    """HBM bytes per launch of the dominant kernel: the fetched + written bytes
    per input byte from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    (calibrated against a plain copy kernel), times the workload size."""
    path = os.path.join(ROOT, "profiles", "codec_kernel_traffic.json")
    if not os.path.exists(path):
        return None
    t = json.load(open(path)).get(kernel)
    if t is None:
        return None
    return (t["fetch_bytes_per_input_byte"] + t["write_bytes_per_input_byte"]) * n


def physical_rates(kernel, n, seconds):
    """The dominant kernel's physical HBM rate (measured bytes per input byte
    x input bytes / kernel time), its L2 hit rate on the state gathers, and
    its 128-byte line fetches per second as a fraction of the measured
    random one-byte read ceiling on the same footprint."""
    path = os.path.join(ROOT, "profiles", "codec_kernel_traffic.json")
    if not os.path.exists(path):
        return {}
    d = json.load(open(path))
    t, ceil = d.get(kernel), d.get("gather_ceiling")
    if t is None or ceil is None:
        return {}
    fetch = t["fetch_bytes_per_input_byte"] * n
    write = t["write_bytes_per_input_byte"] * n
    lines_per_s = fetch / ceil["line_bytes"] / seconds
    return {"hbm_GBps": (fetch + write) / seconds / 1e9,
            "hbm_frac_of_peak": (fetch + write) / seconds / 1e9 / HBM_PEAK_GBS,
            "l2_hit_rate": t["l2_hit_rate"],
            "line_fetches_per_s": lines_per_s,
            "gather_ceiling_frac": lines_per_s / ceil["read_touches_per_s"]}


def chunk_payloads(container, nchunks_max):
    """Yields (chunk index, payload) of the first chunks of a container."""
    import struct
    P = struct.unpack(">I", container[16:20])[0]
    Pc = struct.unpack(">I", container[12:16])[0]
    flag = container[20] & 1  # bit 0: coded prelearn segment (bit 1: frozen model)
    N = struct.unpack(">I", container[0:4])[0]
    batch = struct.unpack(">H", container[4:6])[0]
    S = struct.unpack(">I", container[8:12])[0]
    pos = 21 + (Pc if flag else P)
    C = (N - P + S - 1) // S
    out = []
    c0 = 0
    while c0 < C and len(out) < nchunks_max:
        Cb = min(batch, C - c0)
        blen = struct.unpack("<I", container[pos:pos + 4])[0]
        blob = container[pos + 4:pos + 4 + blen]
        prev = 4 * Cb - 1
        for c in range(Cb):
            e = struct.unpack(">I", blob[4 * c:4 * c + 4])[0]
            out.append((c0 + c, blob[prev + 1:e + 1]))
            prev = e
            if len(out) >= nchunks_max:
                break
        pos += 4 + blen
        c0 += Cb
    return out


def cpu_baseline(data, container, S, U, tb, P, args):
    """Times the oracle (a port of the reference model) on a bounded sample of
    the same workload: `sample` whole chunks after the P prelearn bytes, one
    chunk per core (OpenMP), from the same initial model."""
    from oracle import oracle as O

    threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
    k = min(args.cpu_chunks or threads, (len(data) - P + S - 1) // S)
    sample = data[P: P + k * S].tobytes()
    model = O.Model(tb)
    if P:
        O.prelearn_train(data[:P].tobytes(), model)
    t0 = time.perf_counter()
    if args.frozen:
        blob = O.encode_batch_frozen(sample, S, model, threads=threads)
    else:
        blob = O.encode_batch(sample, S, U, model, tail_pad=True, threads=threads)
    t1 = time.perf_counter()
    if args.frozen:
        back = O.decode_batch_frozen(blob, len(sample), S, model, threads=threads)
    else:
        back = O.decode_batch(blob, len(sample), S, U, model, threads=threads)
    t2 = time.perf_counter()
    ok = back == sample
    # compare the oracle's payloads with the GPU container's first k chunks
    import struct
    gpu = dict(chunk_payloads(container, k))
    prev = 4 * k - 1
    same = True
    for c in range(k):
        e = struct.unpack(">I", blob[4 * c:4 * c + 4])[0]
        if gpu.get(c) != blob[prev + 1:e + 1]:
            same = False
        prev = e
    cpu = {"value": len(sample) / 1e6 / (t1 - t0), "unit": "MB/s (compress)",
           "decompress_MBps": len(sample) / 1e6 / (t2 - t1), "cores": threads, "kind": "port",
           "sample": "%d chunks x %d KiB of the same input (its first %d KiB after the "
                     "prelearn bytes), oracle encode_batch/decode_batch%s, OpenMP one chunk per "
                     "core" % (k, S // 1024, len(sample) // 1024,
                               "_frozen" if args.frozen else ""),
           "lossless": ok}
    return cpu, same


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batch",
                    choices=["stream", "structured", "small", "batch", "serve", "gather", "copy",
                             "replay"])
    ap.add_argument("--mib", type=int, default=0, help="input size (0 = the workload's own)")
    ap.add_argument("--spec", type=int, default=0,
                    help="decoder bits per speculative round (0: automatic, chosen with the launch)")
    ap.add_argument("--frozen", action="store_true",
                    help="single-stream workloads: code with the frozen model")
    ap.add_argument("--prelearn-kib", type=int, default=-1,
                    help="single-stream workloads: prelearn bytes (default 1024 KiB with "
                         "--frozen, else 0)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-chunks", type=int, default=0, help="stream workloads: CPU sample chunks")
    ap.add_argument("--files", type=int, default=10000, help="batch workload: number of files")
    ap.add_argument("--plain-files", type=int, default=40,
                    help="batch workload: files in the one-launch-per-file sample (0: skip)")
    ap.add_argument("--autotune", action="store_true",
                    help="batch workload: tune chunk size, table bits and each direction's launch "
                         "settings on a sample first instead of using cucm_amd/batch_settings.json")
    ap.add_argument("--save-settings", action="store_true",
                    help="with --autotune: write the result to cucm_amd/batch_settings.json")
    ap.add_argument("--cpu-mib", type=int, default=12,
                    help="batch workload: MiB of files in the all-thread CPU baseline sample")
    ap.add_argument("--requests", type=int, default=240, help="serve workload: requests per run")
    ap.add_argument("--serve-requests", type=int, default=240,
                    help="batch workload: requests of the serving mix measured beside it (0 = none)")
    ap.add_argument("--lanes", type=int, default=0, help="serve workload: concurrent lanes (0 = 3)")
    ap.add_argument("--phases", action="store_true",
                    help="batch workload: one more decode in the profiling instantiation, its "
                         "per-phase clocks per decoded bit in the line")
    ap.add_argument("--replay-grid", default="",
                    help="replay workload: D:waves:flags[:chunk KiB] entries, comma separated "
                         "(default: the decoder's residencies and the variants of "
                         "cucm_probe_replay_decode, 128 KiB chunks)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the codec shards by chunk with no exchange on the "
              "hot loop, and this build runs on a single MI355X (use --gpus 1).", file=sys.stderr)
        sys.exit(2)
    if args.workload in STREAMS:
        line = stream_workload(args)
    elif args.workload == "batch":
        from cucm_amd.workloads import batch_workload
        line = batch_workload(args)
    elif args.workload == "gather":
        from cucm_amd.workloads import gather_workload
        line = gather_workload(args)
    elif args.workload == "replay":
        from cucm_amd.workloads import replay_workload
        line = replay_workload(args)
    elif args.workload == "copy":
        from cucm_amd.workloads import copy_workload
        line = copy_workload(args)
    else:
        from cucm_amd.workloads import serve_workload
        line = serve_workload(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
