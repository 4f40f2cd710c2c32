"""Benchmark workloads beyond the default stream (used by bench.py).

batch_workload: 10,000 independent files, lognormal sizes (median 64 KiB,
sigma 1.5, clipped to [512 B, 64 MiB]), file i generated by the order-2
Markov source seeded 1000 + i; 2^18 slots per order table.  All files are
encoded, then decoded, in one persistent launch each (device-resident
buffers); the plain path issues one launch per file, as the reference does
per batch, and is timed on a seeded random sample of the same files because
each single-file launch runs alone on one wave.

serve_workload: Poisson arrivals of mixed compress/decompress requests
through the continuous-batching serving loop (cucm_serve_*).
"""
import os
import sys
import time

import numpy as np

from . import datagen
from .codec import Context, make_params

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
    t0 = time.time()
    sizes = datagen.lognormal_sizes(count, 65536, 1.5, 512, 64 * MiB, 3)
    arena, offs = datagen.markov2_files(sizes, seed0=1000)
    total = int(sizes.sum())
    log("generated %d files, %.1f MiB in %.1fs" % (count, total / MiB, time.time() - t0))
    ctx = Context()
    params = [make_params(pick_chunk_size(int(n), s_max), 128, 100, 0, tb) for n in sizes]
    d_in = ctx.alloc(total)
    d_in.upload(arena)
    caps = [int(n) + int(n) // 2 + 4096 * (1 + int(n) // p.chunk_size) for n, p in zip(sizes, params)]
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
        ctx.decompress_many_device(cs, cl, outs, lens, args.spec)
    ctx.synchronize()
    a = time.perf_counter()
    cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
    b = time.perf_counter()
    sc = ctx.stats()
    dl, dst = ctx.decompress_many_device(cs, cl, outs, lens, args.spec)
    c = time.perf_counter()
    sd = ctx.stats()
    assert all(s == 0 for s in st) and all(s == 0 for s in dst)
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
    pdl, _ = ctx.decompress_many_device(scs, pl, souts, slen, args.spec, plain=True)
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
        "config": {"workload": "batch: %d lognormal files (median 64 KiB, sigma 1.5), 2^18 slots, "
                               "chunk = min(file, %d KiB), U=128" % (count, s_max // 1024),
                   "files": count, "bytes": total},
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
    res["batched_vs_plain_compress"] = res["batched_compress_MBps"] / res["plain_compress_MBps"]
    res["batched_vs_plain_decompress"] = res["batched_decompress_MBps"] / res["plain_decompress_MBps"]
    return res


def serve_workload(args):
    raise SystemExit("the serving workload needs the serving loop (cucm_serve_start)")


def gather_workload(args):
    """Random one-byte read-modify-write touch rate over a footprint equal to
    the stream workload's live model (resident waves x 7 order tables of
    2^table_bits bytes): the ceiling for the codec's state gathers."""
    import ctypes
    from ._native import lib
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
        rc = lib().cucm_gather_probe(buf.addr, bytes_, blocks, thr, steps, width, wb,
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
    from ._native import lib
    n = args.mib * MiB
    ctx = Context()
    a = ctx.alloc(n)
    b = ctx.alloc(n)
    out = []
    for _ in range(3):
        ms = ctypes.c_double()
        assert lib().cucm_copy_probe(b.addr, a.addr, n, ctypes.byref(ms)) == 0
        out.append(ms.value)
    return {"metric": "copy kernel", "bytes": n, "ms": out, "GBps": 2 * n / (min(out) / 1e3) / 1e9}
