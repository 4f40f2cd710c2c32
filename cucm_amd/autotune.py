"""Autotuner for batched compression of many files of uneven size.

It measures candidate settings on a size-stratified sample of the files
(device-resident, one persistent launch per direction, kernel time from HIP
events) and picks:

  - stream settings: the largest chunk size (a file up to it is one chunk;
    longer files are cut into chunks of that size) and the table bits.  These
    change the container, so a candidate is eligible only if its compression
    ratio on the sample is at least (1 - ratio_tol) x the ratio of the
    reference setting (2^18 slots, 256 KiB chunks);
  - launch settings that never change the container: the decoder's
    speculation width (bits decided per round), the waves per workgroup
    (how many waves share one LDS copy of the lookup tables) and team
    execution (one chunk per workgroup of four waves instead of one wave per
    chunk).

The score is sample bytes per second of compress + decompress kernel time.
`choose` is the pure selection rule; `autotune_batch` runs the measurements.
"""
import json
import os
import time

import numpy as np

from .codec import make_params, output_cap

KiB = 1024
REFERENCE = {"chunk_max": 256 * KiB, "table_bits": 18}
STREAM_GRID = [(c, t) for c in (128 * KiB, 256 * KiB, 1024 * KiB) for t in (16, 18)]
SPEC_GRID = (2, 4, 8)
WPB_GRID = (0, 2, 4, 8)
TEAM_SPEC_GRID = (2, 4)  # speculation widths the team decoder runs


def stratified_sample(sizes, budget, seed=0):
    """Indices of files spread over the size distribution (every size
    quantile represented) whose total stays within `budget` bytes."""
    sizes = np.asarray(sizes, dtype=np.int64)
    order = np.argsort(sizes, kind="stable")
    rng = np.random.default_rng(seed)
    mean = max(float(sizes.mean()), 1.0)
    k = int(min(len(sizes), max(1, budget // mean)))
    strata = np.array_split(order, k)
    pick, total = [], 0
    for s in rng.permutation(len(strata)):
        if not len(strata[s]):
            continue
        i = int(rng.choice(strata[s]))
        if total + sizes[i] > budget and pick:
            continue
        pick.append(i)
        total += int(sizes[i])
    return sorted(pick)


def choose(trials, ratio_tol):
    """trials: dicts with chunk_max, table_bits, ratio, score (bytes/s).
    Returns the best-scoring trial whose ratio is within ratio_tol of the
    reference setting's (the reference trial itself is always eligible)."""
    ref = [t for t in trials if t["chunk_max"] == REFERENCE["chunk_max"]
           and t["table_bits"] == REFERENCE["table_bits"]]
    if not ref:
        raise ValueError("the reference setting was not measured")
    floor = ref[0]["ratio"] * (1.0 - ratio_tol)
    ok = [t for t in trials if t["ratio"] >= floor]
    return max(ok, key=lambda t: t["score"])


def file_params(n, chunk_max, table_bits, unit_bits=128, batch_chunks=100):
    """Per-file parameters: one chunk up to chunk_max bytes."""
    return make_params(int(max(1, min(n, chunk_max))), unit_bits, batch_chunks, 0, table_bits)


def _measure(ctx, d_in, ins, lens, d_c, d_o, chunk_max, table_bits, spec, wpb, coffs, team=False):
    params = [file_params(n, chunk_max, table_bits) for n in lens]
    caps = [output_cap(n, p) for n, p in zip(lens, params)]
    if any(c > int(coffs[i + 1] - coffs[i]) for i, c in enumerate(caps)):
        raise RuntimeError("sample output arena too small")
    cs = [d_c.addr + int(o) for o in coffs[:-1]]
    outs = [d_o.addr + (a - d_in.addr) for a in ins]
    ctx.set_waves_per_block(wpb)
    ctx.set_team(2 if team else 1)
    cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
    kc = ctx.stats()["codec_ms"]
    dl, dst = ctx.decompress_many_device(cs, cl, outs, lens, spec)
    kd = ctx.stats()["codec_ms"]
    if any(st) or any(dst):
        raise RuntimeError("autotune trial failed")
    total = sum(lens)
    return {"chunk_max": chunk_max, "table_bits": table_bits, "spec_bits": spec,
            "waves_per_block": wpb, "team": team, "ratio": total / float(sum(cl)),
            "compress_ms": kc, "decompress_ms": kd,
            "score": total / ((kc + kd) / 1e3)}


def autotune_batch(ctx, arena, offs, sizes, sample_bytes=96 << 20, ratio_tol=0.01, cache=None,
                   log=None):
    """Tunes the batch settings on a sample of the files (arena/offs/sizes as
    datagen.markov2_files returns them).  Returns a dict with chunk_max,
    table_bits, spec_bits, waves_per_block, team and the trials."""
    key = "%d files, %d bytes, tol %.4f" % (len(sizes), int(np.sum(sizes)), ratio_tol)
    if cache and os.path.exists(cache):
        got = json.load(open(cache)).get(key)
        if got:
            return got
    t0 = time.time()
    idx = stratified_sample(sizes, sample_bytes)
    lens = [int(sizes[i]) for i in idx]
    soffs = np.concatenate([[0], np.cumsum(np.array(lens, dtype=np.uint64) + 255 & ~np.uint64(255))])
    sample = np.zeros(int(soffs[-1]), dtype=np.uint8)
    for j, i in enumerate(idx):
        sample[int(soffs[j]):int(soffs[j]) + lens[j]] = arena[int(offs[i]):int(offs[i]) + lens[j]]
    d_in = ctx.alloc(len(sample))
    d_in.upload(sample)
    d_o = ctx.alloc(len(sample))
    ins = [d_in.addr + int(o) for o in soffs[:-1]]
    # the largest output of any candidate: chunk-padded 1.5x plus per-chunk slack
    caps = [max(output_cap(n, file_params(n, c, t)) for c, t in STREAM_GRID) for n in lens]
    coffs = np.concatenate([[0], np.cumsum(np.array(caps, dtype=np.uint64) + 255 & ~np.uint64(255))])
    d_c = ctx.alloc(int(coffs[-1]))
    trials = []
    for c, t in STREAM_GRID:
        trials.append(_measure(ctx, d_in, ins, lens, d_c, d_o, c, t, 4, 0, coffs))
        if log:
            log("autotune: chunk %d KiB, 2^%d slots: ratio %.4f, %.1f MB/s" % (
                c // KiB, t, trials[-1]["ratio"], trials[-1]["score"] / 1e6))
    best = dict(choose(trials, ratio_tol))
    launch = []
    for spec in SPEC_GRID:
        for wpb in WPB_GRID:
            launch.append(_measure(ctx, d_in, ins, lens, d_c, d_o, best["chunk_max"],
                                   best["table_bits"], spec, wpb, coffs))
    for spec in TEAM_SPEC_GRID:
        launch.append(_measure(ctx, d_in, ins, lens, d_c, d_o, best["chunk_max"],
                               best["table_bits"], spec, 0, coffs, team=True))
    fast = max(launch, key=lambda t: t["score"])
    ctx.set_waves_per_block(0)
    ctx.set_team(0)
    res = {"chunk_max": best["chunk_max"], "table_bits": best["table_bits"],
           "spec_bits": fast["spec_bits"], "waves_per_block": fast["waves_per_block"],
           "team": fast["team"],
           "sample_files": len(idx), "sample_bytes": int(sum(lens)), "ratio_tol": ratio_tol,
           "seconds": time.time() - t0, "stream_trials": trials, "launch_trials": launch}
    if log:
        log("autotune: chose chunk %d KiB, 2^%d slots, %d-bit speculation, %s "
            "(%.1fs)" % (res["chunk_max"] // KiB, res["table_bits"], res["spec_bits"],
                         "teams of 4 waves" if res["team"] else
                         "%s waves/workgroup" % (res["waves_per_block"] or "auto"),
                         res["seconds"]))
    if cache:
        d = json.load(open(cache)) if os.path.exists(cache) else {}
        d[key] = res
        json.dump(d, open(cache, "w"), indent=1)
    for b in (d_in, d_o, d_c):
        b.free()
    return res
