"""Autotuner for batched compression of many files of uneven size.

It measures candidate settings on a size-stratified sample of the files
(device-resident, one persistent launch per direction, kernel time from HIP
events) and picks:

  - stream settings: the largest chunk size (a file up to it is one chunk;
    longer files are cut into chunks of that size) and the table bits.  These
    change the container, so a candidate is eligible only if its compression
    ratio on the sample is at least (1 - ratio_tol) x the ratio of the
    reference setting (2^18 slots, 256 KiB chunks); the score is sample bytes
    per second of compress + decompress kernel time;
  - launch settings, which never change the container, per direction: for
    the encoder team execution (one chunk per workgroup of four waves) or
    one wave per chunk with 2, 4, 8 or automatically many waves per
    workgroup; for the decoder the same plus the speculation width (bits
    decided per round: 1, 2, 4 or 8), each scored by its own kernel time.

The sample must hold several passes of resident chunks (sample_bytes
defaults to 640 MiB): with fewer chunks than the device holds at once the
launch is latency-bound and favours wide speculation, while the full batch
is bound by random-line traffic, where narrow speculation (fewer gathers per
bit) wins.  `choose` is the pure selection rule; `autotune_batch` runs the
measurements; `save_settings`/`load_settings` keep the result in a file.
"""
import json
import os
import time

import numpy as np

from .codec import make_params, output_cap

KiB = 1024
REFERENCE = {"chunk_max": 256 * KiB, "table_bits": 18}
STREAM_GRID = [(c, t) for c in (128 * KiB, 256 * KiB, 1024 * KiB) for t in (16, 18)]
SPEC_GRID = (1, 2, 4)  # 8-bit rounds (255 paths) gather ~4x more per bit than 4-bit ones
WPB_GRID = (0, 2, 4, 8)
TEAM_SPEC_GRID = (1, 2, 4)  # speculation widths the team decoder runs
SETTINGS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "batch_settings.json")


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


def _run(ctx, ins, lens, params, cs, caps, outs, direction, launch, cl=None):
    """One launch of one direction with the given launch settings; returns
    (kernel ms, compressed lengths)."""
    ctx.set_waves_per_block(launch["waves_per_block"])
    ctx.set_team(2 if launch["team"] else 1)
    ctx.set_max_waves(launch.get("max_waves", 0))
    if direction == "encoder":
        cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
    else:
        _, st = ctx.decompress_many_device(cs, cl, outs, lens, launch["spec_bits"])
    if any(st):
        raise RuntimeError("autotune trial failed")
    return ctx.stats()["codec_ms"], cl


# max_waves caps the resident waves (teams: 4 per chunk): on the 10,000-file
# batch the encoder's teams run fastest at 2.25 per CU, not at the 3 LDS
# allows (576 teams 3,650 ms, 640 teams 3,770 ms, 512 teams 3,805 ms, 768
# teams ~3,980 ms, measured on one box), since the memory system serves
# fewer chunks in flight better
RESIDENCY_GRID = (2048, 2304, 2560)
ENCODER_LAUNCH = [{"team": True, "waves_per_block": 0}] + \
    [{"team": True, "waves_per_block": 0, "max_waves": m} for m in RESIDENCY_GRID] + \
    [{"team": False, "waves_per_block": w} for w in WPB_GRID]
DECODER_LAUNCH = [{"team": True, "waves_per_block": 0, "spec_bits": d} for d in TEAM_SPEC_GRID] + \
    [{"team": False, "waves_per_block": w, "spec_bits": d} for d in SPEC_GRID for w in WPB_GRID] + \
    [{"team": False, "waves_per_block": 0, "spec_bits": 2, "max_waves": 1536}]


def autotune_batch(ctx, arena, offs, sizes, sample_bytes=640 << 20, ratio_tol=0.01, log=None,
                   stream_grid=None, encoder_launch=None, decoder_launch=None):
    """Tunes the batch settings on a sample of the files (arena/offs/sizes as
    datagen.markov2_files returns them).  Returns a dict with chunk_max,
    table_bits, the encoder's and the decoder's launch settings, and the
    trials."""
    stream_grid = stream_grid or STREAM_GRID
    encoder_launch = encoder_launch or ENCODER_LAUNCH
    decoder_launch = decoder_launch or DECODER_LAUNCH
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
    outs = [d_o.addr + int(o) for o in soffs[:-1]]
    # the largest output of any candidate: chunk-padded 1.5x plus per-chunk slack
    caps = [max(output_cap(n, file_params(n, c, t)) for c, t in stream_grid) for n in lens]
    coffs = np.concatenate([[0], np.cumsum(np.array(caps, dtype=np.uint64) + 255 & ~np.uint64(255))])
    d_c = ctx.alloc(int(coffs[-1]))
    cs = [d_c.addr + int(o) for o in coffs[:-1]]
    total = float(sum(lens))
    enc0 = {"team": True, "waves_per_block": 0}
    dec0 = {"team": True, "waves_per_block": 0, "spec_bits": 4}
    trials = []
    for c, t in stream_grid:
        params = [file_params(n, c, t) for n in lens]
        kc, cl = _run(ctx, ins, lens, params, cs, caps, outs, "encoder", enc0)
        kd, _ = _run(ctx, ins, lens, params, cs, caps, outs, "decoder", dec0, cl)
        trials.append({"chunk_max": c, "table_bits": t, "ratio": total / float(sum(cl)),
                       "compress_ms": kc, "decompress_ms": kd, "score": total / ((kc + kd) / 1e3)})
        if log:
            log("autotune: chunk %d KiB, 2^%d slots: ratio %.4f, %.1f MB/s" % (
                c // KiB, t, trials[-1]["ratio"], trials[-1]["score"] / 1e6))
    best = dict(choose(trials, ratio_tol))
    params = [file_params(n, best["chunk_max"], best["table_bits"]) for n in lens]
    enc_trials, dec_trials = [], []
    cl = None
    for launch in encoder_launch:
        k, cl = _run(ctx, ins, lens, params, cs, caps, outs, "encoder", launch)
        enc_trials.append(dict(launch, kernel_ms=k, MBps=total / 1e3 / k))
    for launch in decoder_launch:
        k, _ = _run(ctx, ins, lens, params, cs, caps, outs, "decoder", launch, cl)
        dec_trials.append(dict(launch, kernel_ms=k, MBps=total / 1e3 / k))
        if log:
            log("autotune: decoder %s: %.1f MB/s" % (launch, dec_trials[-1]["MBps"]))
    enc = min(enc_trials, key=lambda t: t["kernel_ms"])
    dec = min(dec_trials, key=lambda t: t["kernel_ms"])
    ctx.set_waves_per_block(0)
    ctx.set_team(0)
    ctx.set_max_waves(0)
    keys = ("team", "waves_per_block", "max_waves", "spec_bits")
    res = {"chunk_max": best["chunk_max"], "table_bits": best["table_bits"],
           "encoder": {k: enc[k] for k in keys if k in enc},
           "decoder": {k: dec[k] for k in keys if k in dec},
           "sample_files": len(idx), "sample_bytes": int(total), "ratio_tol": ratio_tol,
           "seconds": time.time() - t0, "stream_trials": trials,
           "encoder_trials": enc_trials, "decoder_trials": dec_trials}
    if log:
        log("autotune: chose chunk %d KiB, 2^%d slots; encoder %s; decoder %s (%.1fs)" % (
            res["chunk_max"] // KiB, res["table_bits"], res["encoder"], res["decoder"],
            res["seconds"]))
    for b in (d_in, d_o, d_c):
        b.free()
    return res


def apply_launch(ctx, launch):
    """Sets a direction's launch settings (team: True always, False never,
    None automatic; waves per workgroup, 0 automatic; resident waves at most
    max_waves, 0 as many as fit) on ctx."""
    ctx.set_waves_per_block(launch.get("waves_per_block", 0))
    ctx.set_max_waves(launch.get("max_waves", 0))
    team = launch.get("team")
    ctx.set_team(0 if team is None else (2 if team else 1))


def save_settings(res, path=SETTINGS, note=""):
    keep = {k: res[k] for k in ("chunk_max", "table_bits", "encoder", "decoder", "sample_files",
                                "sample_bytes", "ratio_tol")}
    keep["note"] = note
    keep["trials"] = {"stream": res["stream_trials"], "encoder": res["encoder_trials"],
                      "decoder": res["decoder_trials"]}
    with open(path, "w") as f:
        json.dump(keep, f, indent=1)


def load_settings(path=SETTINGS):
    with open(path) as f:
        return json.load(f)
