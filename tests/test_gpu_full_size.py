"""The workloads at their full sizes on the GPU, checked through properties
that do not depend on size: every round trip lossless, sampled chunk
payloads (first, middle and last) byte-identical to the oracle's, container
framing consistent (batch lengths and end offsets), and 10,000 files in one
launch each decoded back to itself.

  * 256 MiB mixed stream, 1 MiB chunks, 2^22 slots per order (the metric's
    single-GPU configuration);
  * 1 GiB structured binary, 256 KiB chunks (4096 chunks), 2^22;
  * 10,000 lognormal files (median 64 KiB / 32, sigma 1.5) in one launch;
  * the batched configuration at its full size: 10,000 lognormal files
    (median 64 KiB, sigma 1.5, clipped to [512 B, 64 MiB]; 1.96 GiB) with the
    bench's settings, device-resident, one launch each way.
"""
import struct

import numpy as np
import pytest

from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu
MiB = 1 << 20


def _chunks(container, want):
    """Payloads of the chunks with indices in `want` (single-batch-size walk of
    the container's batch records; also checks the framing on the way)."""
    N = struct.unpack(">I", container[0:4])[0]
    batch = struct.unpack(">H", container[4:6])[0]
    S = struct.unpack(">I", container[8:12])[0]
    Pc, P = struct.unpack(">II", container[12:20])
    pos = 21 + (Pc if container[20] & 1 else P)
    C = (N - P + S - 1) // S
    got, c0 = {}, 0
    while c0 < C:
        Cb = min(batch, C - c0)
        blen = struct.unpack("<I", container[pos:pos + 4])[0]
        blob = container[pos + 4:pos + 4 + blen]
        prev = 4 * Cb - 1
        for c in range(Cb):
            e = struct.unpack(">I", blob[4 * c:4 * c + 4])[0]
            assert prev < e < blen
            if c0 + c in want:
                got[c0 + c] = blob[prev + 1:e + 1]
            prev = e
        assert prev == blen - 1
        pos += 4 + blen
        c0 += Cb
    assert pos == len(container)
    return got


def _check(ctx, oracle, data, S, tb, sample):
    p = make_params(S, 128, 100, 0, tb)
    blob = ctx.compress(data, p)
    n = len(data)
    C = (n + S - 1) // S
    want = {c for c in sample if c < C}
    pays = _chunks(blob, want)
    model = oracle.Model(tb)
    for c in sorted(want):
        ref = oracle.encode_chunk(data[c * S:(c + 1) * S], S, 128, model, tail_pad=c == C - 1)
        assert pays[c] == ref, "chunk %d" % c
    back = ctx.decompress(blob, 4)
    assert len(back) == n and back == data  # (plain bool: no diff of large buffers)
    return blob


def test_stream_256mib_1mib_chunks(ctx, oracle):
    data = datagen.mixed(256 * MiB, 2).tobytes()
    _check(ctx, oracle, data, MiB, 22, (0, 1, 128, 255))


def test_structured_1gib_256kib_chunks(ctx, oracle):
    data = datagen.structured(1024 * MiB, 4).tobytes()
    # one chunk of each region type and the last chunk
    _check(ctx, oracle, data, 256 << 10, 22, (0, 16, 32, 4095))


def test_ten_thousand_files_one_launch(ctx, oracle):
    sizes = datagen.lognormal_sizes(10000, 65536 // 32, 1.5, 16, 1 << 20, 7)
    arena, offs = datagen.markov2_files(sizes, seed0=1000)
    files = [arena[int(o):int(o) + int(n)].tobytes() for o, n in zip(offs, sizes)]
    p = make_params(16384, 128, 100, 0, 18)
    blobs = ctx.compress_many(files, p)
    assert ctx.stats()["launches"] == 1
    for i in np.linspace(0, len(files) - 1, 12).astype(int):
        assert blobs[i] == oracle.compress(files[i], 16384, 128, 100, 0, 18), "file %d" % i
    back = ctx.decompress_many(blobs)
    assert ctx.stats()["launches"] == 1
    bad = [i for i, (a, b) in enumerate(zip(files, back)) if a != b]
    assert not bad, bad[:10]


def test_batch_configuration_full_size(ctx, oracle):
    """All 10,000 files compressed in one launch and decompressed in another
    (the bench's settings and launch shapes), every file lossless, and a
    size-stratified sample of 24 files plus the largest byte-identical to the
    oracle's containers."""
    from cucm_amd.autotune import apply_launch, load_settings, stratified_sample
    from cucm_amd.codec import output_cap
    from cucm_amd.workloads import batch_files, pick_chunk_size
    cfg = load_settings()
    sizes, arena, offs = batch_files(10000)
    total = int(sizes.sum())
    assert sizes.min() >= 512 and sizes.max() <= 64 * MiB and total > 1900 * MiB
    S, tb = cfg["chunk_max"], cfg["table_bits"]
    params = [make_params(pick_chunk_size(int(n), S), 128, 100, 0, tb) for n in sizes]
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
    try:
        apply_launch(ctx, cfg["encoder"])
        cl, st = ctx.compress_many_device(ins, lens, params, cs, caps)
        assert not any(st) and ctx.stats()["launches"] == 1
        apply_launch(ctx, cfg["decoder"])
        dl, dst = ctx.decompress_many_device(cs, cl, outs, lens, cfg["decoder"]["spec_bits"])
        assert not any(dst) and ctx.stats()["launches"] == 1
        assert list(dl) == lens
        back = np.frombuffer(d_o.download(total), dtype=np.uint8)
        assert bool(np.array_equal(back, arena))
        pick = set(stratified_sample(sizes, 24 << 20, seed=5)[:24]) | {int(np.argmax(sizes))}
        for i in sorted(pick):
            n = int(sizes[i])
            ref = oracle.compress(arena[int(offs[i]):int(offs[i]) + n].tobytes(),
                                  pick_chunk_size(n, S), 128, 100, 0, tb)
            assert d_c.download(int(cl[i]), offset=int(coffs[i])) == ref, "file %d (%d B)" % (i, n)
    finally:
        ctx.set_team(0)
        ctx.set_waves_per_block(0)
