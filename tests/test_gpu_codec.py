"""GPU tests beyond the basic parity cases: golden fixtures, samples of the
large configurations, prelearn, many-file batches, plain vs batched launch,
device-resident buffers and error handling.  Everything goes through the
C ABI of libcucm.so."""
import hashlib
import json
import os

import numpy as np
import pytest

import cucm_amd
from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(GOLDEN, "manifest.json")))


@pytest.mark.parametrize("fx", MANIFEST, ids=[f["name"] for f in MANIFEST])
def test_golden_fixture(ctx, fx):
    data = open(os.path.join(GOLDEN, fx["name"] + ".in"), "rb").read()
    blob = open(os.path.join(GOLDEN, fx["name"] + ".cucm"), "rb").read()
    p = make_params(fx["chunk_size"], fx["unit_bits"], fx["batch_chunks"], fx["prelearn_size"],
                    fx["table_bits"], frozen=fx["frozen"])
    assert ctx.compress(data, p) == blob
    assert ctx.decompress(blob) == data


def test_mixed_stream_sample_2pow22(ctx, oracle):
    """4 chunks of the 256 MiB mixed stream (1 MiB chunks, 2^22 slots)."""
    data = datagen.mixed(4 << 20, 2).tobytes()
    ref = oracle.compress(data, 1 << 20, 128, 100, 0, 22)
    got = ctx.compress(data, make_params(1 << 20, 128, 100, 0, 22))
    assert got == ref
    assert ctx.decompress(got) == data


def test_structured_sample_2pow22(ctx, oracle):
    """Int32 walk / float32 field / record regions in 256 KiB chunks."""
    data = datagen.structured(12 << 20, 4)
    sample = np.concatenate([data[0:256 << 10], data[4 << 20:(4 << 20) + (256 << 10)],
                             data[8 << 20:(8 << 20) + (256 << 10)]]).tobytes()
    ref = oracle.compress(sample, 256 << 10, 128, 100, 0, 22)
    got = ctx.compress(sample, make_params(256 << 10, 128, 100, 0, 22))
    assert got == ref
    assert ctx.decompress(got) == sample


@pytest.mark.parametrize("P", [64, 3000])
def test_prelearn(ctx, oracle, P):
    data = datagen.markov2(20000, 21).tobytes()
    ref = oracle.compress(data, 4096, 128, 2, P, 18)
    got = ctx.compress(data, make_params(4096, 128, 2, P, 18))
    assert got == ref
    assert ctx.decompress(got) == data


def _files(count, seed, scale=64):
    sizes = datagen.lognormal_sizes(count, 65536 // scale, 1.5, 16, 1 << 20, seed)
    arena, offs = datagen.markov2_files(sizes, seed0=1000)
    return [arena[int(o):int(o) + int(n)].tobytes() for o, n in zip(offs, sizes)]


def test_many_files_one_launch_identical_to_oracle(ctx, oracle):
    files = _files(200, 3)
    p = make_params(4096, 128, 100, 0, 18)
    blobs = ctx.compress_many(files, p)
    assert ctx.stats()["launches"] == 1
    for f, b in zip(files[:40], blobs[:40]):
        assert b == oracle.compress(f, 4096, 128, 100, 0, 18)
    assert ctx.decompress_many(blobs) == files


def test_plain_path_equals_batched(ctx):
    files = _files(40, 4)
    p = make_params(8192, 128, 100, 0, 18)
    n = len(files)
    din = [ctx.alloc(len(f)) for f in files]
    for d, f in zip(din, files):
        d.upload(f)
    caps = [len(f) + len(f) // 2 + 8192 for f in files]
    dout_a = [ctx.alloc(c) for c in caps]
    dout_b = [ctx.alloc(c) for c in caps]
    la, sa = ctx.compress_many_device([d.addr for d in din], [len(f) for f in files], p,
                                      [d.addr for d in dout_a], caps)
    lb, _ = ctx.compress_many_device([d.addr for d in din], [len(f) for f in files], p,
                                     [d.addr for d in dout_b], caps, plain=True)
    assert la == lb and all(s == 0 for s in sa)
    assert ctx.stats()["launches"] == n
    for i in range(n):
        assert dout_a[i].download(la[i]) == dout_b[i].download(lb[i])
    back = [ctx.alloc(len(f)) for f in files]
    lens, st = ctx.decompress_many_device([d.addr for d in dout_a], la, [b.addr for b in back],
                                          [len(f) for f in files])
    assert all(s == 0 for s in st)
    for b, f in zip(back, files):
        assert b.download(len(f)) == f


def test_batch_records(ctx, oracle):
    data = datagen.markov2(50000, 22).tobytes()
    blob = ctx.compress_batch(data, 8192, 64, 16)
    assert blob == oracle.encode_batch(data, 8192, 64, oracle.Model(16))
    assert ctx.decompress_batch(blob, len(data), 8192, 64, 16) == data


@pytest.mark.parametrize("n,S", [(100, 1 << 20), (5000, 1 << 20), (100, 65536), (70000, 65536)])
def test_batch_records_much_smaller_than_chunk(ctx, oracle, n, S):
    """The last chunk is tail-padded to the full chunk size, so a blob of a
    few bytes can code to most of a chunk; the binding sizes its buffer from
    the padded length (and retries with the strict bound)."""
    data = datagen.markov2(n, 28).tobytes()
    blob = ctx.compress_batch(data, S, 128, 12)
    assert blob == oracle.encode_batch(data, S, 128, oracle.Model(12))
    assert ctx.decompress_batch(blob, n, S, 128, 12) == data


def test_no_tail_pad_flag(ctx, oracle):
    data = datagen.markov2(5000, 23).tobytes()
    a = ctx.compress(data, make_params(4096, 128, 100, 0, 18, tail_pad=False))
    b = ctx.compress(data, make_params(4096, 128, 100, 0, 18))
    assert len(a) < len(b)
    assert ctx.decompress(a) == data and oracle.decompress(a) == data


def test_corrupt_and_truncated_containers_are_reported(ctx):
    data = datagen.markov2(30000, 24).tobytes()
    blob = ctx.compress(data, make_params(4096, 128, 100, 0, 18))
    uf = (blob[6] << 8 | blob[7]) & ~31
    # table bits outside 8..25 (no encoder writes them)
    tbs = [blob[:6] + (uf | t).to_bytes(2, "big") + blob[8:] for t in (3, 26)]
    bad = [blob[:10], blob[:200], blob[:21] + b"\xff" * 40 + blob[61:]] + tbs
    n = len(bad)
    import ctypes
    from cucm_amd._native import lib
    outs = [np.empty(40000, dtype=np.uint8) for _ in bad]
    ins = (ctypes.c_void_p * n)(*[ctypes.cast(ctypes.c_char_p(b), ctypes.c_void_p).value for b in bad])
    lens = (ctypes.c_uint64 * n)(*[len(b) for b in bad])
    outp = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    caps = (ctypes.c_uint64 * n)(*[40000] * n)
    olen = (ctypes.c_uint64 * n)()
    st = (ctypes.c_int32 * n)()
    rc = lib().cucm_decompress_many(ctx.handle, n, ins, lens, outp, caps, olen, st, 4)
    assert rc == 0
    assert list(st) == [-2] * n
    small = (ctypes.c_uint64 * 1)(100)
    ins1 = (ctypes.c_void_p * 1)(ctypes.cast(ctypes.c_char_p(blob), ctypes.c_void_p).value)
    lens1 = (ctypes.c_uint64 * 1)(len(blob))
    st1 = (ctypes.c_int32 * 1)()
    rc = lib().cucm_decompress_many(ctx.handle, 1, ins1, lens1, outp, small, olen, st1, 4)
    assert rc == 0 and st1[0] == -3
    # the context is still usable
    assert ctx.decompress(blob) == data


def test_crafted_chunk_counts_are_corrupt_not_sized(ctx):
    """A 21-byte header claiming N = 2^31 bytes in 1-byte chunks has no room
    for its 2^31 end offsets; two of them in one call once wrapped the 32-bit
    task count.  Both must come back CUCM_E_CORRUPT while the valid container
    beside them decodes (each chunk needs a 4-byte offset and a payload
    byte, so a count above (len - header) / 5 is refused before sizing)."""
    import ctypes
    import struct
    from cucm_amd._native import lib
    data = datagen.markov2(20000, 27).tobytes()
    good = ctx.compress(data, make_params(4096, 128, 100, 0, 18))
    crafted = (struct.pack(">IHHIII", 1 << 31, 100, 128 | 18, 1, 0, 0) + b"\x00" +
               struct.pack("<I", 0) + b"\x00" * 11)
    bad = [crafted, crafted, good]
    n = len(bad)
    outs = [np.empty(len(data), dtype=np.uint8) for _ in bad]
    keep = [ctypes.create_string_buffer(b, len(b)) for b in bad]
    ins = (ctypes.c_void_p * n)(*[ctypes.addressof(k) for k in keep])
    lens = (ctypes.c_uint64 * n)(*[len(b) for b in bad])
    outp = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
    caps = (ctypes.c_uint64 * n)(*([1 << 31, 1 << 31, len(data)]))
    olen = (ctypes.c_uint64 * n)()
    st = (ctypes.c_int32 * n)()
    # device-resident decode: the crafted containers' capacity is not backed
    # by memory, so only a refusal before any task is built keeps this safe
    d_in = [ctx.alloc(len(b)) for b in bad]
    for d, b in zip(d_in, bad):
        d.upload(b)
    d_out = ctx.alloc(len(data))
    dins = (ctypes.c_void_p * n)(*[d.addr for d in d_in])
    douts = (ctypes.c_void_p * n)(*[d_out.addr] * n)
    rc = lib().cucm_decompress_many_device(ctx.handle, n, dins, lens, douts, caps, olen, st, 4)
    assert rc == 0
    assert list(st) == [-2, -2, 0]
    assert olen[2] == len(data)
    assert d_out.download(len(data)) == data


def test_bad_params_rejected(ctx):
    with pytest.raises(cucm_amd.CucmError):
        ctx.compress(b"abc", make_params(0, 128))
    with pytest.raises(cucm_amd.CucmError):
        ctx.compress(b"abc", make_params(4096, 100))
    p = make_params(4096, 128)
    p.flags = 0x4  # no such option
    with pytest.raises(cucm_amd.CucmError) as e:
        ctx.compress(b"abc", p)
    assert e.value.code == -1


def test_one_shot_api(oracle):
    data = datagen.markov2(9000, 25).tobytes()
    blob = cucm_amd.compress(data, 2048, 128, 2, 0, 16)
    assert blob == oracle.compress(data, 2048, 128, 2, 0, 16)
    assert cucm_amd.decompress(blob) == data


def test_file_api(oracle, tmp_path):
    """compress_file / decompress_file (the reference's file-level calls):
    the container equals the oracle's, and every maxassumed_length, and the
    automatic width, decode it."""
    data = datagen.markov2(60000, 27).tobytes()
    src, enc, dec = tmp_path / "in", tmp_path / "in.cucm", tmp_path / "out"
    src.write_bytes(data)
    n = cucm_amd.compress_file(str(src), str(enc), chunksize=8192, unitlength=128,
                               prelearn_size=2000, batch_num=3, table_bits=16)
    assert enc.read_bytes() == oracle.compress(data, 8192, 128, 3, 2000, 16)
    assert n == enc.stat().st_size
    for m in (None, 0, 1, 3, 7):
        dec.unlink(missing_ok=True)
        assert cucm_amd.decompress_file(str(enc), str(dec), maxassumed_length=m) == len(data)
        assert dec.read_bytes() == data, m


def test_small_budget_is_refused_cleanly():
    with cucm_amd.Context(0, hbm_budget=1 << 20) as c:
        with pytest.raises(cucm_amd.CucmError) as e:
            c.compress(datagen.markov2(100000, 26).tobytes(), make_params(16384, 128, 100, 0, 22))
        assert e.value.code == -5
