"""Streaming container I/O and the drop-in command-line tools on the GPU:
containers produced segment by segment (and by cucm_compress) are
byte-identical to one-shot compression, and decode back through the stream
API (and cucm_decompress) at every speculation width."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import cucm_amd
from cucm_amd import datagen, make_params
from cucm_amd._native import lib

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
COMP = os.path.join(ROOT, "cucm_amd", "bin", "cucm_compress")
DECOMP = os.path.join(ROOT, "cucm_amd", "bin", "cucm_decompress")

CASES = [
    # (N, S, batch, P, tb, tail_pad)
    (300_000, 16384, 4, 0, 18, True),
    (300_000, 16384, 3, 5000, 16, True),
    (123_457, 8192, 5, 0, 14, False),
    (65_536, 16384, 4, 0, 18, True),
]


@pytest.mark.parametrize("N,S,batch,P,tb,pad", CASES)
def test_stream_equals_one_shot(ctx, N, S, batch, P, tb, pad):
    data = datagen.markov2(N, 41).tobytes()
    p = make_params(S, 128, batch, P, tb, tail_pad=pad)
    ref = ctx.compress(data, p)
    for seg in (1, 3):
        assert ctx.compress_stream(data, p, segment_batches=seg) == ref
    for per_call, spec in ((1, 4), (2, 1), (5, 8)):
        assert ctx.decompress_stream(ref, spec_bits=spec, records_per_call=per_call) == data


def test_stream_rejects_misaligned_segments_and_partial_records(ctx):
    data = datagen.markov2(100_000, 5).tobytes()
    p = make_params(8192, 128, 2, 0, 14)
    head = np.empty(64, dtype=np.uint8)
    hl = ctypes.c_uint64()
    s = ctypes.c_void_p()
    assert lib().cucm_stream_compress_begin(ctx.handle, len(data), ctypes.byref(p), data,
                                            head.ctypes.data, 64, ctypes.byref(hl),
                                            ctypes.byref(s)) == 0
    out = np.empty(1 << 20, dtype=np.uint8)
    olen = ctypes.c_uint64()
    # 10000 bytes is neither a whole number of batch records nor the rest
    assert lib().cucm_stream_compress_next(s, data, 10000, out.ctypes.data, out.size,
                                           ctypes.byref(olen)) == -1
    lib().cucm_stream_end(s)
    blob = ctx.compress(data, p)
    n0 = ctypes.c_uint64()
    q = cucm_amd.Params()
    assert lib().cucm_stream_decompress_begin(ctx.handle, blob, 21, 4, ctypes.byref(n0),
                                              ctypes.byref(q), None, 0, ctypes.byref(s)) == 0
    first = 4 + int.from_bytes(blob[21:25], "little")
    rec = blob[21:21 + first - 3]  # a record cut short
    assert lib().cucm_stream_decompress_next(s, rec, len(rec), out.ctypes.data, out.size,
                                             ctypes.byref(olen)) == -2
    lib().cucm_stream_end(s)
    # table bits outside 8..25 in the header
    uf = (blob[6] << 8 | blob[7]) & ~31
    for t in (3, 26):
        bad = blob[:6] + (uf | t).to_bytes(2, "big") + blob[8:21]
        assert lib().cucm_stream_decompress_begin(ctx.handle, bad, 21, 4, ctypes.byref(n0),
                                                  ctypes.byref(q), None, 0,
                                                  ctypes.byref(s)) == -2


def test_stream_refuses_sizes_beyond_the_container_fields(ctx):
    p = make_params(1 << 20, 128, 100, 0, 18)
    head = np.empty(64, dtype=np.uint8)
    hl = ctypes.c_uint64()
    s = ctypes.c_void_p()
    assert lib().cucm_stream_compress_begin(ctx.handle, 1 << 32, ctypes.byref(p), None,
                                            head.ctypes.data, 64, ctypes.byref(hl),
                                            ctypes.byref(s)) == -1


def _tool(args):
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr + r.stdout
    return r.stdout


def test_cli_round_trip_identical_to_library(ctx, tmp_path):
    data = datagen.markov2(200_000, 77).tobytes()
    src, enc, dec = tmp_path / "in", tmp_path / "in.cucm", tmp_path / "out"
    src.write_bytes(data)
    out = _tool([COMP, "-f", str(src), "-o", str(enc), "-s", "16K", "-b", "4", "-t", "18",
                 "-p", "3000"])
    assert "Compression ratio:" in out
    assert enc.read_bytes() == ctx.compress(data, make_params(16384, 128, 4, 3000, 18))
    for a in (None, "0", "1", "3", "7"):
        _tool([DECOMP, "-i", str(enc), "-o", str(dec)] + (["-a", a] if a else []))
        assert dec.read_bytes() == data


def test_cli_chunknum_truncates_the_input(ctx, tmp_path):
    data = datagen.markov2(100_000, 78).tobytes()
    src, enc = tmp_path / "in", tmp_path / "in.cucm"
    src.write_bytes(data)
    _tool([COMP, "-f", str(src), "-o", str(enc), "-s", "8K", "-c", "5", "-t", "16", "-u", "64"])
    assert enc.read_bytes() == ctx.compress(data[:5 * 8192], make_params(8192, 64, 100, 0, 16))


def test_cli_frozen_round_trip(ctx, oracle, tmp_path):
    """--frozen: the container equals the oracle's frozen container, and the
    decompressor reads the mode from the header."""
    data = datagen.markov2(150_000, 79).tobytes()
    src, enc, dec = tmp_path / "in", tmp_path / "in.cucm", tmp_path / "out"
    src.write_bytes(data)
    _tool([COMP, "-f", str(src), "-o", str(enc), "-s", "16K", "-b", "3", "-t", "16",
           "-p", "20000", "--frozen"])
    blob = enc.read_bytes()
    assert blob == oracle.compress(data, 16384, 128, 3, 20000, 16, frozen=True)
    _tool([DECOMP, "-i", str(enc), "-o", str(dec), "-a", "3"])
    assert dec.read_bytes() == data
