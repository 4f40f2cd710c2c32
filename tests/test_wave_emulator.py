"""The codec's per-wave device code (cucm_amd/csrc/cucm_wave.h), compiled for
the host and run under the fiber wave emulation of tests/wave_emu with
AddressSanitizer, against the oracle: byte-identical chunk payloads, lossless
decoding at every speculation width, and no out-of-bounds access of any LDS or
global buffer.  CPU only; the GPU parity tests run the same source on the
device."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from cucm_amd import datagen

HERE = os.path.dirname(os.path.abspath(__file__))
EMU_DIR = os.path.join(HERE, "wave_emu")
EMU = os.path.join(EMU_DIR, "build", "wave_emu")
CXX = "/opt/rocm/llvm/bin/clang++"

pytestmark = pytest.mark.skipif(not os.path.exists(CXX), reason="host clang++ not available")


@pytest.fixture(scope="module")
def emu():
    r = subprocess.run(["make", "-s", "-C", EMU_DIR], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    return EMU


def _payloads(blob):
    pos, out = 0, []
    while pos < len(blob):
        n = struct.unpack("<I", blob[pos:pos + 4])[0]
        out.append(blob[pos + 4:pos + 4 + n])
        pos += 4 + n
    return out


CASES = [
    # (kind, N, S, U, tb, tail_pad, spec_bits)
    ("markov2", 6000, 2048, 128, 12, True, 4),
    ("markov2", 4100, 1024, 64, 10, True, 1),
    ("mixed", 5000, 2048, 256, 14, True, 8),
    ("structured", 3000, 1024, 96, 11, False, 2),
    ("random", 1500, 512, 32, 9, True, 4),
    ("markov2", 40, 4096, 128, 8, True, 4),
]


def _data(kind, n):
    if kind == "zeros":
        return bytes(n)
    if kind == "runs":  # runs of 1..39 equal bytes: a slot hit again by the next byte
        rng = np.random.default_rng(4)
        return b"".join(bytes([int(rng.integers(0, 256))]) * int(rng.integers(1, 40))
                        for _ in range(n // 4 + 1))[:n]
    if kind == "markov2":
        return datagen.markov2(n, 3).tobytes()
    if kind == "mixed":
        return datagen.mixed(n, 5).tobytes()
    if kind == "structured":
        return datagen.structured(n, 6).tobytes()
    return np.random.default_rng(9).integers(0, 256, n, dtype=np.uint8).tobytes()


@pytest.mark.parametrize("kind,N,S,U,tb,pad,spec", CASES)
def test_emulated_wave_code_matches_oracle(emu, oracle, tmp_path, kind, N, S, U, tb, pad, spec):
    data = _data(kind, N)
    inp, enc, dec = tmp_path / "in", tmp_path / "enc", tmp_path / "dec"
    inp.write_bytes(data)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([emu, "enc", str(inp), str(S), str(U), str(tb), "1" if pad else "0", str(enc)],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    pays = _payloads(enc.read_bytes())
    C = (N + S - 1) // S
    assert len(pays) == C
    for c in range(C):
        ref = oracle.encode_chunk(data[c * S:(c + 1) * S], S, U, oracle.Model(tb),
                                  tail_pad=pad and c == C - 1)
        assert pays[c] == ref, "chunk %d" % c
    r = subprocess.run([emu, "dec", str(enc), str(N), str(S), str(U), str(tb), str(spec), str(dec)],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert dec.read_bytes() == data


TEAM_CASES = [
    # (kind, N, S, U, tb, spec_bits, seeds)
    ("markov2", 6000, 2048, 128, 12, 4, (1, 2, 3)),
    ("markov2", 4100, 1024, 64, 10, 1, (4,)),
    ("structured", 3000, 1024, 96, 11, 2, (5, 6)),
    ("mixed", 5000, 2048, 256, 14, 4, (7,)),
    ("random", 1500, 512, 32, 9, 4, (8, 9)),
    ("markov2", 40, 4096, 128, 8, 4, (10,)),
]


@pytest.mark.parametrize("kind,N,S,U,tb,spec,seeds", TEAM_CASES)
def test_emulated_team_decoder_is_lossless(emu, oracle, tmp_path, kind, N, S, U, tb, spec, seeds):
    """The team decoder (cucm_team.h): every chunk in turn on one emulated
    workgroup of four waves sharing one LDS, the waves stepped in orders drawn
    from several seeds, decodes the oracle's payloads losslessly."""
    data = _data(kind, N)
    C = (N + S - 1) // S
    blob = b""
    for c in range(C):
        p = oracle.encode_chunk(data[c * S:(c + 1) * S], S, U, oracle.Model(tb),
                                tail_pad=c == C - 1)
        blob += struct.pack("<I", len(p)) + p
    enc, dec = tmp_path / "enc", tmp_path / "dec"
    enc.write_bytes(blob)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    for seed in seeds:
        r = subprocess.run([emu, "tdec", str(enc), str(N), str(S), str(U), str(tb), str(spec),
                            str(seed), str(dec)], capture_output=True, text=True, env=env)
        assert r.returncode == 0, r.stderr[-4000:]
        assert dec.read_bytes() == data, "seed %d" % seed


# 2^8- and 2^9-slot tables make two bits of a byte, or bits of consecutive
# bytes, share slots often (the resolve's slow paths); zeros and runs make
# the same bit hit the same slot byte after byte (its forwarded case).
TEAM_ENC_CASES = TEAM_CASES + [
    ("zeros", 3000, 1024, 128, 8, 4, (1, 2)),
    ("random", 3000, 1024, 128, 8, 4, (3,)),
    ("markov2", 8000, 4096, 256, 8, 4, (4,)),
    ("runs", 6000, 2048, 128, 10, 4, (5, 6)),
]


@pytest.mark.parametrize("kind,N,S,U,tb,spec,seeds", TEAM_ENC_CASES)
def test_emulated_team_encoder_matches_oracle(emu, oracle, tmp_path, kind, N, S, U, tb, spec,
                                              seeds):
    """The team encoder (cucm_team.h): every chunk in turn on one emulated
    workgroup of four waves sharing one LDS, with the waves stepped in orders
    drawn from several seeds, produces the oracle's chunk payloads."""
    data = _data(kind, N)
    inp, enc = tmp_path / "in", tmp_path / "enc"
    inp.write_bytes(data)
    C = (N + S - 1) // S
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    for seed in seeds:
        r = subprocess.run([emu, "tenc", str(inp), str(S), str(U), str(tb), "1", str(seed),
                            str(enc)], capture_output=True, text=True, env=env)
        assert r.returncode == 0, r.stderr[-4000:]
        pays = _payloads(enc.read_bytes())
        assert len(pays) == C
        for c in range(C):
            ref = oracle.encode_chunk(data[c * S:(c + 1) * S], S, U, oracle.Model(tb),
                                      tail_pad=c == C - 1)
            assert pays[c] == ref, "seed %d chunk %d" % (seed, c)


FROZEN_CASES = [
    # (kind, N, S, tb, prelearn bytes, spec_bits)
    ("markov2", 6000, 2048, 12, 3000, 4),
    ("mixed", 5000, 1000, 14, 0, 1),
    ("structured", 4000, 4096, 11, 2500, 8),
    ("random", 1500, 512, 9, 700, 2),
    ("markov2", 33, 4096, 16, 400, 4),
    ("markov2", 700, 64, 10, 0, 8),
]


@pytest.mark.parametrize("kind,N,S,tb,P,spec", FROZEN_CASES)
def test_emulated_frozen_codec_matches_oracle(emu, oracle, tmp_path, kind, N, S, tb, P, spec):
    """The frozen-model encoder and decoder (cucm_frozen.h) on one emulated
    wave that keeps its model loaded across chunks: every chunk payload equal
    to the oracle's frozen payload, every chunk decoded losslessly."""
    data = _data(kind, N)
    pre = datagen.markov2(P, 21).tobytes() if P else b""
    inp, enc, dec, pf = tmp_path / "in", tmp_path / "enc", tmp_path / "dec", tmp_path / "pre"
    inp.write_bytes(data)
    pf.write_bytes(pre)
    model = oracle.Model(tb)
    if P:
        oracle.prelearn_train(pre, model)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    pre_arg = str(pf) if P else "-"
    r = subprocess.run([emu, "fenc", str(inp), str(S), str(tb), pre_arg, str(enc)],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    pays = _payloads(enc.read_bytes())
    C = (N + S - 1) // S
    assert len(pays) == C
    for c in range(C):
        assert pays[c] == oracle.encode_chunk_frozen(data[c * S:(c + 1) * S], model), "chunk %d" % c
    r = subprocess.run([emu, "fdec", str(enc), str(N), str(S), str(tb), str(spec), pre_arg, str(dec)],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert dec.read_bytes() == data


BIG_CASES = [
    # (kind, N, S, U, tb, tail_pad): units above the main kernels' 256 bits
    ("markov2", 3000, 1024, 288, 12, True),
    ("mixed", 5000, 2048, 512, 14, True),
    ("structured", 2100, 4096, 1024, 10, True),
    ("markov2", 9000, 4096, 4096, 12, True),
    ("zeros", 3000, 1024, 320, 8, True),
    ("runs", 4000, 2048, 2048, 9, False),
    ("random", 40, 4096, 65504, 8, True),
]


@pytest.mark.parametrize("kind,N,S,U,tb,pad", BIG_CASES)
def test_emulated_large_unit_codec_matches_oracle(emu, oracle, tmp_path, kind, N, S, U, tb, pad):
    """The large-unit codec (cucm_bigunit.h, U > 256) on one emulated wave
    whose model slot and unit records persist across chunks: chunk payloads
    equal to the oracle's (tail padding included), decodes lossless; zeros,
    runs and 2^8-slot tables make slots repeat inside an update block."""
    data = _data(kind, N)
    inp, enc, dec = tmp_path / "in", tmp_path / "enc", tmp_path / "dec"
    inp.write_bytes(data)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1")
    r = subprocess.run([emu, "benc", str(inp), str(S), str(U), str(tb), "1" if pad else "0",
                        str(enc)], capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    pays = _payloads(enc.read_bytes())
    C = (N + S - 1) // S
    assert len(pays) == C
    for c in range(C):
        ref = oracle.encode_chunk(data[c * S:(c + 1) * S], S, U, oracle.Model(tb),
                                  tail_pad=pad and c == C - 1)
        assert pays[c] == ref, "chunk %d" % c
    r = subprocess.run([emu, "bdec", str(enc), str(N), str(S), str(U), str(tb), str(dec)],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert dec.read_bytes() == data
