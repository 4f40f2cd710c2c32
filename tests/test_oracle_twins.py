"""Cross-checks of independently structured restatements inside the oracle.

The unit schedule, the reference's stale-slot tail padding and the framing
are pinned only by code reading (the reference has no tests and cannot run
here).  Two encoders (clean per-chunk sequential; literal kernel-by-kernel
over all chunks with the reference's staging buffers and guards) and two
decoders (clean sequential; literal lock-step speculative restatement of the
reference decoder) must agree, and every container must round-trip.
"""
import random

import numpy as np
import pytest

from cucm_amd import datagen
from oracle import oracle as O


def rnd_bytes(n, seed, alphabet=None):
    r = random.Random(seed)
    if alphabet is None:
        return bytes(r.getrandbits(8) for _ in range(n))
    return bytes(r.choice(alphabet) for _ in range(n))


CASES = [
    # N, S, U, tb, data kind
    (1000, 256, 32, 18, "text"), (5000, 1024, 128, 12, "text"), (777, 100, 64, 12, "text"),
    (4096, 4096, 128, 18, "text"), (10, 4096, 128, 18, "text"), (3000, 1000, 96, 10, "text"),
    (2500, 512, 256, 9, "zeros"), (3333, 333, 32, 14, "random"), (6000, 1500, 160, 16, "markov"),
    (100, 8, 32, 8, "text"), (64, 64, 512, 10, "text"), (9000, 2048, 64, 13, "markov"),
]


def make(kind, n, seed):
    if kind == "text":
        return rnd_bytes(n, seed, b"abcab cde\x00\x01the ")
    if kind == "zeros":
        return bytes(n)
    if kind == "random":
        return rnd_bytes(n, seed)
    return datagen.markov2(n, seed).tobytes()


@pytest.mark.parametrize("N,S,U,tb,kind", CASES)
def test_clean_and_literal_encoders_agree(N, S, U, tb, kind):
    data = make(kind, N, N + S)
    m = O.Model(tb)
    clean = O.encode_batch(data, S, U, m)
    literal = O.encode_batch(data, S, U, m, literal=True)
    assert clean == literal
    assert O.decode_batch(clean, N, S, U, m) == data


@pytest.mark.parametrize("N,S,U,tb,kind", [c for c in CASES if c[1] >= 8])
@pytest.mark.parametrize("m", [0, 1, 3])
def test_reference_decoder_restatement_accepts_stream(N, S, U, tb, kind, m):
    """Every case's chunk payloads are at least 4 bytes, so the reference
    decoder's 4-byte x initialisation stays inside each payload (flags 0);
    its restatement must then return the input for every speculation depth."""
    data = make(kind, N, N + S)
    model = O.Model(tb)
    blob = O.encode_batch(data, S, U, model)
    C = (N + S - 1) // S
    ends = [int.from_bytes(blob[4 * c:4 * c + 4], "big") for c in range(C)]
    sizes = [e - p for e, p in zip(ends, [4 * C - 1] + ends[:-1])]
    assert min(sizes) >= 4
    out, flags = O.decode_batch(blob, N, S, U, model, literal_m=m)
    assert flags == 0
    assert out == data


def test_chunks_below_8_bytes_round_trip():
    """With S < 8 the reference decoder's global previous-byte guard can read
    an already decoded byte of the previous chunk where the encoder saw 0
    (cucm_decompress_v2.cu:716); the codec uses the encoder's chunk-relative
    rule, which always round-trips."""
    for S in (1, 2, 3, 4, 7):
        data = make("text", 64, S)
        m = O.Model(10)
        blob = O.encode_batch(data, S, 32, m)
        assert O.decode_batch(blob, 64, S, 32, m) == data


def test_unit_of_one_bit_equals_prelearn_coder():
    data = make("markov", 3000, 11)
    m1 = O.Model(16)
    seg = O.prelearn_encode(data, m1)
    m0 = O.Model(16)
    assert O.encode_chunk(data, len(data), 1, m0) == seg
    m2 = O.Model(16)
    assert O.prelearn_decode(seg, len(data), m2) == data
    assert m2.digest() == m1.digest()
    m3 = O.Model(16)
    O.prelearn_train(data, m3)
    assert m3.digest() == m1.digest()


def test_chunk_payload_depends_only_on_its_bytes():
    data = make("markov", 9000, 12)
    S, U, m = 2048, 128, O.Model(14)
    blob = O.encode_batch(data, S, U, m)
    C = (len(data) + S - 1) // S
    ends = [int.from_bytes(blob[4 * c:4 * c + 4], "big") for c in range(C)]
    prev = 4 * C - 1
    for c in range(C):
        chunk = data[c * S:(c + 1) * S]
        alone = O.encode_chunk(chunk, S, U, m, tail_pad=(c == C - 1), L=len(chunk))
        assert blob[prev + 1:ends[c] + 1] == alone
        prev = ends[c]


def test_batch_size_changes_only_framing():
    data = make("markov", 20000, 13)
    a = O.compress(data, 2048, 128, 100, 0, 14)
    b = O.compress(data, 2048, 128, 3, 0, 14)
    assert a != b
    assert O.decompress(a) == data and O.decompress(b) == data

    def payloads(blob):
        import struct
        N = struct.unpack(">I", blob[:4])[0]
        batch = struct.unpack(">H", blob[4:6])[0]
        S = struct.unpack(">I", blob[8:12])[0]
        C = (N + S - 1) // S
        pos, out = 21, []
        for c0 in range(0, C, batch):
            Cb = min(batch, C - c0)
            blen = struct.unpack("<I", blob[pos:pos + 4])[0]
            bl = blob[pos + 4:pos + 4 + blen]
            prev = 4 * Cb - 1
            for c in range(Cb):
                e = struct.unpack(">I", bl[4 * c:4 * c + 4])[0]
                out.append(bl[prev + 1:e + 1])
                prev = e
            pos += 4 + blen
        return out

    assert payloads(a) == payloads(b)


def test_tail_padding_only_changes_last_chunk():
    data = make("markov", 5000, 14)
    m = O.Model(14)
    padded = O.encode_batch(data, 2048, 128, m, tail_pad=True)
    plain = O.encode_batch(data, 2048, 128, m, tail_pad=False)
    assert len(plain) < len(padded)
    assert O.decode_batch(plain, 5000, 2048, 128, m) == data
    assert O.decode_batch(padded, 5000, 2048, 128, m) == data
    # the first two chunks' payloads are identical
    assert padded[12:12 + 100] == plain[12:12 + 100]


@pytest.mark.parametrize("P", [1, 100, 2048, 4000])
def test_prelearn_containers_round_trip(P):
    data = make("markov", 4000, 15)
    blob = O.compress(data, 1024, 64, 2, P, 16)
    assert O.decompress(blob) == data
    flag = blob[20]
    Pc = int.from_bytes(blob[12:16], "big")
    assert flag == (1 if Pc < P else 0)


def test_header_layout_reference_fields():
    data = make("text", 3000, 16)
    blob = O.compress(data, 1000, 96, 2, 0, 25)
    assert blob[:4] == (3000).to_bytes(4, "big")
    assert blob[4:6] == (2).to_bytes(2, "big")
    assert blob[6:8] == (96).to_bytes(2, "big")   # table_bits 25: U field unchanged
    assert blob[8:12] == (1000).to_bytes(4, "big")
    assert blob[12:21] == bytes(9)
    # batch records: LE length then BE end offsets
    blen = int.from_bytes(blob[21:25], "little")
    assert 21 + 4 + blen < len(blob)
    tb = O.compress(data, 1000, 96, 2, 0, 18)
    assert tb[6:8] == (96 | 18).to_bytes(2, "big")


def test_empty_input_is_header_only():
    blob = O.compress(b"", 4096, 128, 100, 0, 18)
    assert len(blob) == 21 and O.decompress(blob) == b""


@pytest.mark.parametrize("n", [1, 11, 12, 999, 1000, 1001, 1999, 2000, 2001])
def test_edge_sizes_twins_agree_and_round_trip(n):
    """Around one unit (U = 96: 12 bytes) and one and two chunks (S = 1000;
    U does not divide 8S): the clean and literal encoders agree, and both
    decoders (clean, and the restatement of the reference's lock-step
    decoder) give the input back."""
    S, U, tb = 1000, 96, 12
    data = make("text", n, n)
    m = O.Model(tb)
    clean = O.encode_batch(data, S, U, m)
    assert clean == O.encode_batch(data, S, U, m, literal=True)
    assert O.decode_batch(clean, n, S, U, m) == data
    back, _ = O.decode_batch(clean, n, S, U, m, literal_m=3)
    assert back == data
