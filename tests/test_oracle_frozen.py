"""The oracle's frozen-model mode (orc_*_frozen): every bit predicted from
the initial model, no updates, exactly 8L coded bits per chunk.

The reference has no code for this mode (it is the paper's "frozen model",
described but not shipped), so nothing from the reference pins it.  What
pins the restatement instead:
  * a closed form: with the zero model every prediction is pr = 2048, which
    halves the coder interval exactly, so a chunk's payload is its bytes
    complemented (y = 1 takes the lower half) followed by a 0x00 flush byte;
  * the normal path: within a chunk's first unit the normal encoder predicts
    from the unit-start model, i.e. the initial model, and updates nothing
    that any of those predictions reads, so a chunk of at most one unit
    coded without tail padding is byte-identical in both modes;
  * a Python twin composed from the oracle's primitives (slots, StateMap
    prediction, stretch, mixer, coder step), each pinned by its own
    known-answer tests;
  * round trips through the container, with the flag bit read back.
"""
import random

import numpy as np
import pytest

from cucm_amd import datagen
from oracle import oracle as O


def _prelearned(tb, P=6000, seed=3):
    m = O.Model(tb)
    O.prelearn_train(datagen.markov2(P, seed).tobytes(), m)
    return m


@pytest.mark.parametrize("n", [1, 2, 7, 64, 1000, 4097])
def test_zero_model_payload_is_the_complement(n):
    d = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8)
    pay = O.encode_chunk_frozen(d.tobytes(), O.Model(12))
    assert pay == (~d).tobytes() + b"\x00"
    assert O.decode_chunk_frozen(pay, n, O.Model(12)) == d.tobytes()


@pytest.mark.parametrize("L,tb", [(1, 12), (16, 16), (100, 18), (1000, 14), (8188, 20)])
def test_one_unit_chunks_match_the_normal_encoder(L, tb):
    m = _prelearned(tb)
    data = datagen.markov2(L, 40 + L).tobytes()
    U = max(32, (8 * L + 31) // 32 * 32)
    normal = O.encode_chunk(data, L, U, m, tail_pad=False)
    assert O.encode_chunk_frozen(data, m) == normal


def _twin_frozen_payload(chunk, m, tb):
    """Frozen coding restated from the oracle's pinned primitives."""
    st = O.tables()[1]
    x1, x2, out = 0, 0xFFFFFFFF, b""
    buf = bytearray(len(chunk))
    for b in range(8 * len(chunk)):
        # the context reads only the bits before b
        byte = b >> 3
        buf[byte] = chunk[byte] & (0xFF00 >> (b & 7)) & 0xFF
        sl = O.slots(bytes(buf), b, tb)
        smp = [st[m.predict(k, sl[k])] for k in range(8)]
        c1 = chunk[byte - 1] if byte >= 1 else 0
        cxt = (c1 >> 5) * 10 + sum(1 for v in smp if v != 1)
        pr = m.mix(cxt, smp)
        pr += 1 if pr < 2048 else 0
        y = (chunk[byte] >> (7 - (b & 7))) & 1
        x1, x2, emitted = O.coder_step(x1, x2, pr, y)
        out += emitted
        buf[byte] = chunk[byte] & (0xFF00 >> ((b & 7) + 1)) & 0xFF
    return out + bytes([x1 >> 24])


@pytest.mark.parametrize("L,tb", [(150, 14), (40, 10)])
def test_python_twin_from_primitives(L, tb):
    m = _prelearned(tb, P=3000)
    chunk = datagen.markov2(L, 77).tobytes()
    assert O.encode_chunk_frozen(chunk, m) == _twin_frozen_payload(chunk, m, tb)


def test_frozen_differs_from_normal_past_the_first_unit():
    m = _prelearned(16)
    data = datagen.markov2(4000, 5).tobytes()
    assert O.encode_chunk_frozen(data, m) != O.encode_chunk(data, 4000, 128, m, tail_pad=False)


@pytest.mark.parametrize("N,S,U,batch,P,tb", [
    (0, 4096, 128, 100, 0, 16), (1, 4096, 128, 100, 0, 16), (30000, 4096, 64, 3, 0, 14),
    (30000, 4096, 128, 100, 5000, 16), (12345, 1000, 32, 4, 12345, 12),
    (20000, 8192, 256, 2, 300, 18), (9000, 2048, 96, 100, 10, 25),
])
def test_container_round_trip_and_flag(N, S, U, batch, P, tb):
    data = datagen.markov2(N, N + P).tobytes() if N else b""
    blob = O.compress(data, S, U, batch, P, tb, frozen=True)
    assert blob[20] & 2
    assert O.decompress(blob) == data
    normal = O.compress(data, S, U, batch, P, tb)
    assert blob[20] == normal[20] | 2
    assert blob[:20] == normal[:20]
    if P and N > P:
        assert blob != normal


def test_container_batches_equal_the_frozen_blob():
    m = _prelearned(14, P=2000, seed=9)
    data = datagen.markov2(2000 + 9000, 9).tobytes()
    blob = O.compress(data, 2048, 64, 100, 2000, 14, frozen=True)
    hdr = 21 + (int.from_bytes(blob[12:16], "big") if blob[20] & 1 else 2000)
    blen = int.from_bytes(blob[hdr:hdr + 4], "little")
    body = blob[hdr + 4:hdr + 4 + blen]
    assert body == O.encode_batch_frozen(data[2000:], 2048, m)
    assert O.decode_batch_frozen(body, 9000, 2048, m) == data[2000:]


def test_unknown_flag_bits_are_corrupt():
    blob = bytearray(O.compress(b"abc" * 100, 128, 32, 100, 0, 12, frozen=True))
    blob[20] = 4
    with pytest.raises(O.OracleError):
        O.decompress(bytes(blob))


def test_random_bytes_round_trip_with_prelearn():
    r = random.Random(5)
    for i in range(6):
        n = r.randrange(1, 6000)
        P = r.randrange(0, n + 1)
        data = bytes(r.getrandbits(8) for _ in range(n))
        blob = O.compress(data, r.choice([64, 1000, 4096]), 32 * r.randrange(1, 9),
                          r.randrange(1, 5), P, r.randrange(8, 20), frozen=True)
        assert O.decompress(blob) == data
