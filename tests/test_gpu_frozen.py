"""The frozen-model codec (cucm_frozen.h, CUCM_FLAG_FROZEN) on the GPU:
containers byte-identical to the oracle's frozen containers and lossless at
every speculation width, with the zero model (whose payload is the chunk
complemented), raw and coded prelearn segments, more chunks than resident
waves, frozen and normal jobs in one launch, and the streaming API."""
import numpy as np
import pytest

import cucm_amd
from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu


def same(a, b):
    """Equality as a plain bool (a failing assert on large byte strings would
    otherwise spend minutes diffing them)."""
    return a == b


@pytest.fixture(scope="module")
def fctx():
    c = cucm_amd.Context()
    yield c
    c.close()


def _data(kind, n, seed):
    if kind == "markov2":
        return datagen.markov2(n, seed).tobytes()
    if kind == "mixed":
        return datagen.mixed(n, seed).tobytes()
    if kind == "structured":
        return datagen.structured(n, seed).tobytes()
    return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8).tobytes()


CASES = [
    # kind, N, S, U, batch, P, tb
    ("markov2", 90000, 16384, 128, 100, 20000, 18),
    ("mixed", 70000, 4096, 64, 4, 9000, 16),
    ("structured", 50000, 8192, 256, 2, 30000, 20),
    ("random", 9000, 4096, 96, 100, 3000, 14),
    ("markov2", 30000, 1000, 32, 7, 0, 12),
    ("markov2", 5001, 5000, 128, 100, 5000, 18),
    ("markov2", 7, 4096, 128, 100, 3, 10),
    ("mixed", 4000, 3000, 128, 1, 1000, 25),
]


@pytest.mark.parametrize("kind,N,S,U,batch,P,tb", CASES)
def test_frozen_identical_to_oracle_and_lossless(fctx, oracle, kind, N, S, U, batch, P, tb):
    data = _data(kind, N, 13)
    blob = oracle.compress(data, S, U, batch, P, tb, frozen=True)
    got = fctx.compress(data, make_params(S, U, batch, P, tb, frozen=True))
    assert same(got, blob)
    for spec in (1, 2, 4, 8):
        assert same(fctx.decompress(blob, spec), data), "spec %d" % spec


def test_zero_model_payload_is_the_complement(fctx):
    d = np.random.default_rng(4).integers(0, 256, 3000, dtype=np.uint8)
    blob = fctx.compress(d.tobytes(), make_params(4096, 128, 100, 0, 16, frozen=True))
    assert blob[20] == 2
    # one batch: LE length, one BE end offset, the payload
    assert blob[25:29] == (4 + 3001 - 1).to_bytes(4, "big")
    assert same(blob[29:], (~d).tobytes() + b"\x00")


def test_more_chunks_than_waves(fctx, oracle):
    data = _data("markov2", 400_000, 8)
    blob = oracle.compress(data, 2048, 128, 100, 10000, 14, frozen=True)
    fctx.set_max_waves(5)
    try:
        assert same(fctx.compress(data, make_params(2048, 128, 100, 10000, 14, frozen=True)), blob)
        assert fctx.stats()["waves"] == 5
        assert same(fctx.decompress(blob, 4), data)
    finally:
        fctx.set_max_waves(0)


def test_frozen_and_normal_jobs_in_one_launch(fctx, oracle):
    datas = [_data("mixed", n, 50 + i) for i, n in enumerate((5000, 20000, 1, 65536, 30000, 7000))]
    frozen = [True, False, True, False, True, False]
    P = [1000, 0, 0, 2000, 0, 3000]
    blobs = [oracle.compress(d, 4096, 64, 2, p, 14, frozen=f) for d, p, f in zip(datas, P, frozen)]
    got = []
    for d, p, f in zip(datas, P, frozen):
        got.append(fctx.compress(d, make_params(4096, 64, 2, p, 14, frozen=f)))
    assert same(got, blobs)
    # one decode launch pair over all six containers
    assert same(fctx.decompress_many(blobs, 4), datas)
    assert fctx.stats()["launches"] == 2


def test_frozen_stream_equals_one_shot(fctx, oracle):
    data = _data("markov2", 200_000, 31)
    p = make_params(8192, 128, 3, 15000, 16, frozen=True)
    one = fctx.compress(data, p)
    assert same(one, oracle.compress(data, 8192, 128, 3, 15000, 16, frozen=True))
    assert same(fctx.compress_stream(data, p, segment_batches=2), one)
    assert same(fctx.decompress_stream(one, 4, records_per_call=2), data)
    orig, hp = cucm_amd.peek_header(one)
    assert orig == len(data) and hp.flags & 2


def test_expanding_chunks_rerun_with_the_strict_bound(fctx, oracle):
    """Chunks whose payload outgrows the 1.5 S + 256 staging region go
    through the second launch into the strict-bound region: a model
    prelearned on zeros pays ~4 bytes per 0x1f byte (its mixer trusts states
    that predict 0).  Expanding and ordinary chunks, and a normal job, in one
    call; containers equal the oracle's and decode losslessly."""
    S, P = 4096, 5000
    body = (datagen.markov2(3 * S, 3).tobytes() + b"\x1f" * (2 * S)
            + datagen.markov2(S + 100, 4).tobytes())
    data = bytes(P) + body
    model = oracle.Model(14)
    oracle.prelearn_train(bytes(P), model)
    assert len(oracle.encode_chunk_frozen(body[3 * S:4 * S], model)) > S + S // 2 + 256
    blob = oracle.compress(data, S, 128, 2, P, 14, frozen=True)
    other = datagen.mixed(30000, 6).tobytes()
    oblob = oracle.compress(other, S, 128, 100, 0, 14)
    got = fctx.compress_many([data, other], [make_params(S, 128, 2, P, 14, frozen=True),
                                             make_params(S, 128, 100, 0, 14)])
    assert same(got, [blob, oblob])
    for spec in (1, 4, 8):
        assert same(fctx.decompress(blob, spec), data)
