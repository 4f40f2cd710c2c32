"""Units above 256 bits (any multiple of 32 up to 65,504, as the reference's
16-bit U field allows) through the large-unit codec: containers
byte-identical to the oracle's and lossless, alone, with prelearned models,
with tail padding off, and in launches that mix them with ordinary units."""
import numpy as np
import pytest

from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu

CASES = [
    # (bytes, chunk, U, batch, prelearn, table bits, tail pad)
    (30000, 8192, 288, 100, 0, 16, True),
    (50000, 16384, 512, 2, 0, 18, True),
    (21000, 4096, 1024, 100, 3000, 14, True),
    (40000, 16384, 4096, 3, 0, 12, True),
    (9000, 1 << 20, 2048, 100, 0, 25, True),
    (12345, 4096, 65504, 100, 0, 10, False),
    (7000, 1000, 320, 100, 500, 9, False),
]


@pytest.mark.parametrize("n,S,U,batch,P,tb,pad", CASES)
def test_large_units_identical_to_oracle(ctx, oracle, n, S, U, batch, P, tb, pad):
    data = datagen.markov2(n, n + U).tobytes()
    p = make_params(S, U, batch, P, tb, tail_pad=pad)
    blob = ctx.compress(data, p)
    assert blob == oracle.compress(data, S, U, batch, P, tb, tail_pad=pad)
    assert ctx.stats()["tasks"] == (n - P + S - 1) // S
    assert ctx.decompress(blob) == data


def test_large_and_ordinary_units_in_one_launch(ctx, oracle):
    """Files with U of 128, 256, 512 and 2048 bits (and a frozen one) in one
    compress call and one decompress call: each container equals its own
    oracle container."""
    specs = [(128, 18, False), (512, 16, False), (256, 14, False), (2048, 12, False),
             (1024, 16, True), (96, 10, False), (4096, 18, False)]
    files, params = [], []
    for i, (U, tb, frozen) in enumerate(specs):
        data = datagen.markov2(6000 + 1500 * i, 40 + i).tobytes()
        files.append(data)
        params.append(make_params(4096, U, 2, 0, tb, frozen=frozen))
    blobs = ctx.compress_many(files, params)
    for f, p, b in zip(files, params, blobs):
        ref = oracle.compress(f, p.chunk_size, p.unit_bits, p.batch_chunks, 0, p.table_bits,
                              frozen=bool(p.flags & 2))
        assert b == ref, "U=%d" % p.unit_bits
    assert ctx.decompress_many(blobs) == files


def test_large_units_device_resident_many_chunks(ctx, oracle):
    """More chunks than the large-unit launch keeps resident waves for a small
    input each: every chunk through the device work queue."""
    rng = np.random.default_rng(3)
    files = [datagen.markov2(int(rng.integers(1, 9000)), 90 + i).tobytes() for i in range(40)]
    p = make_params(1024, 640, 100, 0, 12)
    blobs = ctx.compress_many(files, p)
    for i in (0, 7, 39):
        assert blobs[i] == oracle.compress(files[i], 1024, 640, 100, 0, 12)
    assert ctx.decompress_many(blobs) == files
