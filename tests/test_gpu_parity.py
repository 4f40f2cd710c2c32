"""GPU parity: the HIP codec's containers are byte-identical to the oracle's
(the CPU restatement of the reference model run with the same chunking) and
every round trip is lossless.  All calls go through libcucm.so's C ABI."""
import numpy as np
import pytest

from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu


def _data(kind, n, seed):
    if kind == "markov2":
        return datagen.markov2(n, seed).tobytes()
    if kind == "mixed":
        return datagen.mixed(n, seed).tobytes()
    if kind == "structured":
        return datagen.structured(n, seed).tobytes()
    if kind == "random":
        return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8).tobytes()
    if kind == "zeros":
        return bytes(n)
    raise ValueError(kind)


CASES = [
    # kind, N, S, U, batch, tb
    ("markov2", 70000, 16384, 128, 100, 18),
    ("markov2", 5000, 1024, 32, 3, 12),
    ("mixed", 40000, 4096, 64, 4, 16),
    ("random", 9000, 4096, 96, 100, 14),
    ("zeros", 20000, 2048, 256, 5, 18),
    ("structured", 30000, 8192, 128, 2, 20),
    ("markov2", 10, 4096, 128, 100, 18),      # tail padding dominates
    ("markov2", 1, 1, 32, 100, 8),
    ("markov2", 3000, 3000, 128, 1, 25),      # one chunk, reference table size
]


@pytest.mark.parametrize("kind,N,S,U,batch,tb", CASES)
def test_container_identical_and_lossless(ctx, oracle, kind, N, S, U, batch, tb):
    data = _data(kind, N, 7)
    ref = oracle.compress(data, S, U, batch, 0, tb)
    got = ctx.compress(data, make_params(S, U, batch, 0, tb))
    assert len(got) == len(ref)
    assert got == ref
    back = ctx.decompress(got)
    assert back == data


@pytest.mark.parametrize("spec", [1, 2, 4, 8])
def test_speculation_width_does_not_change_output(ctx, oracle, spec):
    data = _data("mixed", 20000, 3)
    blob = oracle.compress(data, 4096, 128, 100, 0, 16)
    assert ctx.decompress(blob, spec) == data


def test_small_config(ctx, oracle):
    """1 MiB order-2 Markov source, 64 x 16 KiB chunks, 2^18 slots per order."""
    data = _data("markov2", 1 << 20, 1)
    ref = oracle.compress(data, 16384, 128, 100, 0, 18)
    got = ctx.compress(data, make_params(16384, 128, 100, 0, 18))
    assert got == ref
    assert ctx.decompress(got) == data


@pytest.mark.parametrize("P", [0, 1 << 20], ids=["P0", "P1M"])
def test_reference_default_configuration(ctx, oracle, P):
    """The reference command line's defaults (cucm_compress_v2.cu:1288-1366)
    with its hard-coded 2^25 slots per order (statemap_size,
    cucm_compress_v2.cu:553): S = 1 MiB, U = 128, batch 100, with and without
    a 1 MiB prelearn segment (-p 1M).  3 MiB + 4321 bytes of the mixed
    stream, so the last chunk is ragged and tail-padded; the container is
    the oracle's byte for byte and decodes at the automatic and a fixed
    speculation width."""
    data = _data("mixed", 3 * (1 << 20) + 4321, 91)
    ref = oracle.compress(data, 1 << 20, 128, 100, P, 25)
    got = ctx.compress(data, make_params(1 << 20, 128, 100, P, 25))
    assert got == ref
    assert ctx.decompress(got) == data
    assert ctx.decompress(got, 2) == data


def test_empty_input(ctx, oracle):
    ref = oracle.compress(b"", 4096, 128, 100, 0, 18)
    got = ctx.compress(b"", make_params(4096, 128, 100, 0, 18))
    assert got == ref and len(got) == 21
    assert ctx.decompress(got) == b""
