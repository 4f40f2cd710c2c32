"""The allocation-free entry point (cucm_compress_device_noalloc) through the
C ABI: its containers equal cucm_compress_many_device's byte for byte, on
consecutive calls that reuse the device scratch (the second smaller than the
first, so stale scratch contents would show), and it refuses what its fixed
scratch cannot hold."""
import numpy as np
import pytest

import cucm_amd
from cucm_amd import datagen, make_params
from cucm_amd.codec import output_cap

pytestmark = pytest.mark.gpu


def _both(ctx, data, p):
    n = len(data)
    d_in = ctx.alloc(max(n, 1))
    if n:
        d_in.upload(data)
    cap = output_cap(n, p)
    d_a = ctx.alloc(cap)
    d_b = ctx.alloc(cap)
    lens, st = ctx.compress_many_device([d_in.addr], [n], p, [d_a.addr], [cap])
    assert st[0] == 0
    m = ctx.compress_device_noalloc(d_in.addr, n, p, d_b.addr, cap)
    assert m == lens[0]
    return d_a.download(m), d_b.download(m)


CASES = [
    # (bytes, chunk, U, batch, table bits, frozen, tail pad), each followed by a smaller one
    [(3 << 20, 16384, 128, 100, 18, False, True), (100000, 4096, 64, 7, 16, False, True)],
    [(2 << 20, 1 << 20, 128, 100, 22, False, True), (300000, 65536, 256, 2, 20, False, False)],
    [(1 << 20, 65536, 128, 3, 18, True, True), (5000, 1024, 96, 100, 12, True, True)],
]


@pytest.mark.parametrize("pair", CASES, ids=["small-cfg", "tb22", "frozen"])
def test_noalloc_equals_many_device_twice(ctx, oracle, pair):
    for i, (n, S, U, batch, tb, frozen, pad) in enumerate(pair):
        data = datagen.markov2(n, 50 + i + n % 97).tobytes()
        p = make_params(S, U, batch, 0, tb, tail_pad=pad, frozen=frozen)
        a, b = _both(ctx, data, p)
        assert a == b, (i, n)
        if i == 1:
            assert b == oracle.compress(data, S, U, batch, 0, tb, tail_pad=pad, frozen=frozen)
        assert ctx.decompress(b) == data


def test_noalloc_empty_input(ctx):
    p = make_params(4096, 128, 100, 0, 18)
    d_out = ctx.alloc(64)
    d_in = ctx.alloc(16)
    assert ctx.compress_device_noalloc(d_in.addr, 0, p, d_out.addr, 64) == 21
    assert d_out.download(21) == ctx.compress(b"", p)


def test_noalloc_refusals(ctx):
    p = make_params(4096, 128, 100, 100, 18)  # prelearn needs device memory
    d = ctx.alloc(1 << 20)
    o = ctx.alloc(2 << 20)
    with pytest.raises(cucm_amd.CucmError) as e:
        ctx.compress_device_noalloc(d.addr, 1 << 20, p, o.addr, 2 << 20)
    assert e.value.code == -7
    big = ctx.alloc(48 << 20)
    with pytest.raises(cucm_amd.CucmError) as e:  # staging beyond the scratch
        ctx.compress_device_noalloc(big.addr, 48 << 20, make_params(1 << 20, 128, 100, 0, 18),
                                    o.addr, 2 << 20)
    assert e.value.code == -5
    with pytest.raises(cucm_amd.CucmError) as e:  # output too small
        data = datagen.markov2(50000, 3).tobytes()
        d.upload(data)
        ctx.compress_device_noalloc(d.addr, len(data), make_params(4096, 128, 100, 0, 16),
                                    o.addr, 100)
    assert e.value.code == -3
    # still correct after the refusals
    data = datagen.markov2(70000, 4).tobytes()
    a, b = _both(ctx, data, make_params(8192, 128, 100, 0, 18))
    assert a == b
