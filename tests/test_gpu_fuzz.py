"""Edge sizes and seeded random configurations on the GPU, each container
byte-identical to the oracle's and decoded back losslessly.

Edge sizes (chunk size S = 1000, U = 96 so U does not divide 8S): empty,
1 byte, one unit less a byte, one unit, S - 1, S, S + 1, 2S - 1, 2S, 2S + 1,
as single waves, as teams and in automatic mode.  The fuzz draws input size
(log-uniform up to 300 KB), chunk size, U, batch size, prelearn size, table
bits, frozen model, tail padding, speculation width, team mode and data
kind from a fixed seed."""
import math
import random

import numpy as np
import pytest

from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu


def _data(kind, n, seed):
    if n == 0:
        return b""
    if kind == "markov2":
        return datagen.markov2(n, seed).tobytes()
    if kind == "mixed":
        return datagen.mixed(n, seed).tobytes()
    if kind == "structured":
        return datagen.structured(n, seed).tobytes()
    if kind == "zeros":
        return bytes(n)
    return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8).tobytes()


S_EDGE, U_EDGE = 1000, 96
EDGE_SIZES = [0, 1, U_EDGE // 8 - 1, U_EDGE // 8, S_EDGE - 1, S_EDGE, S_EDGE + 1,
              2 * S_EDGE - 1, 2 * S_EDGE, 2 * S_EDGE + 1]


@pytest.mark.parametrize("team", [0, 1, 2])
def test_edge_sizes(ctx, oracle, team):
    ctx.set_team(team)
    try:
        for n in EDGE_SIZES:
            data = _data("markov2", n, 60 + n)
            ref = oracle.compress(data, S_EDGE, U_EDGE, 2, 0, 12)
            got = ctx.compress(data, make_params(S_EDGE, U_EDGE, 2, 0, 12))
            assert got == ref, "n=%d" % n
            assert ctx.decompress(got) == data, "n=%d" % n
    finally:
        ctx.set_team(0)


def _draw(r):
    n = 0 if r.random() < 0.05 else int(math.exp(r.uniform(0, math.log(300_000))))
    S = r.choice([64, 100, 512, 1000, 4096, 8192, 16384, 65536])
    U = 32 * r.randint(1, 8)
    batch = r.choice([1, 2, 3, 7, 100, 200])
    P = 0 if r.random() < 0.6 else r.randint(0, n)
    tb = r.choice([8, 9, 12, 14, 16, 18, 20, 22]) if n > 5000 else r.choice([8, 12, 16, 25])
    frozen = r.random() < 0.3
    pad = r.random() < 0.8
    spec = r.choice([1, 2, 4, 8])
    team = r.choice([0, 1, 2])
    kind = r.choice(["markov2", "mixed", "structured", "zeros", "random"])
    return n, S, U, batch, P, tb, frozen, pad, spec, team, kind


def test_random_configurations(ctx, oracle):
    r = random.Random(20261203)
    for case in range(40):
        n, S, U, batch, P, tb, frozen, pad, spec, team, kind = _draw(r)
        what = "case %d: n=%d S=%d U=%d batch=%d P=%d tb=%d frozen=%s pad=%s spec=%d team=%d %s" % (
            case, n, S, U, batch, P, tb, frozen, pad, spec, team, kind)
        data = _data(kind, n, case)
        ref = oracle.compress(data, S, U, batch, P, tb, tail_pad=pad, frozen=frozen)
        ctx.set_team(team)
        try:
            got = ctx.compress(data, make_params(S, U, batch, P, tb, spec, pad, frozen))
            ok = got == ref
            assert ok, what
            back = ctx.decompress(got, spec)
            ok = back == data
            assert ok, what
        finally:
            ctx.set_team(0)
