"""The team encoder and decoder (one chunk per workgroup of four waves,
cucm_team.h) on the GPU: containers byte-identical to the oracle's and
round trips lossless on every configuration and speculation width they
support, with a prelearned model, with more chunks than resident teams
(workgroups reuse their LDS and model slot from chunk to chunk), over many
containers in one launch, chosen automatically for a launch whose chunks
all fit, and the automatic speculation width."""
import numpy as np
import pytest

import cucm_amd
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
    if kind == "runs":  # runs of 1..39 equal bytes
        rng = np.random.default_rng(seed)
        return b"".join(bytes([int(rng.integers(0, 256))]) * int(rng.integers(1, 40))
                        for _ in range(n // 4 + 1))[:n]
    return bytes(n)


@pytest.fixture(scope="module")
def tctx():
    c = cucm_amd.Context()
    yield c
    c.close()


CASES = [
    # kind, N, S, U, batch, P, tb, spec
    ("markov2", 70000, 16384, 128, 100, 0, 18, 4),
    ("mixed", 40000, 4096, 64, 4, 0, 16, 2),
    ("random", 9000, 4096, 96, 100, 0, 14, 1),
    ("zeros", 20000, 2048, 256, 5, 0, 18, 4),
    ("structured", 30000, 8192, 128, 2, 0, 20, 4),
    ("markov2", 10, 4096, 128, 100, 0, 18, 4),
    ("markov2", 1, 1, 32, 100, 0, 8, 2),
    ("markov2", 60000, 8192, 128, 3, 7000, 16, 4),  # prelearned initial model
    ("mixed", 3000, 3000, 128, 1, 0, 25, 4),        # reference table size
    # 2^8 / 2^9 slots: bits of a byte, and of consecutive bytes, share slots
    ("random", 20000, 4096, 128, 100, 0, 8, 4),
    ("runs", 30000, 8192, 256, 100, 0, 9, 2),
    ("zeros", 12000, 4096, 64, 100, 0, 8, 4),
]


@pytest.mark.parametrize("kind,N,S,U,batch,P,tb,spec", CASES)
def test_team_codec_identical_and_lossless(tctx, oracle, kind, N, S, U, batch, P, tb, spec):
    data = _data(kind, N, 11)
    blob = oracle.compress(data, S, U, batch, P, tb)
    tctx.set_team(2)
    try:
        assert tctx.compress(data, make_params(S, U, batch, P, tb)) == blob
        assert tctx.decompress(blob, spec) == data
    finally:
        tctx.set_team(0)


def test_team_workgroups_cycle_through_many_chunks(tctx, oracle):
    data = _data("markov2", 300_000, 5)
    blob = oracle.compress(data, 4096, 128, 100, 0, 16)
    tctx.set_team(2)
    tctx.set_max_waves(8)  # two teams for 74 chunks
    try:
        assert tctx.compress(data, make_params(4096, 128, 100, 0, 16)) == blob
        assert tctx.stats()["waves"] == 8
        assert tctx.decompress(blob, 4) == data
        assert tctx.stats()["waves"] == 8
    finally:
        tctx.set_team(0)
        tctx.set_max_waves(0)


def test_team_many_containers_one_launch(tctx, oracle):
    datas = [_data("mixed", n, 100 + i) for i, n in enumerate((1, 700, 5000, 20000, 65536))]
    blobs = [oracle.compress(d, 4096, 64, 2, 0, 14) for d in datas]
    tctx.set_team(2)
    try:
        assert tctx.compress_many(datas, make_params(4096, 64, 2, 0, 14)) == blobs
        assert tctx.decompress_many(blobs, 2) == datas
    finally:
        tctx.set_team(0)


def test_auto_mode_runs_few_chunks_as_teams(tctx, oracle):
    data = _data("structured", 40000, 9)
    p = make_params(8192, 128, 100, 0, 16)
    blob = tctx.compress(data, p)
    assert tctx.stats()["waves"] == 20  # 5 chunks as teams
    assert blob == oracle.compress(data, 8192, 128, 100, 0, 16)
    tctx.set_team(1)
    try:
        assert tctx.compress(data, p) == blob
        assert tctx.stats()["waves"] == 5
    finally:
        tctx.set_team(0)
    for mode, spec in ((0, 4), (1, 4), (0, 8)):
        tctx.set_team(mode)
        try:
            assert tctx.decompress(blob, spec) == data
            # 5 chunks: teams of 4 waves in automatic mode with 1, 2 or 4
            # speculation bits; one wave per chunk otherwise
            want = 20 if (mode == 0 and spec <= 4) else 5
            assert tctx.stats()["waves"] == want
        finally:
            tctx.set_team(0)


def test_auto_speculation_width_follows_residency(tctx, oracle):
    """spec_bits 0: 4-bit rounds as teams while every chunk fits resident as
    a team; with more chunks (here 1,200 of 256 B against at most 2 teams
    per CU), one wave per chunk with 2-bit rounds.  Forced team modes keep
    their execution.  The output never changes."""
    few = _data("structured", 40000, 9)
    pf = make_params(8192, 128, 100, 0, 16)
    bf = tctx.compress(few, pf)
    assert tctx.decompress(bf, 0) == few
    st = tctx.stats()
    assert (st["spec_bits"], st["waves"]) == (4, 20)
    many = _data("markov2", 1200 * 256, 12)
    pm = make_params(256, 128, 100, 0, 10)
    bm = tctx.compress(many, pm)
    assert bm == oracle.compress(many, 256, 128, 100, 0, 10)
    assert tctx.decompress(bm, 0) == many
    st = tctx.stats()
    assert st["spec_bits"] == 2 and st["teams"] == 0 and st["tasks"] == 1200
    for mode, want in ((1, (2, 0)), (2, (4, 1))):
        tctx.set_team(mode)
        try:
            assert tctx.decompress(bm, 0) == many
            st = tctx.stats()
            assert (st["spec_bits"], min(st["teams"], 1)) == want
        finally:
            tctx.set_team(0)


def test_fixed_waves_per_block_plans_only_resident_workgroups(tctx, oracle):
    """cucm_ctx_set_waves_per_block: the launch is sized by how many
    workgroups of that size the CUs' LDS holds at once, so every planned
    workgroup is resident (no model slot is reserved for a workgroup that
    could only start once the queue is empty).  One wave per chunk, 2-bit
    rounds, 2,400 chunks: seven waves per workgroup fit one workgroup per
    CU, and so do six, so six-wave launches plan no more workgroups than
    seven-wave ones."""
    many = _data("markov2", 2400 * 256, 13)
    pm = make_params(256, 128, 100, 0, 10)
    bm = tctx.compress(many, pm)
    assert bm == oracle.compress(many, 256, 128, 100, 0, 10)
    tctx.set_team(1)
    try:
        shapes = {}
        for wpb in (7, 6):
            tctx.set_waves_per_block(wpb)
            assert tctx.decompress(bm, 2) == many
            st = tctx.stats()
            assert st["waves"] % wpb == 0
            shapes[wpb] = st["waves"] // wpb
        assert shapes[6] <= shapes[7], shapes
    finally:
        tctx.set_waves_per_block(0)
        tctx.set_team(0)


def test_many_pass_encoder_teams_are_capped(tctx, oracle):
    """An encoder launch with four or more full passes of teams keeps at most
    9 teams per 4 CUs resident by default (fewer than LDS allows: measured
    faster); an explicit resident-wave cap replaces that default.  The
    output never changes."""
    many = _data("markov2", 4000 * 256, 14)
    pm = make_params(256, 128, 100, 0, 10)
    ref = oracle.compress(many, 256, 128, 100, 0, 10)
    tctx.set_team(2)
    try:
        assert tctx.compress(many, pm) == ref
        capped = tctx.stats()["teams"]
        tctx.set_max_waves(4 * 4000)
        assert tctx.compress(many, pm) == ref
        full = tctx.stats()["teams"]
    finally:
        tctx.set_max_waves(0)
        tctx.set_team(0)
    assert 0 < capped < full and capped * 4 <= full * 3, (capped, full)
    assert tctx.decompress(ref) == many
