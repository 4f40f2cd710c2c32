"""Known-answer and structural checks that pin the oracle's primitives.

Sources of the expected values:
  - FNV-1a-32 published test vectors ("", "a", "foobar");
  - the closed forms of squash/stretch/dt at cucm_compress_v2.cu:681-705
    evaluated at points whose values are fixed by the formula's structure
    (endpoints, the midpoint, the inverse relation);
  - the first StateMap update and first coder step computed by hand from
    cucm_compress_v2.cu:592-601 and :855-862;
  - the state table checked against all four copies in the reference
    sources when they are present (cucm_compress_v2.cu:35-285, :287-537,
    cucm_decompress_v2.cu:16-265, :267-517) and against a recorded digest.
"""
import hashlib
import os
import re

import pytest

from oracle import oracle as O

NEX_SHA256 = "33e19144bc4f54d211c34a695655d391fdf04eb9922149d32d72ed676d98966e"
REF = "/root/reference"


@pytest.fixture(scope="module")
def tables():
    return O.tables()


def test_fnv1a32_published_vectors():
    assert O.fnv1a32(b"") == 0x811C9DC5
    assert O.fnv1a32(b"a") == 0xE40C292C
    assert O.fnv1a32(b"foobar") == 0xBF9CF968


def test_squash_endpoints_and_monotone(tables):
    sq, st, dt, nex = tables
    assert (sq[0], sq[2047], sq[4094], sq[4095]) == (1, 2047, 4094, 4095)
    assert all(sq[i] <= sq[i + 1] for i in range(4095))
    # knots are hit exactly at multiples of 128 from -2047+1..: d = k - 2047
    knots = [1, 2, 3, 6, 10, 16, 27, 45, 73, 120, 194, 310, 488, 747, 1101, 1546, 2047, 2549,
             2994, 3348, 3607, 3785, 3901, 3975, 4022, 4050, 4068, 4079, 4085, 4089, 4092, 4093,
             4094]
    for i in range(1, 32):
        assert sq[2047 + (i - 16) * 128] == knots[i]


def test_stretch_is_inverse_of_squash(tables):
    sq, st, dt, nex = tables
    assert (st[0], st[2047], st[2048], st[4095]) == (-2047, 0, 1, 2047)
    assert all(st[i] <= st[i + 1] for i in range(4095))
    for x in range(-2047, 2048):
        p = sq[x + 2047]
        # stretch(p) is the smallest x' with squash(x') >= p
        assert st[p] <= x
        assert sq[st[p] + 2047] >= p


def test_dt_values(tables):
    sq, st, dt, nex = tables
    assert dt[:4] == [5461, 3276, 2340, 1820] and dt[1023] == 7
    assert all(d * (2 * n + 3) <= 16384 < (d + 1) * (2 * n + 3) for n, d in enumerate(dt))


def test_state_table_digest_and_invariants(tables):
    nex = tables[3]
    assert hashlib.sha256(nex).hexdigest() == NEX_SHA256
    assert (nex[0], nex[1]) == (1, 2)
    assert max(nex) == 252
    assert nex[2 * 253:] == bytes(6)


def test_state_table_matches_product_copy(tables):
    nex = tables[3]
    src = open(os.path.join(os.path.dirname(__file__), "..", "cucm_amd", "csrc", "cucm_model.h")).read()
    body = src[src.index("kNextState[256] = {"):]
    words = [int(w, 16) for w in re.findall(r"0x([0-9a-f]{4})", body[: body.index("};")])]
    assert len(words) == 256
    assert bytes(b for w in words for b in (w & 0xFF, w >> 8)) == nex


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_state_table_matches_all_reference_copies(tables):
    nex = tables[3]
    copies = []
    for fn in ("cucm_compress_v2.cu", "cucm_decompress_v2.cu"):
        text = open(os.path.join(REF, fn)).read()
        for m in re.finditer(r"[sS]tate_table\[256\]\[2\]\s*=\s*\{(.*?)\};", text, re.S):
            body = re.sub(r"//[^\n]*", "", m.group(1))
            copies.append(bytes(int(x) for x in re.findall(r"\d+", body)))
    assert len(copies) == 4
    assert all(c == nex for c in copies)


def test_first_bit_slots_at_25_bits():
    # chunk start: h0 = 1, previous bytes all zero
    assert O.slots(b"\x00" * 8, 0, 25) == [809868, 24386916, 30033004, 23705092, 23693388,
                                          18976676, 30744364, 8393284]


def test_slots_follow_fnv_prefix_chain():
    chunk = bytes([0x41, 0x42, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0xA5])
    b = 8 * 9 + 3  # bit 3 of byte 9: h0 = 0b1000 | 0xA5 >> 5
    h = [0b1000 | (0xA5 >> 5)] + [chunk[9 - i] for i in range(1, 8)]
    for k, s in enumerate(O.slots(chunk, b, 22)):
        assert s == O.fnv1a32(bytes(h[: k + 1])) & ((1 << 22) - 1)


def test_order0_slots_are_distinct_for_all_table_sizes():
    for tb in range(8, 26):
        s = {O.fnv1a32(bytes([h0])) & ((1 << tb) - 1) for h0 in range(1, 256)}
        assert len(s) == 255


def test_first_statemap_update():
    m = O.Model(25)
    m.sm_update(0, 5, 1)
    assert m.statep(0, 0) == 0xD5540001 and m.state(0, 5) == 2
    m = O.Model(25)
    m.sm_update(3, 77, 0)
    assert m.statep(3, 0) == 0x2AAC0001 and m.state(3, 77) == 1
    # an untouched slot is in state 0, whose statep the update just moved
    assert m.predict(3, 78) == 0x2AAC0001 >> 20 == 682
    assert m.predict(2, 78) == 2048


def test_zero_weight_mixer_and_update():
    m = O.Model(12)
    assert m.mix(0, [100, -50, 7, 0, 1, 2, 3, 4]) == 2047  # dot 0 -> squash(0)
    m.mix_update(1, 2048, 5, [2047, -2047, 1, 0, 100, -100, 16, -16])
    err = ((1 << 12) - 2048) * 7
    exp = [(err * s) >> 16 for s in [2047, -2047, 1, 0, 100, -100, 16, -16]]
    assert [m.weight(5 * 8 + i) for i in range(8)] == exp
    assert all(m.weight(i) == 0 for i in range(640) if not 40 <= i < 48)


def test_first_coder_step():
    assert O.coder_step(0, 0xFFFFFFFF, 2048, 1) == (0, 0x7FFFFFFF, b"")
    assert O.coder_step(0, 0xFFFFFFFF, 2048, 0) == (0x80000000, 0xFFFFFFFF, b"")
    # a certain bit narrows the interval until the top byte is settled
    x1, x2, out = 0, 0xFFFFFFFF, b""
    for _ in range(12):
        x1, x2, o = O.coder_step(x1, x2, 1, 1)
        out += o
    assert len(out) >= 1 and out[0] == 0
