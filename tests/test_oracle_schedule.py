"""Pins of the oracle's unit schedule, stale-slot tail padding and
zero-initialised staging, from mathematics rather than from the oracle's
own code paths.

The reference (cucm_compress_v2.cu:1067-1090) freezes the model for a unit of
U bits: every prediction of the unit reads the model as it stood at the unit
start, and the updates are applied after the unit, StateMaps in bit order and
mixer deltas summed.  A partial last chunk is coded up to the nominal chunk
size with the (pr, y) left in the per-slot staging buffer `prm`
(:764-767, :850-851), which is zeroed once per batch (:1039-1040).  Nothing
from the reference can run here, so these facts are fixed by:

  * closed forms: the zero model predicts pr = 2048 for every bit of a unit
    that starts from it, which halves the coder interval exactly, so such a
    unit codes its bytes complemented; a zeroed staging slot codes
    (pr, y) = (0, 0), whose step is x1 += 1;
  * a hand-derived two-unit known answer on all-zero input whose second unit
    depends on the first unit's deferred updates;
  * the stale-slot rule checked on a per-bit trace: padding bit b codes the
    (pr, y) of the last data bit b' < 8L with b' = b (mod U), or (0, 0) when
    there is none;
  * a Python restatement of the unit schedule composed only of primitives
    that their own known-answer tests pin (slots, StateMap predict/update,
    stretch, mixer predict/update, coder step), run on small inputs.
A mistake such as updating immediately instead of at the unit end, applying
StateMap updates out of bit order, a stale slot taken from the wrong unit, or
staging that is not zeroed changes bytes in at least one of these tests.
"""
import numpy as np
import pytest

from cucm_amd import datagen
from oracle import oracle as O

SQ, ST, DT, NEX = O.tables()


def payloads(blob, C):
    ends = [int.from_bytes(blob[4 * c:4 * c + 4], "big") for c in range(C)]
    out, prev = [], 4 * C - 1
    for e in ends:
        out.append(blob[prev + 1:e + 1])
        prev = e
    return out


def code(trace, x1=0, x2=0xFFFFFFFF):
    """Runs the KAT-pinned coder step over (pr, y) pairs; returns the emitted
    bytes and the final interval."""
    out = bytearray()
    for pr, y in trace:
        x1, x2, e = O.coder_step(x1, x2, pr, y)
        out += e
    return bytes(out), x1, x2


def bits_of(data):
    return [(data[b >> 3] >> (7 - (b & 7))) & 1 for b in range(8 * len(data))]


def data_trace(chunk, U, model):
    """(pr, y) of every data bit under the unit schedule, restated from the
    pinned primitives: predictions of a unit from the unit-start model, then
    mixer updates from the recorded (pr, cxt, smp), then each order's StateMap
    updates in bit order (cucm_compress_v2.cu:1067-1090)."""
    m = model.copy()
    tb = m.tb
    nbits = 8 * len(chunk)
    ys = bits_of(chunk)
    trace = []
    for base in range(0, nbits, U):
        recs = []
        for b in range(base, min(base + U, nbits)):
            sl = O.slots(chunk, b, tb)
            smp = [ST[m.predict(k, sl[k])] for k in range(8)]
            c1 = chunk[(b >> 3) - 1] if b >= 8 else 0
            cxt = (c1 >> 5) * 10 + sum(1 for v in smp if v != 1)
            pr = m.mix(cxt, smp)
            pr += 1 if pr < 2048 else 0
            recs.append((sl, smp, cxt, pr, ys[b]))
            trace.append((pr, ys[b]))
        for sl, smp, cxt, pr, y in recs:
            m.mix_update(y, pr, cxt, smp)
        for k in range(8):
            for sl, smp, cxt, pr, y in recs:
                m.sm_update(k, sl[k], y)
    return trace


def padded_trace(dtrace, L, S, U):
    """The coded (pr, y) sequence of a tail-padded chunk: data bits, then for
    each padding bit b in [8L, 8S) the staging slot r = b mod U, which holds
    the last data bit written to it (b' = b - U*k < 8L) or zero."""
    out = list(dtrace)
    for b in range(8 * L, 8 * S):
        r = b % U
        last = r + ((8 * L - 1 - r) // U) * U if r < 8 * L else None
        out.append(dtrace[last] if last is not None else (0, 0))
    return out


# ---------------------------------------------------------------------------
# (a) closed forms of the zero model
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("N,S,U", [(5 * 12 + 7, 12, 128), (3 * 32, 32, 256), (1000, 100, 800),
                                   (4 * 8 + 1, 8, 64)])
def test_zero_model_single_unit_chunks_are_the_complement(N, S, U):
    """U >= 8S: every chunk of the batch is one unit predicted from the zero
    model.  A fresh state has statep = 2^31, p = 2048, stretch(2048) = 1, so
    no order is active and the zero-weight dot is 0; squash(0) = 2047 and the
    bump gives pr = 2048 (cucm_compress_v2.cu:558-566, 636-644, 786-788).  For
    r = x2 - x1 = 4096a + c the split is x1 + 2048a + floor(2048c/4096) =
    x1 + floor(r/2): pr = 2048 halves [0, 2^32 - 1] exactly, y = 1 keeping the
    lower half, so 8 bits emit the complemented byte and restore
    (0, 2^32 - 1); the flush is x1 >> 24 = 0.  Padding of the last chunk codes
    zeroed slots, (pr, y) = (0, 0): xmid = x1, x1 = xmid + 1, so x1 ends at
    8(S - L) < 2^24 and the flush stays 0."""
    data = np.random.default_rng(N).integers(0, 256, N, dtype=np.uint8)
    C = (N + S - 1) // S
    for pad in (True, False):
        blob = O.encode_batch(data.tobytes(), S, U, O.Model(12), tail_pad=pad)
        for c, pay in enumerate(payloads(blob, C)):
            chunk = data[c * S:(c + 1) * S]
            assert pay == (~chunk).tobytes() + b"\x00", (c, pad)


@pytest.mark.parametrize("S,L,tb", [(64, 64, 14), (256, 200, 16), (1000, 999, 12)])
def test_prelearned_single_unit_chunks_equal_the_frozen_coder(S, L, tb):
    """U >= 8S: each chunk is one unit, so every bit is predicted from the
    initial (prelearned) model with no update visible -- the frozen coder's
    definition, pinned by its own closed form and twin.  Non-last chunks have
    no padding; the last one is compared without it."""
    m = O.Model(tb)
    O.prelearn_train(datagen.markov2(5000, 31).tobytes(), m)
    U = (8 * S + 31) // 32 * 32
    data = datagen.markov2(3 * S + L, 32).tobytes()
    C = 4
    blob = O.encode_batch(data, S, U, m, tail_pad=False)
    for c, pay in enumerate(payloads(blob, C)):
        assert pay == O.encode_chunk_frozen(data[c * S:(c + 1) * S], m), c


# ---------------------------------------------------------------------------
# (b) zeroed staging: padding that codes (pr, y) = (0, 0)
# ---------------------------------------------------------------------------

def _check_zero_slot_padding(chunk, L, S, U, m):
    """Padded payload = data bits, then 8(S - L) steps of (0, 0) through the
    pinned coder step, then the flush.  When x1_end + 8(S - L) keeps its top
    byte below x2's, no step emits and the closed form holds: the padded
    payload is the unpadded one with the flush (x1_end + 8(S - L)) >> 24."""
    plain = O.encode_chunk(chunk, S, U, m, tail_pad=False, L=L)
    padded = O.encode_chunk(chunk, S, U, m, tail_pad=True, L=L)
    emitted, x1, x2 = code(data_trace(chunk, U, m))
    assert plain == emitted + bytes([x1 >> 24])
    pad, px1, _ = code([(0, 0)] * (8 * (S - L)), x1, x2)
    assert padded == emitted + pad + bytes([px1 >> 24])
    if ((x1 + 8 * (S - L)) >> 24) < (x2 >> 24):
        assert pad == b"" and px1 == x1 + 8 * (S - L)
        assert padded == plain[:-1] + bytes([(x1 + 8 * (S - L)) >> 24])
    return padded != plain


@pytest.mark.parametrize("L,S,U", [(10, 32, 256), (3, 1024, 8192), (7, 200, 2048)])
def test_zeroed_slots_pad_with_x1_increments(L, S, U):
    """8S <= U: every padding bit sits in the chunk's only unit at a slot no
    data bit wrote, so it codes (0, 0), i.e. x1 += 1, emitting nothing while
    x1's top byte stays below x2's.  The padded payload is therefore the
    unpadded payload without its flush byte, followed by the flush
    (x1_end + 8(S - L)) >> 24, where x1_end is the interval after the data
    bits (taken from the data trace run through the pinned coder step)."""
    m = O.Model(14)
    O.prelearn_train(datagen.markov2(4000, 33).tobytes(), m)
    for seed in range(6):
        _check_zero_slot_padding(datagen.markov2(L, 100 + seed).tobytes(), L, S, U, m)


def test_zeroed_slot_padding_carries_into_the_flush():
    """The same rule where it is visible: 8(S - L) = 65,496 padding
    increments (S = 8188, U = 65,504, the largest unit) carry into x1's top
    byte when x1_end's low 24 bits lie within 65,496 of 2^24.  Inputs are
    searched (by the padded and unpadded payloads differing) until two such
    cases are found; the rule must hold for them and for the others."""
    L, S, U = 1, 8188, 65504
    m = O.Model(12)
    O.prelearn_train(datagen.markov2(3000, 39).tobytes(), m)
    carried, checked = 0, 0
    for seed in range(6000):
        chunk = datagen.markov2(L, 1000 + seed).tobytes()
        plain = O.encode_chunk(chunk, S, U, m, tail_pad=False, L=L)
        padded = O.encode_chunk(chunk, S, U, m, tail_pad=True, L=L)
        if padded != plain or checked < 3:
            assert _check_zero_slot_padding(chunk, L, S, U, m) == (padded != plain)
            checked += 1
            carried += padded != plain
        if carried >= 2:
            break
    assert carried >= 2


# ---------------------------------------------------------------------------
# (c) stale slots: padding that repeats the last unit's data bits
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("L,S,U,tb", [(40, 100, 64, 14), (37, 64, 96, 12), (16, 200, 32, 16),
                                      (129, 300, 256, 13), (5, 20, 128, 10)])
def test_padding_repeats_the_last_data_bit_of_each_slot(L, S, U, tb):
    """Padding bit b in [8L, 8S) codes staging slot r = b mod U, which keeps
    the (pr, y) of the last data bit b' = b - U*k < 8L written to it, or the
    zero it was initialised with when no data bit used slot r (8L <= r).  The
    whole padded payload must equal that sequence through the coder."""
    m = O.Model(tb)
    O.prelearn_train(datagen.markov2(3000, 34).tobytes(), m)
    chunk = datagen.markov2(L, 35 + L).tobytes()
    dtrace = data_trace(chunk, U, m)
    full = padded_trace(dtrace, L, S, U)
    emitted, x1, _ = code(full)
    assert O.encode_chunk(chunk, S, U, m, tail_pad=True, L=L) == emitted + bytes([x1 >> 24])
    # and the same chunk as the last one of a batch (only it is padded)
    head = datagen.markov2(2 * S, 36).tobytes()
    blob = O.encode_batch(head + chunk, S, U, m, tail_pad=True)
    assert payloads(blob, 3)[2] == emitted + bytes([x1 >> 24])


def test_stale_slot_rule_distinguishes_units():
    """The rule picks the LAST unit's bit for a slot: taking the first unit's
    (a plausible misreading) gives different bytes on this input."""
    L, S, U, tb = 40, 100, 64, 14
    m = O.Model(tb)
    O.prelearn_train(datagen.markov2(3000, 34).tobytes(), m)
    chunk = datagen.markov2(L, 35 + L).tobytes()
    dtrace = data_trace(chunk, U, m)
    wrong = list(dtrace)
    for b in range(8 * L, 8 * S):
        r = b % U
        wrong.append(dtrace[r] if r < 8 * L else (0, 0))
    emitted, x1, _ = code(wrong)
    assert O.encode_chunk(chunk, S, U, m, tail_pad=True, L=L) != emitted + bytes([x1 >> 24])


# ---------------------------------------------------------------------------
# (d) a hand-derived two-unit known answer
# ---------------------------------------------------------------------------

def test_two_units_of_zero_bytes_known_answer():
    """8 zero bytes, S = 8, U = 32, 2^18 slots, zero model: two units.

    Unit 0 (bits 0..31): predictions from the zero model, pr = 2048 (see the
    complement test); y = 0 throughout, so 4 bytes 0xFF are emitted and the
    interval returns to (0, 2^32 - 1).

    Its deferred updates (cucm_compress_v2.cu:1084-1085):
      * mixer: every bit has cxt = 0 (no active order, previous byte 0) and
        err = (0 - 2048) * 7 = -14336; each delta is (-14336 * 1) >> 16 = -1,
        so after 32 bits w[0][k] = -32 for all 8 inputs;
      * StateMaps: with all-zero bytes every order hashes (h0, 0, ..., 0) with
        h0 = 1 << (b mod 8), so bits b, b+8, b+16, b+24 share one slot per
        order; y = 0 walks each slot 0 -> 1 -> 3 -> 7 -> 15 through
        nex[s][0] (cucm_compress_v2.cu:36-43: {1,2}, {3,5}, {7,10}, {15,19}),
        and statep[k][15] is never updated, so it stays 2^31.
    Unit 1 (bits 32..63): every slot is in state 15, p = 2^31 >> 20 = 2048,
    stretch = 1, no order active, cxt = 0; dot = 8 * (-32) = -256 and
    -256 >> 16 = -1; pr = squash(-1) = sq[2046] =
    (t[15] * 1 + t[16] * 127 + 64) >> 7 = (1546 + 259969 + 64) >> 7 = 2043
    (cucm_compress_v2.cu:681-691), bumped to 2044.  So unit 1 codes 32 bits
    (pr, y) = (2044, 0); then the flush byte.  With immediate updates
    (U = 1) bit 8 would already see state 1 and a different prediction."""
    data = bytes(8)
    sl = [O.slots(data, b, 18) for b in range(8)]
    for k in range(8):
        assert len({s[k] for s in sl}) == 8  # the 8 slots of an order are distinct
    assert [NEX[2 * s] for s in (0, 1, 3, 7)] == [1, 3, 7, 15]
    assert SQ[2046] == 2043
    trace = [(2048, 0)] * 32 + [(2044, 0)] * 32
    emitted, x1, _ = code(trace)
    assert emitted[:4] == b"\xff" * 4
    expect = emitted + bytes([x1 >> 24])
    assert O.encode_chunk(data, 8, 32, O.Model(18), tail_pad=True) == expect
    assert O.encode_batch(data, 8, 32, O.Model(18)) == (4 + len(expect) - 1).to_bytes(4, "big") + expect
    # the model after both units: w[0][k] = -32 - 32 * ((-14336 * 7 ... )) is
    # not needed; the decoder must give the zeros back
    assert O.decode_chunk(expect, 8, 8, 32, O.Model(18)) == data
    assert O.encode_chunk(data, 8, 1, O.Model(18)) != expect


# ---------------------------------------------------------------------------
# the schedule twin on small inputs
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("L,U,tb,pre", [(300, 32, 12, False), (200, 96, 14, True),
                                        (123, 128, 10, True), (64, 256, 16, False),
                                        (90, 64, 8, True)])
def test_unit_schedule_restated_from_primitives(L, U, tb, pre):
    """Without padding, the oracle's chunk payload equals the data trace of
    the unit schedule (restated above from the pinned primitives) run
    through the pinned coder step, for several U, table sizes and initial
    models; U not dividing 8L leaves a short last unit."""
    m = O.Model(tb)
    if pre:
        O.prelearn_train(datagen.markov2(2500, 37).tobytes(), m)
    chunk = datagen.markov2(L, 38 + U).tobytes()
    emitted, x1, _ = code(data_trace(chunk, U, m))
    assert O.encode_chunk(chunk, L, U, m, tail_pad=False) == emitted + bytes([x1 >> 24])
