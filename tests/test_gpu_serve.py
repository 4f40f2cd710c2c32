"""Serving loop (cucm_serve_start / cucm_submit / cucm_poll): requests of
mixed kinds, sizes and parameters (prelearned- and frozen-model ones among
them) submitted together share the serving kernel's running launch (or, with static lanes,
shared batches), requests that arrive mid-launch are admitted into it, and
every result equals the same request run alone -- bit-exact, since the
codec's output depends only on the request."""
import numpy as np
import pytest

import cucm_amd
from cucm_amd import datagen, make_params

pytestmark = pytest.mark.gpu


def _requests():
    rng = np.random.default_rng(23)
    reqs = []
    for i in range(40):
        n = int(rng.choice([0, 1, 7, 300, 5000, 40000, 150000, 400000]))
        data = datagen.markov2(n, 300 + i).tobytes() if n else b""
        kind = i % 6
        if kind == 0:
            p = make_params(16384, 128, 100, 0, 18)
        elif kind == 5:  # prelearned model, taken by the serving kernel (U <= 128, 2^18)
            p = make_params(16384, 128, 5, min(n, 4000), 18)
        elif kind == 1:
            p = make_params(32768, 64, 3, 0, 16)
        elif kind == 2:
            p = make_params(8192, 256, 100, min(n, 2000), 14)
        elif kind == 3:
            p = make_params(65536, 96, 7, 0, 20, tail_pad=False)
        else:  # frozen model: runs in the same launches as the others
            p = make_params(16384, 128, 4, min(n, 3000), 16, frozen=True)
        reqs.append((data, p))
    return reqs


@pytest.mark.parametrize("static_lanes", [False, True], ids=["kernel", "static-lanes"])
def test_serve_results_identical_alone_and_in_mix(ctx, static_lanes):
    reqs = _requests()
    alone_c = [ctx.compress(d, p) for d, p in reqs]
    for (d, _), b in zip(reqs, alone_c):
        assert ctx.decompress(b) == d
    sc = cucm_amd.Context()
    try:
        sc.serve_start(lanes=3, static_lanes=static_lanes)
        ids = []
        for i, (d, p) in enumerate(reqs):
            ids.append(("c", i, sc.submit_compress(d, p)))
            # interleave decompressions of containers made alone
            ids.append(("d", i, sc.submit_decompress(alone_c[i], spec_bits=[1, 2, 4, 8][i % 4])))
        batch = []
        for kind, i, jid in ids:
            out, info = sc.poll(jid, wait=True)
            batch.append(info["batch_jobs"])
            if kind == "c":
                assert out == alone_c[i], "compress request %d differs in the mix" % i
            else:
                assert out == reqs[i][0], "decompress request %d differs in the mix" % i
        launches = sc.serve_launches()
        sc.serve_stop()
    finally:
        sc.close()
    assert launches < len(ids), "requests were never batched"
    assert max(batch) > 1


def test_serve_poll_nowait_and_errors(ctx):
    sc = cucm_amd.Context()
    try:
        sc.serve_start(lanes=2, max_inflight=4)
        d = datagen.markov2(100000, 77).tobytes()
        jid = sc.submit_compress(d, make_params(16384, 128, 100, 0, 18))
        got = None
        while got is None:
            got = sc.poll(jid, wait=False)
        blob, info = got
        assert info["latency_ms"] >= info["queue_ms"] >= 0
        assert blob == ctx.compress(d, make_params(16384, 128, 100, 0, 18))
        # a corrupt container fails alone, without hurting its neighbours
        bad = bytearray(blob)
        bad[25:29] = b"\xff\xff\xff\xff"
        j_bad = sc.submit_decompress(bytes(bad))
        j_ok = sc.submit_decompress(blob)
        with pytest.raises(cucm_amd.CucmError):
            sc.poll(j_bad, wait=True)
        assert sc.poll(j_ok, wait=True)[0] == d
        sc.serve_stop()
    finally:
        sc.close()


def test_serve_poisson_mix_of_the_serving_configuration(ctx):
    """The serving configuration's request mix: Poisson arrivals, half
    compress requests of fresh payloads and half decompress requests of
    containers made earlier, sizes min(round(256 KiB * exp(1.2 Z)), 128 MiB)
    with order-2 Markov content, 16 KiB chunks and 2^18 slots.  Requests are
    submitted from a thread at their arrival times while earlier ones run;
    every result equals the same request run alone."""
    import threading
    import time
    count = 48
    sizes = datagen.lognormal_sizes(count, 262144, 1.2, 1, 128 << 20, 13)
    pool_sizes = datagen.lognormal_sizes(count // 2, 262144, 1.2, 1, 128 << 20, 17)
    arena, offs = datagen.markov2_files(sizes, seed0=5000)
    parena, poffs = datagen.markov2_files(pool_sizes, seed0=9000)
    fresh = [arena[int(o):int(o) + int(n)].tobytes() for o, n in zip(offs, sizes)]
    pool = [parena[int(o):int(o) + int(n)].tobytes() for o, n in zip(poffs, pool_sizes)]
    p = make_params(16384, 128, 100, 0, 18)
    blobs = ctx.compress_many(pool, p)
    rng = np.random.default_rng(11)
    ops = rng.random(count) < 0.5
    reqs = [("c", fresh[i], ctx.compress(fresh[i], p)) if ops[i] else
            ("d", blobs[i % len(blobs)], pool[i % len(pool)]) for i in range(count)]
    mean = float(np.mean([len(r[1]) if r[0] == "c" else len(r[2]) for r in reqs]))
    arrivals = datagen.poisson_arrivals(count, 40e6 / mean, 19)  # ~40 MB/s offered
    sc = cucm_amd.Context()
    ids = [None] * count
    try:
        sc.serve_start()

        def submitter():
            t0 = time.perf_counter()
            for i, (kind, x, _) in enumerate(reqs):
                d = t0 + arrivals[i] - time.perf_counter()
                if d > 0:
                    time.sleep(d)
                ids[i] = sc.submit_compress(x, p) if kind == "c" else sc.submit_decompress(x)

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        for i, (kind, x, want) in enumerate(reqs):
            out, info = sc.poll(ids[i], wait=True)
            assert out == want, "request %d (%s, %d B) differs from the request alone" % (
                i, kind, len(x))
        sc.serve_stop()
    finally:
        sc.close()


def test_requests_admitted_into_the_running_kernel(ctx):
    """A large compress request is submitted first; small requests submitted
    while it runs start on free waves of the same launch and finish before
    it (they do not wait for it), and every result equals the request run
    alone."""
    p = make_params(16384, 128, 100, 0, 18)
    big = datagen.markov2(24 << 20, 61).tobytes()
    smalls = [datagen.markov2(20000 + 997 * i, 70 + i).tobytes() for i in range(12)]
    big_alone = ctx.compress(big, p)
    small_alone = [ctx.compress(s, p) for s in smalls]
    sc = cucm_amd.Context()
    try:
        sc.serve_start()
        launches0 = sc.serve_launches()
        jb = sc.submit_compress(big, p)
        assert sc.poll(jb, wait=False) is None  # still running
        ids = [sc.submit_compress(s, p) for s in smalls]
        done_small = []
        for i, jid in enumerate(ids):
            out, info = sc.poll(jid, wait=True)
            assert out == small_alone[i], "small request %d" % i
            assert info["batch_jobs"] >= 2  # admitted while the large one was in flight
            done_small.append(info["latency_ms"])
        assert sc.poll(jb, wait=False) is None, "the small requests waited for the large one"
        out, info = sc.poll(jb, wait=True)
        assert out == big_alone
        assert max(done_small) < info["latency_ms"]
        assert sc.serve_launches() - launches0 <= 2  # one running launch took them all
        # decompressions of the results through the same kernel
        ids = [sc.submit_decompress(b) for b in [big_alone] + small_alone]
        for jid, want in zip(ids, [big] + smalls):
            assert sc.poll(jid, wait=True)[0] == want
        sc.serve_stop()
    finally:
        sc.close()
