"""The autotuner on the GPU, on a small uneven batch: its stream choice
keeps the ratio within tolerance of the reference setting and gives
containers equal to the oracle's with those parameters, and the launch
settings it tries never change the bytes."""
import numpy as np
import pytest

from cucm_amd import datagen
from cucm_amd.autotune import REFERENCE, apply_launch, autotune_batch, file_params
from cucm_amd.codec import output_cap

pytestmark = pytest.mark.gpu
KiB = 1024


def test_autotune_small_uneven_batch(ctx, oracle):
    sizes = datagen.lognormal_sizes(300, 16384, 1.5, 512, 1 << 20, 21)
    arena, offs = datagen.markov2_files(sizes, seed0=700)
    grid = [(32 * KiB, 16), (32 * KiB, 18), (REFERENCE["chunk_max"], REFERENCE["table_bits"])]
    enc = [{"team": True, "waves_per_block": 0}, {"team": False, "waves_per_block": 2}]
    dec = [{"team": False, "waves_per_block": 0, "spec_bits": d} for d in (1, 2, 4)] + \
        [{"team": True, "waves_per_block": 0, "spec_bits": 2}]
    res = autotune_batch(ctx, arena, offs, sizes, sample_bytes=3 << 20, ratio_tol=0.01,
                         stream_grid=grid, encoder_launch=enc, decoder_launch=dec)
    trials = res["stream_trials"]
    ref = [t for t in trials if (t["chunk_max"], t["table_bits"]) ==
           (REFERENCE["chunk_max"], REFERENCE["table_bits"])][0]
    chosen = [t for t in trials if (t["chunk_max"], t["table_bits"]) ==
              (res["chunk_max"], res["table_bits"])][0]
    assert chosen["ratio"] >= ref["ratio"] * 0.99
    assert res["encoder"] in [{k: e[k] for k in ("team", "waves_per_block")} for e in enc]
    assert res["decoder"]["spec_bits"] in (1, 2, 4)
    # the chosen stream settings give the oracle's containers; every launch
    # setting tried gives the same bytes and decodes losslessly
    idx = list(range(0, 300, 25)) + [int(np.argmax(sizes))]
    files = [arena[int(offs[i]):int(offs[i]) + int(sizes[i])].tobytes() for i in idx]
    params = [file_params(len(f), res["chunk_max"], res["table_bits"]) for f in files]
    n = len(files)
    d_in = [ctx.alloc(len(f)) for f in files]
    for d, f in zip(d_in, files):
        d.upload(f)
    caps = [output_cap(len(f), p) for f, p in zip(files, params)]
    d_c = [ctx.alloc(c) for c in caps]
    d_o = [ctx.alloc(len(f)) for f in files]
    try:
        outs = None
        for e in enc:
            apply_launch(ctx, e)
            cl, st = ctx.compress_many_device([d.addr for d in d_in], [len(f) for f in files],
                                              params, [d.addr for d in d_c], caps)
            assert not any(st)
            got = [d_c[i].download(cl[i]) for i in range(n)]
            if outs is None:
                outs = got
                for i, f in enumerate(files):
                    assert got[i] == oracle.compress(f, params[i].chunk_size, 128, 100, 0,
                                                     res["table_bits"]), "file %d" % idx[i]
            assert got == outs, "launch setting %s changed the bytes" % e
        for d in dec:
            apply_launch(ctx, d)
            dl, st = ctx.decompress_many_device([x.addr for x in d_c], [len(o) for o in outs],
                                                [x.addr for x in d_o], [len(f) for f in files],
                                                d["spec_bits"])
            assert not any(st)
            assert [d_o[i].download(len(files[i])) for i in range(n)] == files, d
    finally:
        ctx.set_team(0)
        ctx.set_waves_per_block(0)
