"""Host-side logic that needs no GPU."""
import os
import subprocess
import sys

import cucm_amd
from cucm_amd import codec

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_more_than_one_gpu():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True)
    assert p.returncode == 2
    assert "one GPU" in p.stderr


def test_bench_container_parser_matches_oracle_layout():
    sys.path.insert(0, ROOT)
    import bench
    from oracle import oracle as O
    data = bytes(range(256)) * 40
    blob = O.compress(data, 1000, 64, 3, 0, 12)
    pays = bench.chunk_payloads(blob, 100)
    assert [c for c, _ in pays] == list(range(11))
    m = O.Model(12)
    for c, p in pays:
        chunk = data[c * 1000:(c + 1) * 1000]
        assert O.encode_chunk(chunk, 1000, 64, m, tail_pad=(c == 10), L=len(chunk)) == p


def test_make_params_flags():
    p = codec.make_params(tail_pad=False)
    assert p.flags == cucm_amd._native.FLAG_NO_TAIL_PAD
    assert codec.make_params().flags == 0
