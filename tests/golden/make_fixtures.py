"""Writes the golden fixtures: small seeded inputs and the containers the
oracle produces for them.

These are self-pinned vectors (the reference ships no tests, fixtures or
known answers, and its CUDA sources cannot be built here): they freeze the
oracle's output so that any later change to the oracle or to the HIP codec
that alters a single byte is caught.  Run from the repository root:

    python tests/golden/make_fixtures.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from cucm_amd import datagen  # noqa: E402
from oracle import oracle as O  # noqa: E402

# name, generator, n, seed, S, U, batch, P, tb, frozen model
FIXTURES = [
    ("markov2_64k", "markov2", 65536, 1, 16384, 128, 100, 0, 18, False),
    ("mixed_32k", "mixed", 32768, 2, 4096, 64, 3, 0, 16, False),
    ("structured_24k", "structured", 24576, 4, 8192, 128, 2, 0, 20, False),
    ("tail_10b", "markov2", 10, 5, 4096, 128, 100, 0, 18, False),
    ("prelearn_8k", "markov2", 8192, 6, 2048, 32, 2, 2048, 18, False),
    ("tb25_3000", "markov2", 3000, 7, 3000, 128, 1, 0, 25, False),
    ("u256_20k", "mixed", 20000, 8, 5000, 256, 4, 0, 14, False),
    ("empty", "markov2", 0, 9, 4096, 128, 100, 0, 18, False),
    ("frozen_prelearn_24k", "markov2", 24576, 10, 4096, 128, 3, 8192, 16, True),
    ("frozen_zero_5k", "mixed", 5000, 11, 2048, 64, 100, 0, 12, True),
]


def make_input(gen, n, seed):
    if n == 0:
        return b""
    return getattr(datagen, gen)(n, seed).tobytes()


def main():
    manifest = []
    for name, gen, n, seed, S, U, batch, P, tb, frozen in FIXTURES:
        data = make_input(gen, n, seed)
        blob = O.compress(data, S, U, batch, P, tb, frozen=frozen)
        assert O.decompress(blob) == data
        with open(os.path.join(HERE, name + ".in"), "wb") as f:
            f.write(data)
        with open(os.path.join(HERE, name + ".cucm"), "wb") as f:
            f.write(blob)
        manifest.append({"name": name, "generator": gen, "n": n, "seed": seed, "chunk_size": S,
                         "unit_bits": U, "batch_chunks": batch, "prelearn_size": P,
                         "table_bits": tb, "frozen": frozen,
                         "input_sha256": hashlib.sha256(data).hexdigest(),
                         "container_sha256": hashlib.sha256(blob).hexdigest(),
                         "container_bytes": len(blob)})
        print("%-16s %6d -> %6d bytes" % (name, len(data), len(blob)))
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1)


if __name__ == "__main__":
    main()
