"""The seeded generators are deterministic and match the workload specs."""
import hashlib

import numpy as np

from cucm_amd import datagen


def digest(a):
    return hashlib.sha256(a.tobytes()).hexdigest()[:16]


def test_generators_are_deterministic():
    assert digest(datagen.markov2(4096, 1)) == digest(datagen.markov2(4096, 1))
    assert digest(datagen.mixed(1 << 16, 2)) == digest(datagen.mixed(1 << 16, 2))
    assert digest(datagen.structured(1 << 20, 4)) == digest(datagen.structured(1 << 20, 4))
    assert digest(datagen.markov2(4096, 1)) != digest(datagen.markov2(4096, 2))


def test_markov2_alphabet():
    a = datagen.markov2(1 << 16, 1)
    assert a.min() >= 0x20 and a.max() < 0x20 + 64


def test_mixed_has_random_blocks_and_text():
    a = datagen.mixed(1 << 20, 2)
    blocks = a.reshape(-1, 4096)
    printable = ((blocks >= 0x20) & (blocks < 0x80)).mean(axis=1)
    frac_random = (printable < 0.7).mean()
    assert 0.08 < frac_random < 0.25


def test_lognormal_sizes_match_spec():
    s = datagen.lognormal_sizes(10000, 65536, 1.5, 512, 64 << 20, 3)
    assert s.min() >= 512 and s.max() <= 64 << 20
    med = float(np.median(s))
    assert 55000 < med < 76000
    assert 1.6 < s.sum() / 2 ** 30 < 2.3


def test_markov2_files_are_per_file_seeded():
    sizes = [100, 2000, 33]
    arena, offs = datagen.markov2_files(sizes, seed0=1000)
    for i, n in enumerate(sizes):
        assert arena[offs[i]:offs[i] + n].tobytes() == datagen.markov2(n, 1000 + i).tobytes()
