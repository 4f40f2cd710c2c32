"""The oracle reproduces every committed golden container byte for byte
(regression pin; the fixtures are self-pinned, see make_fixtures.py)."""
import hashlib
import json
import os

import pytest

from oracle import oracle as O

HERE = os.path.join(os.path.dirname(__file__), "golden")
MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))


@pytest.mark.parametrize("fx", MANIFEST, ids=[f["name"] for f in MANIFEST])
def test_oracle_reproduces_fixture(fx):
    data = open(os.path.join(HERE, fx["name"] + ".in"), "rb").read()
    blob = open(os.path.join(HERE, fx["name"] + ".cucm"), "rb").read()
    assert hashlib.sha256(data).hexdigest() == fx["input_sha256"]
    assert hashlib.sha256(blob).hexdigest() == fx["container_sha256"]
    got = O.compress(data, fx["chunk_size"], fx["unit_bits"], fx["batch_chunks"],
                     fx["prelearn_size"], fx["table_bits"], frozen=fx["frozen"])
    assert got == blob
    assert O.decompress(blob) == data


@pytest.mark.parametrize("fx", MANIFEST, ids=[f["name"] for f in MANIFEST])
def test_fixture_inputs_regenerate(fx):
    from cucm_amd import datagen
    data = open(os.path.join(HERE, fx["name"] + ".in"), "rb").read()
    if fx["n"]:
        assert getattr(datagen, fx["generator"])(fx["n"], fx["seed"]).tobytes() == data
