"""The golden fixtures in tests/golden: each file matches the SHA-256 its
manifest records, and the oracle, rerun on the fixture's seeded inputs,
reproduces every stored array bit for bit (the fixtures freeze the pinned
oracle; the GPU side is checked against the files in test_gpu_golden.py)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, HERE)
import make_golden  # noqa: E402

MANIFEST = json.load(open(os.path.join(HERE, "manifest.json")))


def test_manifest_lists_every_case():
    assert sorted(MANIFEST) == sorted(make_golden.CASES)


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_fixture_hash(name):
    with open(os.path.join(HERE, name + ".npz"), "rb") as f:
        assert hashlib.sha256(f.read()).hexdigest() == MANIFEST[name]["sha256"]


@pytest.mark.parametrize("name", sorted(make_golden.CASES))
def test_oracle_reproduces_fixture(name):
    par, arrs = make_golden.CASES[name]()
    stored = np.load(os.path.join(HERE, name + ".npz"), allow_pickle=False)
    assert sorted(stored.files) == sorted(arrs) == MANIFEST[name]["arrays"]
    for k, v in arrs.items():
        np.testing.assert_array_equal(stored[k], np.asarray(v, dtype=np.float64), err_msg=f"{name}:{k}")
    assert json.loads(json.dumps(par)) == MANIFEST[name]["params"]
