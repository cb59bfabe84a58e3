"""The committed known answers (tests/golden/mayo_golden.json) are what the
oracle produces today: any change to the oracle's outputs shows up here, and
the GPU tests compare the kernels with the same file."""
import hashlib
import json
import os

import pytest

from tests.golden import make_golden as G

HERE = os.path.dirname(os.path.abspath(__file__))


def load():
    with open(os.path.join(HERE, "golden", "mayo_golden.json")) as f:
        return json.load(f)


def test_oracle_reproduces_golden_file():
    assert G.build() == load()


def test_golden_seed_rule_and_shapes():
    """sk = SHAKE128(LE64(2i + 1)) for rand_index 0 (keypair.cu:32-40), and
    every recorded signature verifies; forced failures retry or exhaust."""
    from oracle import oracle as O
    gold = load()
    for name, d in gold.items():
        v = int(name.split("_")[1])
        p = O.params(v)
        for k in d["keys"]:
            want = hashlib.shake_128((2 * k["item"] + 1).to_bytes(8, "little")).digest(p.sk_bytes)
            assert bytes.fromhex(k["sk"]) == want
        assert all(s["verify"] == 0 for s in d["signatures"])
        assert [f["attempts"] for f in d["forced_failures"]] == [2, 4, 256]
