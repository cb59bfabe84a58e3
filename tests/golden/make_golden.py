"""Writes mayo_golden.json: known answers of the oracle (the CPU
restatement of the reference algorithm) for fixed seeds and messages, all
four parameter sets. Signing uses R = 0, so every value is deterministic.

    python tests/golden/make_golden.py

Keys come from the reference's counter seed rule with rand_index 0 (items
0..2). Messages are SHAKE256("mayo-golden/<set>/<j>") of the listed lengths
(empty, one byte, one rate block of 136 bytes and one past it). Large
outputs are recorded as SHA-256 digests.
"""
import hashlib
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

SETS = (1, 2, 3, 5)
LENS = (0, 1, 32, 136, 137, 1000)
FORCED = (1, 3, 255)
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "mayo_golden.json")


def message(v: int, j: int, n: int) -> bytes:
    return hashlib.shake_256(f"mayo-golden/{v}/{j}".encode()).digest(n)


def sha(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def build() -> dict:
    out = {}
    for v in SETS:
        keys, esks, epks = [], [], []
        for i in range(3):
            pk, sk = O.keypair(v, 0, i)
            esk, epk = O.expandsk(v, sk), O.expandpk(v, pk)
            esks.append(esk)
            epks.append(epk)
            keys.append({"item": i, "sk": sk.hex(), "pk_sha256": sha(pk),
                         "esk_sha256": sha(esk), "epk_sha256": sha(epk)})
        sigs = []
        for j, n in enumerate(LENS):
            k = j % 3
            m = message(v, j, n)
            sig, att = O.sign(v, esks[k], m)
            sigs.append({"j": j, "key": k, "msg_len": n, "sig_sha256": sha(sig),
                         "sig": sig.hex() if j < 2 else None, "attempts": att,
                         "verify": O.verify(v, epks[k], m, sig)})
        forced = []
        m = message(v, 100, 64)
        for ff in FORCED:
            sig, att = O.sign(v, esks[0], m, fail_first=ff)
            forced.append({"fail_first": ff, "key": 0, "j": 100, "msg_len": 64,
                           "sig_sha256": sha(sig), "attempts": att})
        out[f"MAYO_{v}"] = {"keys": keys, "signatures": sigs, "forced_failures": forced}
    return out


if __name__ == "__main__":
    with open(OUT, "w") as f:
        json.dump(build(), f, indent=1)
    print(OUT)
