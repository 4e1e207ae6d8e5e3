"""Writes the golden fixtures in this folder from the CPU oracle.

Run from the repository root:  python tests/golden/make_golden.py
The oracle is pinned (tests/test_oracle.py) before these files are trusted;
they freeze its outputs so the GPU tests can compare against them without
rebuilding the oracle trie every time.
"""
import json
import os
import random
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

from conftest import load_package  # noqa: E402
import oracle  # noqa: E402


def main():
    wl = load_package().workloads
    rng = random.Random(2024)
    vecs = []
    for n in [0, 1, 31, 32, 55, 56, 100, 135, 136, 137, 271, 272, 273, 500, 1000, 1100]:
        msg = bytes(rng.getrandbits(8) for _ in range(n))
        vecs.append({"msg": msg.hex(), "digest": oracle.keccak256(msg).hex()})
    with open(os.path.join(HERE, "keccak256.json"), "w") as f:
        json.dump(vecs, f, indent=1)

    roots = {}
    roots["config1"] = oracle.root_of(wl.config1_oracle()).hex()
    small = wl.batch_from_bytes([b"doe", b"dog", b"dogglesworth"], [b"reindeer", b"puppy", b"cat"])
    roots["doe_dog_dogglesworth"] = oracle.root_of(small).hex()
    with open(os.path.join(HERE, "roots.json"), "w") as f:
        json.dump(roots, f, indent=1)
    print(json.dumps(roots, indent=1))


if __name__ == "__main__":
    main()
