"""Writes tests/golden/generator.json from the REFERENCE Graph500 generator.

Runs oracle/_ref/gen_ref (the reference's generator/ sources, SEQ variant,
compiled by `make -C oracle ref`; see oracle/Makefile) for a set of
(scale, edge count, seed pair) and stores, per case, the edge count, the
SHA-256 of the sorted multiset of packed edges (u << 32 | v, little-endian
uint64) and a few sorted edges.  Only the fixture file is committed.

    make -C oracle ref && python tests/golden/make_generator_fixtures.py
"""
import hashlib
import json
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
GEN = os.path.join(ROOT, "oracle", "_ref", "gen_ref")

CASES = [
    # (scale, nedges, seed1, seed2)
    (4, 64, 1, 2), (6, 1024, 1, 2), (8, 4096, 1, 2), (8, 2048, 23, 24), (10, 16384, 1, 2),
    (10, 16384, 2, 3), (11, 16384, 4, 5), (12, 65536, 23, 24), (12, 131072, 7, 8),
    (14, 262144, 1, 2), (16, 1048576, 1, 2), (16, 524288, 2, 3), (18, 4194304, 2, 3),
]


def digest(edges: np.ndarray) -> dict:
    k = np.sort((edges[:, 0].astype(np.uint64) << np.uint64(32)) | edges[:, 1].astype(np.uint64))
    return {
        "count": int(k.size),
        "sha256": hashlib.sha256(k.astype("<u8").tobytes()).hexdigest(),
        "head": [[int(x >> np.uint64(32)), int(x & np.uint64(0xFFFFFFFF))] for x in k[:4]],
        "tail": [[int(x >> np.uint64(32)), int(x & np.uint64(0xFFFFFFFF))] for x in k[-4:]],
        "self_loops": int((edges[:, 0] == edges[:, 1]).sum()),
    }


def main():
    out = []
    with tempfile.TemporaryDirectory() as td:
        for scale, ne, s1, s2 in CASES:
            f = os.path.join(td, "e.bin")
            subprocess.run([GEN, str(scale), str(ne), str(s1), str(s2), f], check=True,
                           capture_output=True)
            e = np.fromfile(f, dtype=np.int64).reshape(-1, 2)
            d = digest(e)
            d.update({"scale": scale, "nedges": ne, "seed1": s1, "seed2": s2})
            out.append(d)
            print(scale, ne, s1, s2, d["count"], d["sha256"][:16])
    with open(os.path.join(HERE, "generator.json"), "w") as fh:
        json.dump({"initiator": [0.57, 0.19, 0.19, 0.05], "cases": out}, fh, indent=1)


if __name__ == "__main__":
    main()
