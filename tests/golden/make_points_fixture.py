"""Writes tests/golden/points_sha256.json: SHA-256 of the first points of each
benchmark workload from dgx_gen_points, so that drift of the generator
(or of libm) between machines is detected."""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import dendro4_amd as dgx  # noqa: E402

N = 100000


def digests():
    return {w: hashlib.sha256(dgx.workload_points(w, N).tobytes()).hexdigest() for w in dgx.WORKLOADS}


if __name__ == "__main__":
    with open(os.path.join(HERE, "points_sha256.json"), "w") as f:
        json.dump({"n": N, "sha256": digests()}, f, indent=1)
