"""Writes tests/golden/points_sha256.json: SHA-256 of the first points of each
benchmark configuration from dgx_gen_points, so that drift of the generator
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
    return {str(c): hashlib.sha256(dgx.config_points(c, N).tobytes()).hexdigest() for c in (1, 2, 3, 5)}


if __name__ == "__main__":
    with open(os.path.join(HERE, "points_sha256.json"), "w") as f:
        json.dump({"n": N, "sha256": digests()}, f, indent=1)
