"""Writes tests/golden/type2_stencils.npz from the reference's stencil tables
(Stencils/LapType2Stencils.inp, Stencils/MassType2Stencils.inp; read with the
loop order of handleType2Stencils.C:678-1000: cNum, eType, i, j).  Run in a
container that has the reference checkout; the GPU box only reads the .npz."""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def read(name):
    with open(os.path.join(REF, "Stencils", name)) as f:
        return np.array(f.read().split(), dtype=np.float64).reshape(8, 18, 8, 8)


if __name__ == "__main__":
    L, M = read("LapType2Stencils.inp"), read("MassType2Stencils.inp")
    np.savez_compressed(os.path.join(HERE, "type2_stencils.npz"), lap=L, mass=M)
    h = hashlib.sha256(L.tobytes() + M.tobytes()).hexdigest()
    with open(os.path.join(HERE, "type2_stencils.sha256"), "w") as f:
        f.write(f"{h}  type2_stencils.npz (lap bytes + mass bytes, float64 C order)\n")
