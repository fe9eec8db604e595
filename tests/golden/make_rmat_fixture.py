"""Writes tests/golden/rmat_stencils.npz from the reference's restriction data
files (read in the loop orders of their loaders):
  - Stencils/RmatType1Stencils.inp: rmat1[cNumCoarse][cNumFine][eType][cCtr][fCtr]
    (src/omg/rmatType1StencilsAux.cpp:13-47);
  - Stencils/RmatType2Stencils.inp: rmat2[cNum][eType][cCtr][fCtr]
    (src/omg/rmatType2StencilsAux.cpp:13-44);
  - Stencils/vtxMap.inp: per fine element number f (8), map1[f][cNumCoarse][eType][v]
    then map2[f][cNumFine][cNumCoarse][eType][v]; then per f (7) and scaling s (2),
    map3[f][s][cNumCoarse][eType][v] then map4[f][s][cNumFine][cNumCoarse][eType][v]
    (src/omg/vtxMaps.cpp:13-97).
Run in a container that has the reference checkout; the GPU box only reads the .npz."""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"


def numbers(name, dtype):
    with open(os.path.join(REF, "Stencils", name)) as f:
        return np.array(f.read().split(), dtype=dtype)


def vtx_maps():
    v = numbers("vtxMap.inp", np.int64)
    m1 = np.zeros((8, 8, 18, 8), np.uint16)
    m2 = np.zeros((8, 8, 8, 18, 8), np.uint16)
    m3 = np.zeros((7, 2, 8, 18, 8), np.uint16)
    m4 = np.zeros((7, 2, 8, 8, 18, 8), np.uint16)
    p = 0

    def take(shape):
        nonlocal p
        n = int(np.prod(shape))
        out = v[p:p + n].reshape(shape)
        p += n
        return out

    for f in range(8):
        m1[f] = take((8, 18, 8))
        m2[f] = take((8, 8, 18, 8))
    for f in range(7):
        for s in range(2):
            m3[f, s] = take((8, 18, 8))
            m4[f, s] = take((8, 8, 18, 8))
    assert p == v.size, (p, v.size)
    return m1, m2, m3, m4


if __name__ == "__main__":
    r1 = numbers("RmatType1Stencils.inp", np.float64).reshape(8, 8, 18, 8, 8)
    r2 = numbers("RmatType2Stencils.inp", np.float64).reshape(8, 18, 8, 8)
    m1, m2, m3, m4 = vtx_maps()
    np.savez_compressed(os.path.join(HERE, "rmat_stencils.npz"), rmat1=r1, rmat2=r2, map1=m1, map2=m2, map3=m3,
                        map4=m4)
    h = hashlib.sha256(b"".join(a.tobytes() for a in (r1, r2, m1, m2, m3, m4))).hexdigest()
    with open(os.path.join(HERE, "rmat_stencils.sha256"), "w") as f:
        f.write(f"{h}  rmat_stencils.npz (rmat1, rmat2 float64; map1..map4 uint16; C order, concatenated)\n")
