"""On-disk formats of the reference (src/octFunctions.cpp:319-488,
src/treenode2vtk.cpp): the library's readers and writers against the
reference's own data files (data/p1K.pts, data/p2K.pts, copied under
tests/golden/), byte-level layouts stated by the reference code, and round
trips.  Host code only: these run without a GPU, except the last test, which
builds octrees from the reference's point files on the GPU."""
import os
import struct

import numpy as np
import pytest

from conftest import GOLDEN

P1K = os.path.join(GOLDEN, "p1K.pts")
P2K = os.path.join(GOLDEN, "p2K.pts")


def raw_pts(path):
    b = open(path, "rb").read()
    n = struct.unpack("<I", b[:4])[0]
    return n, np.frombuffer(b[4:4 + 24 * n], "<f8").reshape(n, 3)


@pytest.mark.parametrize("path,count", [(P1K, 1156), (P2K, 2752)])
def test_reads_reference_point_files(path, count):
    import dendro4_amd as dgx
    pts = dgx.read_pts(path)
    n, ref = raw_pts(path)
    assert n == count and pts.shape == (count, 3)
    assert np.array_equal(pts, ref)
    # the reference's files hold octant anchors of an adaptive tree: a 1/64 grid in [0, 1)
    assert pts.min() == 0.0 and pts.max() < 1.0 and np.all(pts * 64 == np.round(pts * 64))


@pytest.mark.parametrize("path", [P1K, P2K])
def test_pts_round_trip_is_byte_identical(path, tmp_path):
    import dendro4_amd as dgx
    out = tmp_path / "copy.pts"
    dgx.write_pts(out, dgx.read_pts(path))
    assert open(out, "rb").read() == open(path, "rb").read()


def test_pts_zero_points_is_an_empty_file(tmp_path):
    import dendro4_amd as dgx
    out = tmp_path / "none.pts"
    dgx.write_pts(out, np.zeros((0, 3)))
    assert os.path.getsize(out) == 0
    assert dgx.read_pts(out).shape == (0, 3)


def test_pts_truncated_file_is_an_error(tmp_path):
    import dendro4_amd as dgx
    bad = tmp_path / "bad.pts"
    bad.write_bytes(struct.pack("<I", 10) + b"\0" * 24 * 9)
    with pytest.raises(dgx.DgxError, match="points need"):
        dgx.read_pts(bad)
    with pytest.raises(dgx.DgxError, match="cannot open"):
        dgx.read_pts(tmp_path / "missing.pts")


def test_data_points_layout(tmp_path):
    import dendro4_amd as dgx
    rng = np.random.default_rng(0)
    pts, vals = rng.random((5, 3)), rng.random(5)
    f = tmp_path / "d.pts"
    f.write_bytes(struct.pack("<I", 5) + pts.astype("<f8").tobytes() + vals.astype("<f8").tobytes())
    p2, v2 = dgx.read_data_pts(f)
    assert np.array_equal(p2, pts) and np.array_equal(v2, vals)
    # writeDataPtsToFile writes the count and the points, not the values
    g = tmp_path / "w.pts"
    dgx.write_data_pts(g, pts, vals)
    assert open(g, "rb").read() == struct.pack("<I", 5) + pts.astype("<f8").tobytes()
    h = tmp_path / "z.pts"
    dgx.write_data_pts(h, np.zeros((0, 3)))
    assert open(h, "rb").read() == struct.pack("<I", 0)


def tree_rows(oracle):
    _, pts = raw_pts(P1K)
    tn = oracle.points2octree(pts, 30, 1)
    tn[::7, 3] |= 64  # BOUNDARY flag on some octants: writers drop flags from the level
    tn[::5, 4] = 3    # a weight other than 1
    return tn


def test_nodes_text_format_and_round_trip(oracle, tmp_path):
    import dendro4_amd as dgx
    tn = tree_rows(oracle)
    f = tmp_path / "t.ot"
    dgx.write_nodes(f, tn)
    lines = open(f).read().splitlines()
    assert lines[0] == "3 30" and lines[1] == str(len(tn))
    assert lines[2:] == [f"{r[0]} {r[1]} {r[2]} {r[3] & 31}" for r in tn]
    back = dgx.read_nodes(f)
    # TreeNode(x, y, z, level, dim, maxDepth): weight 1, no flags
    want = tn.copy()
    want[:, 3] &= 31
    want[:, 4] = 1
    assert np.array_equal(back, want)
    e = tmp_path / "e.ot"
    dgx.write_nodes(e, tn[:0])
    assert os.path.getsize(e) == 0


def test_nodes_text_rejects_mixed_depths(oracle, tmp_path):
    import dendro4_amd as dgx
    tn = tree_rows(oracle)[:4].copy()
    tn[2, 6] = 20
    with pytest.raises(dgx.DgxError, match="maxDepth"):
        dgx.write_nodes(tmp_path / "m.ot", tn)


def test_nodes_binary_is_raw_treenode_records(oracle, tmp_path):
    import dendro4_amd as dgx
    tn = tree_rows(oracle)
    f = tmp_path / "t.bin"
    dgx.write_nodes_binary(f, tn)
    assert open(f, "rb").read() == tn.astype("<u4").tobytes()
    # all N records (the reference's feof loop would repeat the last one)
    assert np.array_equal(dgx.read_nodes_binary(f), tn)
    f.write_bytes(b"\0" * 30)
    with pytest.raises(dgx.DgxError, match="28-byte"):
        dgx.read_nodes_binary(f)


def expected_vtk(tn, rank, hsorted):
    """treeNodesTovtk's output for already sorted rows (src/treenode2vtk.cpp:20-131)."""
    n = len(tn)
    out = ["# vtk DataFile Version 2.0", "DENDRO OCTREES", "ASCII", "DATASET UNSTRUCTURED_GRID", f"POINTS {8 * n} float"]
    for x, y, z, lev, _, _, md in tn:
        ln = 1 << (int(md) - (int(lev) & 31))
        for dx, dy, dz in ((0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)):
            out.append(f"{x + dx * ln} {y + dy * ln} {z + dz * ln}")
    out.append(f"CELLS {n} {9 * n}")
    out += ["8 " + "".join(f"{8 * i + j} " for j in range(8)) for i in range(n)]
    out.append(f"CELL_TYPES {n}")
    out += ["12"] * n
    out.append(f"FIELD OCTREE_DATA {3 if hsorted else 2}")
    out.append(f"cell_level 1 {n} int")
    out.append("".join(f"{r[3] & 31} " for r in tn))
    out.append(f"weight 1 {n} int")
    out.append("".join(f"{r[4]} " for r in tn))
    out.append(f"mpirank 1 {n} int")
    out.append("".join(f"{rank} " for _ in tn))
    if hsorted:
        out.append(f"hilbert_index 1 {n} int ")
        out.append("".join(f"{i + 1} " for i in range(n)))
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("hsorted", [False, True])
def test_vtk_writer(oracle, tmp_path, hsorted):
    import dendro4_amd as dgx
    tn = tree_rows(oracle)[:300]
    shuffled = tn[np.random.default_rng(1).permutation(len(tn))]
    srt = dgx.write_vtk(shuffled, str(tmp_path / "mesh"), rank=2, hsorted=hsorted)
    assert np.array_equal(srt, oracle.sort(shuffled))  # the reference's std::sort (Hilbert operator<)
    text = open(tmp_path / "mesh_H_2.vtk").read()
    assert text == expected_vtk(srt, 2, hsorted)


@pytest.mark.gpu
@pytest.mark.parametrize("path,max_num_pts", [(P1K, 1), (P2K, 1), (P2K, 4)])
def test_points2octree_on_reference_point_files(dgx, oracle, path, max_num_pts):
    """points2Octree and balanceOctree on the reference's own point files
    (octant anchors on a 1/64 grid, i.e. on octant faces) against the oracle."""
    pts = dgx.read_pts(path)
    o = dgx.points2octree(pts, 30, max_num_pts)
    ref = oracle.points2octree(pts, 30, max_num_pts)
    got = o.treenodes()
    assert np.array_equal(got[:, :3], ref[:, :3]) and np.array_equal(got[:, 3] & 31, ref[:, 3] & 31)
    b = dgx.balance_octree(o)
    rb = oracle.balance(ref)
    gb = b.treenodes()
    assert np.array_equal(gb[:, :3], rb[:, :3]) and np.array_equal(gb[:, 3] & 31, rb[:, 3] & 31)
