"""The C++ facade (include/dendro_mi355x.hpp): the header compiles with g++
on its own, the facade test program links against the library and the
oracle, fails loudly without a GPU, and on the GPU reproduces the oracle's
octrees and DA maps and solves the Dirichlet problem through DAMG."""
import os
import subprocess

import pytest

from conftest import ROOT

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "facade_test")


def test_facade_header_compiles_standalone(tmp_path):
    src = tmp_path / "use.cpp"
    src.write_text('#include "dendro_mi355x.hpp"\n'
                   'int main() { std::vector<ot::TreeNode> a, b; (void)a; (void)b; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


IO_PROGRAM = r"""
#include <cstdio>
#include "dendro_mi355x.hpp"
int main(int argc, char **argv) {
    const std::string dir = argv[2];
    std::vector<double> pts;
    ot::readPtsFromFile(argv[1], pts);
    if (pts.size() != 3 * 1156) { std::printf("IO FAIL pts %zu\n", pts.size()); return 1; }
    ot::writePtsToFile((dir + "/copy.pts").c_str(), pts);
    std::vector<ot::TreeNode> nodes, back, bin;
    for (unsigned i = 0; i < 8; ++i)
        nodes.push_back(ot::TreeNode((i & 1u) << 29, ((i >> 1) & 1u) << 29, ((i >> 2) & 1u) << 29, 1, 3, 30));
    ot::writeNodesToFile((dir + "/t.ot").c_str(), nodes);
    ot::readNodesFromFile((dir + "/t.ot").c_str(), back);
    ot::writeNodesToFile_binary((dir + "/t.bin").c_str(), nodes);
    const int nb = ot::readNodesFromFile_binary((dir + "/t.bin").c_str(), bin);
    if (back != nodes || bin != nodes || nb != 8) { std::printf("IO FAIL nodes\n"); return 1; }
    treeNodesTovtk(nodes, 0, dir + "/mesh", true);
    std::printf("IO OK\n");
    return 0;
}
"""


def test_facade_file_formats_without_gpu(tmp_path):
    """ot::readPtsFromFile / writePtsToFile / readNodesFromFile(_binary) /
    writeNodesToFile(_binary) / treeNodesTovtk through the facade: host code,
    no device needed."""
    src = tmp_path / "io.cpp"
    src.write_text(IO_PROGRAM)
    exe = tmp_path / "io"
    lib = os.path.join(ROOT, "dendro4_amd", "_lib")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-o", str(exe), str(src), "-I",
                        os.path.join(ROOT, "include"), "-L", lib, "-ldendro_mi355x", "-Wl,-rpath," + lib],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    golden = os.path.join(ROOT, "tests", "golden", "p1K.pts")
    r = subprocess.run([str(exe), golden, str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "IO OK" in r.stdout, r.stdout + r.stderr
    assert open(tmp_path / "copy.pts", "rb").read() == open(golden, "rb").read()
    assert open(tmp_path / "t.ot").read().splitlines()[:3] == ["3 30", "8", "0 0 0 1"]
    assert (tmp_path / "mesh_H_0.vtk").exists()


def _need_binary():
    if not os.path.exists(BIN):
        import __graft_entry__
        __graft_entry__.build_facade_test()


def test_facade_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    _need_binary()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "FACADE ERROR" in r.stdout and "dgx_init" in r.stdout


@pytest.mark.gpu
def test_facade_end_to_end_on_gpu():
    _need_binary()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout[-4000:], r.stderr[-2000:])
    assert r.returncode == 0 and "FACADE OK" in r.stdout
