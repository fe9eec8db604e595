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
