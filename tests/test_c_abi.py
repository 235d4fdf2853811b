"""The C ABI used from plain C (tests/c/abi_example.c, C11 with -pedantic):
it compiles and links here against include/kkw.h and libkkw.so; on a GPU its
node2vec dump equals the oracle's paths and its batch replays agree."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "abi_example.c")


def build(tmp_path):
    exe = str(tmp_path / "abi_example")
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", exe, "-L", os.path.join(ROOT, "knightking_amd"), "-lkkw",
                    "-Wl,-rpath," + os.path.join(ROOT, "knightking_amd")], check=True)
    return exe


def test_c_example_compiles_and_links(tmp_path):
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_c_example_matches_oracle(tmp_path, karate):
    exe = build(tmp_path)
    out = str(tmp_path / "n2v.txt")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "karate.txt"), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    src, dst, w = karate
    rp, col, ww = O.csr_from_edges(34, src, dst, w, False, True)
    paths, lens, _, _ = O.walk(O.Graph(rp, col, ww), O.NODE2VEC, 34, 10, True, p=2.0, q=0.5, seed=5)
    rows = [[int(x) for x in line.split()] for line in open(out)]
    assert len(rows) == 34
    for i, row in enumerate(rows):
        assert row[0] == i and row[1] == lens[i] and row[2:] == paths[i, :lens[i]].tolist()
