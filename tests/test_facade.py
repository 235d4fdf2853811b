"""The C++ facade (include/knightking/) compiles as C++11 against kkw.h, and on
a GPU it reproduces the oracle's paths through node2vec()/deepwalk()/ppr()
and a WalkProgram, with WalkConfig output files, epochs and consumers."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "facade_karate.cpp")


def build(tmp_path):
    exe = str(tmp_path / "facade_karate")
    # the HIP runtime API is only for the test's own copy of the device-consumer paths
    subprocess.run(["g++", "-std=c++11", "-O1", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__",
                    "-isystem", "/opt/rocm/include", "-I", os.path.join(ROOT, "include"), SRC,
                    "-o", exe, "-L", os.path.join(ROOT, "knightking_amd"), "-lkkw",
                    "-L", "/opt/rocm/lib", "-lamdhip64",
                    "-Wl,-rpath," + os.path.join(ROOT, "knightking_amd")], check=True)
    return exe


def test_facade_compiles(tmp_path):
    assert os.path.exists(build(tmp_path))


def read_dump(path, head):
    rows = []
    for line in open(path):
        f = [int(x) for x in line.split()]
        rows.append((f[0], f[1], f[2:]) if head else (None, len(f), f))
    return rows


@pytest.mark.gpu
def test_facade_matches_oracle(tmp_path, karate):
    exe = build(tmp_path)
    prefix = str(tmp_path / "out")
    prog_src = os.path.join(ROOT, "knightking_amd", "programs", "return_bias.cpp")
    r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "karate.txt"), prefix, prog_src],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    src, dst, w = karate
    rp, col, ww = O.csr_from_edges(34, src, dst, w, False, True)
    g = O.Graph(rp, col, ww)
    paths, lens, _, _ = O.walk(g, O.NODE2VEC, 34, 10, True, p=2.0, q=0.5, seed=1)
    for name in ("_n2v.0", "_n2v_dev.0"):   # PathSet dump, and the device consumer's rows
        got = read_dump(prefix + name, False)
        assert len(got) == 34
        for i, (_, n, verts) in enumerate(got):
            assert verts == paths[i, :lens[i]].tolist(), name
    rp2, col2, w2 = O.csr_from_edges(34, src, dst, w, False, False)
    g2 = O.Graph(rp2, col2, w2)
    paths, lens, _, _ = O.walk(g2, O.DEEPWALK, 34, 12, True, seed=1)
    got = read_dump(prefix + "_dw.0", True)
    assert [x[0] for x in got] == list(range(34))     # two epochs appended in order
    for i, (wid, n, verts) in enumerate(got):
        assert n == lens[i] and verts == paths[i, :lens[i]].tolist()
    paths, lens, _, _ = O.walk(g2, O.PPR, 20, 0, True, terminate_prob=0.15, seed=2, start_mode=O.START_LIST,
                               starts=np.array([0, 33], np.uint32))
    got = read_dump(prefix + "_ppr.0", True)
    for i, (wid, n, verts) in enumerate(got):
        assert wid == i and verts == paths[i, :lens[i]].tolist()
    pa, la, _ = O.program_walk(O.program_oracle(prog_src, str(tmp_path / "orc")), rp2, col2, w2, params=[10],
                               use_weights=True, seed=1, walker_num=68, max_length=10)
    got = read_dump(prefix + "_prog.0", True)
    assert [x[0] for x in got] == list(range(68))
    for i, (wid, n, verts) in enumerate(got):
        assert n == la[i] and verts == pa[i, :la[i]].tolist()
    # metapath on binary Edge<WeightedMetaData> records, type (src + 2 dst) % 3
    et = ((np.repeat(np.arange(34), np.diff(rp2.astype(np.int64))) + 2 * col2.astype(np.int64)) % 3).astype(np.uint8)
    gm = O.Graph(rp2, col2, w2, et)
    masks = [[0b101, 0b110], [0b011], [0b100, 0b001, 0b010]]   # get_scheme_mask of the schemes written
    paths, lens, _, _ = O.walk(gm, O.METAPATH, 68, 12, True, seed=1, schemes=masks)
    got = read_dump(prefix + "_meta.0", True)
    assert [x[0] for x in got] == list(range(68))
    for i, (wid, n, verts) in enumerate(got):
        assert n == lens[i] and verts == paths[i, :lens[i]].tolist(), f"metapath walker {i}"
    assert (tmp_path / "out_schemes.txt").read_text().split() == \
        "3 3 2 1 0 1 0 1 1 1 1 1 0 3 0 0 1 1 0 0 0 1 0".split()
    vm = np.loadtxt(prefix + "_meta_vmask.txt", dtype=np.int64)
    assert np.array_equal(vm[:, 0], O.vmask_of(rp2, et))
    assert np.array_equal(vm[:, 1], np.diff(rp2.astype(np.int64)))
    assert np.array_equal(vm[:, 2], np.bincount(col2, minlength=34))
    # epochs streamed to a device stream consumer (and to the output file)
    paths, lens, _, _ = O.walk(g2, O.DEEPWALK, 34, 12, True, seed=1)
    got = read_dump(prefix + "_stream_dw.0", True)
    assert [x[0] for x in got] == list(range(34))
    for i, (wid, n, verts) in enumerate(got):
        assert n == lens[i] and verts == paths[i, :lens[i]].tolist(), f"streamed epochs, file, walker {i}"
    got = read_dump(prefix + "_stream_dev.0", False)
    for i, (_, n, verts) in enumerate(got):
        assert verts == paths[i, :lens[i]].tolist(), f"streamed epochs, device copies, walker {i}"
    # lists sorted on the host through csr, then walked on the device
    rp3, col3, w3 = O.csr_from_edges(34, src, dst, w, False, True)
    paths, lens, _, _ = O.walk(O.Graph(rp3, col3, w3), O.DEEPWALK, 34, 12, True, seed=1)
    got = read_dump(prefix + "_sorted_dw.0", True)
    for i, (wid, n, verts) in enumerate(got):
        assert n == lens[i] and verts == paths[i, :lens[i]].tolist(), f"sorted deepwalk walker {i}"
