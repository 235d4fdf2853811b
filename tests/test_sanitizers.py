"""The oracle and the facade's host-only code built with
-fsanitize=address,undefined (gcc, no GPU) and run on small inputs: no
memory error, no undefined behaviour (UBSan aborts on the first report)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="2")


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, (" ".join(cmd)) + "\n" + r.stdout + r.stderr
    return r


def test_oracle_under_asan_ubsan(tmp_path):
    o = str(tmp_path)
    inc = ["-I", os.path.join(ROOT, "oracle"), "-I", os.path.join(ROOT, "include")]
    run(["gcc", *SAN, "-fopenmp", "-ffp-contract=off", *inc, "-c", os.path.join(ROOT, "oracle", "kkw_oracle.c"),
         "-o", o + "/orc.o"])
    prog = os.path.join(ROOT, "knightking_amd", "programs", "node2vec_query.cpp")
    run(["g++", *SAN, "-fopenmp", "-ffp-contract=off", "-std=c++17", *inc, f'-DKKW_PROGRAM_SOURCE="{prog}"',
         "-c", os.path.join(ROOT, "oracle", "program_oracle.cpp"), "-o", o + "/prog.o"])
    run(["gcc", *SAN, "-DWITH_PROGRAM", *inc, "-c", os.path.join(ROOT, "tests", "c", "sanitize_driver.c"),
         "-o", o + "/drv.o"])
    run(["g++", *SAN, "-fopenmp", o + "/drv.o", o + "/orc.o", o + "/prog.o", "-o", o + "/drv"])
    r = run([o + "/drv"], env=ENV, timeout=300)
    assert "ok" in r.stdout


def test_facade_host_code_under_asan_ubsan(tmp_path):
    o = str(tmp_path)
    run(["g++", *SAN, "-std=c++11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
         os.path.join(ROOT, "tests", "cpp", "sanitize_facade.cpp"), "-o", o + "/fac"])
    r = run([o + "/fac", o], env=ENV, timeout=120)
    assert "ok" in r.stdout
    assert open(o + "/paths.txt").read() == "0 3 1 2 3\n1 2 4 5\n"
