"""The C-ABI library builds, loads and exports every entry point include/kkw.h
declares; without a GPU it refuses to create an engine (there is no CPU path)."""
import ctypes
import os
import re

import pytest

import knightking_amd as kk

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "kkw.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(kkw_[a-z_]+)\s*\(", text)))


def test_header_declares_the_python_mirror():
    assert declared_symbols() == sorted(kk.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = kk.load_library()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert lib.kkw_abi_version() == 3


def test_library_marks_its_ranges_for_the_profiler():
    """The marker-trace ranges are compiled in: libkkw imports the roctx range
    calls and carries the names of its ranges."""
    import subprocess
    so = os.path.join(ROOT, "knightking_amd", "libkkw.so")
    und = subprocess.run(["nm", "-D", "--undefined-only", so], capture_output=True, text=True, check=True).stdout
    for sym in ("roctxRangePushA", "roctxRangePop", "roctxRangeStartA", "roctxRangeStop"):
        assert re.search(r"\b" + sym + r"\b", und), sym
    blob = open(so, "rb").read()
    for name in (b"alias_tables", b"neighbour_sets", b"walk_launch", b"walk_wait", b"serve_flight"):
        assert name + b"\0" in blob, name


def test_no_gpu_means_no_engine():
    """Without a GPU the engine cannot be created (no CPU fallback).  torch is
    kept out of this process: its bundled HIP runtime beside libkkw's is what
    Engine() refuses (test below)."""
    try:
        eng = kk.Engine()
    except kk.KkwError as e:
        assert e.code == kk.KKW_E_DEVICE
        return
    eng.close()
    pytest.skip("a GPU is present")


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the POD structs have the sizes and offsets gcc gives the header's."""
    import subprocess
    fields = {"kkw_job": kk._Job, "kkw_paths": kk._Paths, "kkw_engine_opts": kk._Opts,
              "kkw_timing": kk._Timing, "kkw_serve_opts": kk._ServeOpts, "kkw_prog_job": kk._ProgJob}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "kkw.h"', "int main(void){"]
    for cname, py in fields.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for cname, py in fields.items():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert got == want


def test_two_hip_runtimes_are_refused_before_any_gpu_call():
    """/proc/self/maps listing torch's bundled HIP runtime beside the system's
    (libkkw loaded before torch initialised its GPU) makes Engine() and
    paths_torch refuse, naming both; one runtime mapped twice is fine."""
    sys_rt = "/opt/rocm-7.2.0/lib/libamdhip64.so.7.2.0"
    torch_rt = "/usr/local/lib/python3.10/dist-packages/torch/lib/libamdhip64.so"
    line = "7f0000000000-7f0000100000 r-xp 00000000 08:01 123 {}"
    one = "\n".join([line.format(sys_rt), line.format(sys_rt), line.format("/usr/lib/libc.so.6")])
    assert kk.hip_runtimes(one) == [sys_rt]
    kk.check_one_hip_runtime("test", one)
    two = one + "\n" + line.format(torch_rt)
    assert kk.hip_runtimes(two) == sorted([sys_rt, torch_rt])
    with pytest.raises(RuntimeError) as e:
        kk.check_one_hip_runtime("Engine()", two)
    assert sys_rt in str(e.value) and torch_rt in str(e.value)
