"""Compile libkkw.so (HIP kernels + C ABI) for gfx950, in-tree."""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
SOURCES = ["graph.hip", "walk.hip", "engine.hip", "program.hip"]
OUT = os.path.join(HERE, "libkkw.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")
# headers a walk program is compiled with at run time, embedded as text
PROGRAM_HEADERS = [("kkw_rng.h", os.path.join(CSRC, "kkw_rng.h")),
                   ("kkw_program.h", os.path.join(INCLUDE, "kkw_program.h")),
                   ("program_args.h", os.path.join(CSRC, "program_args.h")),
                   ("program_walk.h", os.path.join(CSRC, "program_walk.h"))]
PROGRAM_INC = os.path.join(CSRC, "program_sources.inc")

# -ffp-contract=off: every float operation rounds on its own, as in the CPU
# replay, so the accept/reject comparisons agree bit for bit.
# -packed-fp32-ops off: no packed fp32 instructions.
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.sep not in c or os.path.exists(c)):
            return c
    return "hipcc"


def write_program_sources() -> None:
    """program_sources.inc: the run-time compiled headers as string literals
    (name, text) for hiprtcCreateProgram."""
    parts = ["// generated by knightking_amd/build.py from the headers it names; do not edit",
             "static const char *const kProgramHeaderNames[] = {" +
             ", ".join(f'"{n}"' for n, _ in PROGRAM_HEADERS) + "};",
             "static const char *const kProgramHeaderTexts[] = {"]
    for _, path in PROGRAM_HEADERS:
        with open(path) as f:
            text = f.read()
        if ")KKWSRC\"" in text:
            raise RuntimeError(f"{path} holds the raw-string delimiter")
        parts.append('R"KKWSRC(' + text + ')KKWSRC",')
    parts.append("};")
    parts.append(f"static const int kProgramHeaderCount = {len(PROGRAM_HEADERS)};")
    with open(PROGRAM_INC + ".tmp", "w") as f:
        f.write("\n".join(parts) + "\n")
    os.replace(PROGRAM_INC + ".tmp", PROGRAM_INC)


def build_lib(verbose: bool = False) -> str:
    write_program_sources()
    cmd = [hipcc(), *HIPCC_FLAGS, "-I", INCLUDE, "-o", OUT + ".tmp", *[os.path.join(CSRC, s) for s in SOURCES],
           "-lhiprtc", "-lrocprofiler-sdk-roctx"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed ({r.returncode}):\n{msgs}")
    if verbose and msgs.strip():
        print(msgs, file=sys.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build_lib(verbose=True))
