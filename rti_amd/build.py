"""Builds librti_hip.so (gfx950) in-tree with hipcc.

Usage: python -m rti_amd.build   (or rti_amd.build.build()).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "librti_hip.so")
SOURCES = ["build.hip", "trace.hip", "post.hip", "api.hip"]
# The neighbour predicate and the exposed areas restate FP32 reference
# arithmetic exactly: no fused multiply-add in that translation unit (the
# compiler otherwise fuses the sum of squares even through __fmul_rn).
EXTRA = {"build.hip": ["-ffp-contract=off"]}
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay disabled in all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def build(verbose=False, jobs=None):
    objs = []
    procs = []
    for src in SOURCES:
        obj = os.path.join(CSRC, src.replace(".hip", ".o"))
        cmd = [HIPCC, *FLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # hipcc's host pass does not know the packed-fp32 target feature
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            failed.append(src)
            print("\n".join(lines), file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


CLI_SRC = os.path.join(HERE, "cli", "rti_cli.cpp")
CLI = os.path.join(HERE, "cli", "rti_cli")


def build_cli(verbose=False):
    """The command-line tracer (host C++ over the C ABI, VTK XML I/O, zlib)."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I" + inc, CLI_SRC, "-o", CLI + ".tmp", "-L" + HERE, "-lrti_hip",
           "-lz", "-Wl,-rpath,$ORIGIN/.."]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(CLI + ".tmp", CLI)
    return CLI


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_cli(verbose="-v" in sys.argv))
