#!/usr/bin/env python3
"""Builds libmdrl.so in place (sc2019_md_amd/libmdrl.so) for gfx950 with hipcc.

    python sc2019_md_amd/build.py

Every HIP source is compiled with -ffp-contract=off (fused operations are explicit) and
with packed FP32 disabled; the host sources go through the same hipcc line.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmdrl.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["mdrl_system.hip", "mdrl_kernels.hip", "mdrl_noalloc.hip", "mdrl_api.cpp", "mdrl_io.cpp",
           "mdrl_cells.cpp", "mdrl_gen.cpp"]

COMMON = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-munsafe-fp-atomics",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]


def _compile(src, objdir, verbose):
    obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC] + COMMON + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = [l for l in (r.stdout + r.stderr).splitlines()
            if "packed-fp32-ops" not in l and l.strip()]
    return src, obj, r.returncode, cmd, msgs


def build(verbose=False):
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        results = list(ex.map(lambda s: _compile(s, objdir, verbose), SOURCES))
    for src, obj, rc, cmd, msgs in results:
        if verbose:
            print(" ".join(cmd))
        if rc != 0:
            sys.stderr.write("\n".join(msgs) + "\n")
            raise RuntimeError("hipcc failed on %s" % src)
        if verbose and msgs:
            print("\n".join(msgs))
        objs.append(obj)
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-ldl"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link of libmdrl.so failed")
    os.replace(tmp, OUT)
    build_host_example()
    return OUT


HOST_SRC = os.path.join(HERE, "examples", "host_c.c")
HOST_OUT = os.path.join(HERE, "examples", "host_c")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")


def build_host_example(out=HOST_OUT, cc=None):
    """The plain C host of examples/host_c.c, linked against libmdrl.so (found next to the
    package through an $ORIGIN-relative rpath, so the pair can move together)."""
    cc = cc or os.environ.get("CC", "gcc")
    cmd = [cc, "-std=c99", "-O2", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I", INCLUDE, HOST_SRC,
           "-o", out, "-L", HERE, "-lmdrl", "-Wl,-rpath,$ORIGIN/.." if out == HOST_OUT else "-Wl,-rpath," + HERE]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build of %s failed" % os.path.basename(out))
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
