"""Build the native libraries of annie_amd in-tree.

libannie_hip.so   gfx950 kernels + host engine + C ABI (include/annie_hip.h)
libannie_host.so  the test protocol host (Fermat/Euler/Proth check, Gerbicz-Li,
                  checkpoints) over the C ABI, plus the `annie` command-line tool

Run as `python -m annie_amd.build` or through __graft_entry__.build().
"""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(ROOT)
CSRC = os.path.join(ROOT, "csrc")
HOST = os.path.join(ROOT, "host")
LIBDIR = os.path.join(ROOT, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
GMP_INC = "/opt/conda/include"
# The system GMP runtime (same 6.2.1 ABI as the conda header).
GMP_LIB = ["-L/usr/lib/x86_64-linux-gnu", "-l:libgmp.so.10"]

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _newer(target, sources):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(s) > t for s in sources)


def build_hip(force=True):
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libannie_hip.so")
    srcs = [os.path.join(CSRC, f) for f in ("kernels.hip", "engine.hip")]
    deps = srcs + [os.path.join(CSRC, f) for f in ("arith.h", "plan.h", "kernels.h")]
    deps.append(os.path.join(REPO, "include", "annie_hip.h"))
    if not force and not _newer(out, deps):
        return out
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", "-I" + GMP_INC, "-I" + os.path.join(REPO, "include"),
          *srcs, "-o", tmp, *GMP_LIB])
    os.replace(tmp, out)
    return out


def build_host(force=True):
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libannie_host.so")
    if not os.path.isdir(HOST):
        return None
    srcs = sorted(os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".cpp") and f != "main.cpp")
    if not srcs:
        return None
    deps = srcs + [os.path.join(HOST, f) for f in os.listdir(HOST) if f.endswith(".h")]
    if force or _newer(out, deps):
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall", "-I" + GMP_INC,
              "-I" + os.path.join(REPO, "include"), *srcs, "-o", tmp, *GMP_LIB,
              "-L" + LIBDIR, "-lannie_hip", "-Wl,-rpath,$ORIGIN"])
        os.replace(tmp, out)
    main = os.path.join(HOST, "main.cpp")
    exe = os.path.join(LIBDIR, "annie")
    if os.path.exists(main) and (force or _newer(exe, [main, out])):
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I" + GMP_INC, "-I" + os.path.join(REPO, "include"),
              main, "-o", exe, "-L" + LIBDIR, "-lannie_host", "-lannie_hip", *GMP_LIB,
              "-Wl,-rpath,$ORIGIN"])
    return out


def build(force=True):
    build_hip(force)
    build_host(force)


if __name__ == "__main__":
    build(force="--if-stale" not in sys.argv)
