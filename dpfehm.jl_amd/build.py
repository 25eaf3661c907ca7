"""Builds libdpfgpu.so in-tree for gfx950.

    python dpfehm.jl_amd/build.py

hipcc cross-compiles without a GPU; the library lands next to this file
(``dpfehm.jl_amd/libdpfgpu.so``) so it travels with the source tree.
"""
import os
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
LIB = os.path.join(HERE, "libdpfgpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# packed FP32 instructions are disabled for every device pass
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(ROOT, "include")]

SOURCES = [
    ("dpf_gpu.hip", True),
    ("host_setup.cpp", False),
    ("inspect.cpp", False),
]


def build(verbose=False):
    objdir = tempfile.mkdtemp(prefix="dpfgpu_build_")
    procs, objs = [], []
    for src, device in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        cmd = [HIPCC] + COMMON + ["-c", os.path.join(CSRC, src), "-o", obj]
        if device:
            cmd += ["--offload-arch=" + ARCH] + DEVICE_FLAGS
        else:
            cmd = ["g++"] + COMMON + ["-fopenmp", "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        objs.append(obj)
    failed = []
    for src, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((src, out.decode(errors="replace")))
        elif verbose and out:
            print(out.decode(errors="replace"))
    if failed:
        msg = "\n".join(f"--- {s}\n{o}" for s, o in failed)
        raise RuntimeError("libdpfgpu build failed:\n" + msg)
    tmp = LIB + ".tmp"
    # host setup uses OpenMP through libgomp (the runtime PyTorch also loads)
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lgomp"]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    os.rmdir(objdir)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
