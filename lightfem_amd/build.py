"""In-tree build of the gfx950 library, the host C++ library and CLI, and the oracle.

``python -m lightfem_amd.build`` compiles every HIP source for gfx950 with hipcc
(objects in lightfem_amd/_build, shared library in lightfem_amd/lib), the host
C++ library and ``lightfem`` CLI, and the C oracle (``make -C oracle port``);
when the reference sources are present it also compiles them into oracle/_ref.
"""
from __future__ import annotations

import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
DEVICE = os.path.join(PKG, "csrc", "device")
HOST = os.path.join(PKG, "csrc", "host")
CLI = os.path.join(PKG, "csrc", "cli")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
GPU_LIB = os.path.join(LIB, "liblightfem_gpu.so")
HOST_LIB = os.path.join(LIB, "liblightfem.so")
CLI_BIN = os.path.join(BIN, "lightfem")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -ffp-contract=off keeps every multiply and add separate, as the CPU reference
# computes them, so element, assembly, elimination and preconditioner kernels are
# bit-identical to it. Packed FP32 instructions are disabled for all device code.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I", os.path.join(ROOT, "include"), "-I", DEVICE]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wextra",
              "-I", os.path.join(ROOT, "include"), "-I", HOST]


def _run(cmd: list[str]) -> None:
    result = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if result.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + result.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    if result.stdout.strip():
        # keep warnings visible but short
        lines = [ln for ln in result.stdout.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if lines:
            sys.stderr.write("\n".join(lines[:40]) + "\n")


def _sources(directory: str, exts: tuple[str, ...]) -> list[str]:
    if not os.path.isdir(directory):
        return []
    return sorted(os.path.join(directory, f) for f in os.listdir(directory) if f.endswith(exts))


def build_gpu() -> str:
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    jobs = []
    for src in _sources(DEVICE, (".hip", ".cpp")):
        obj = os.path.join(BUILD, os.path.basename(src) + ".o")
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        jobs.append(([HIPCC, *HIP_FLAGS, *lang, "-c", src, "-o", obj], obj))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as pool:
        list(pool.map(lambda job: _run(job[0]), jobs))
    tmp = GPU_LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *[obj for _, obj in jobs]])
    os.replace(tmp, GPU_LIB)
    return GPU_LIB


def build_host() -> None:
    """Host C++ library (model, parser, validator, pattern, RCM, colouring, solver, writers) and CLI."""
    srcs = _sources(HOST, (".cpp",))
    if not srcs:
        return
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    objs = []
    jobs = []
    for src in srcs:
        obj = os.path.join(BUILD, "host_" + os.path.basename(src) + ".o")
        objs.append(obj)
        jobs.append(["g++", *HOST_FLAGS, "-c", src, "-o", obj])
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as pool:
        list(pool.map(_run, jobs))
    tmp = HOST_LIB + ".tmp"
    _run(["g++", "-shared", "-fopenmp", "-o", tmp, *objs, "-L", LIB, "-llightfem_gpu",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64"])
    os.replace(tmp, HOST_LIB)
    cli_src = os.path.join(CLI, "main.cpp")
    if os.path.exists(cli_src):
        tmp = CLI_BIN + ".tmp"
        _run(["g++", *HOST_FLAGS, cli_src, "-o", tmp, "-L", LIB, "-llightfem", "-llightfem_gpu",
              "-Wl,-rpath,$ORIGIN/../lib", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
        os.replace(tmp, CLI_BIN)


def build_oracle() -> None:
    make = shutil.which("make") or "make"
    _run([make, "-C", os.path.join(ROOT, "oracle"), "port", "probe"])
    if os.path.isdir("/root/reference/src"):
        _run([make, "-C", os.path.join(ROOT, "oracle"), "-j8", "ref"])


def build_all() -> None:
    build_gpu()
    build_host()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", GPU_LIB)
