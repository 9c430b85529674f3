"""Builds libvfr.so (HIP kernels for gfx950 + host C++), the vfRay CLI, and the
CPU oracle.  Run as ``python -m vfray_amd.build``; every call recompiles.

Compiler flags that matter for results:
  -ffp-contract=off   no FMA contraction, so every float operation that decides
                      an entry face, exit face or step count is the IEEE
                      operation the CPU oracle performs;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions in device code.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "vfray_amd")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(ROOT, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_lib(verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    jobs = [
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "vfr_kernels.hip"), "-o", os.path.join(OBJ, "vfr_kernels.o")],
        [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, "vfr_api.hip"), "-o", os.path.join(OBJ, "vfr_api.o")],
        ["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, "vfr_host.cpp"), "-o", os.path.join(OBJ, "vfr_host.o")],
    ]
    for extra in ("vfr_server.hip", "vfr_noalloc.hip"):
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, extra),
                     "-o", os.path.join(OBJ, extra.replace(".hip", ".o"))])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(_run, jobs))
    objs = [j[-1] for j in jobs]
    lib = os.path.join(PKG, "libvfr.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, lib)
    cli = os.path.join(PKG, "vfRay")
    _run(["g++", *CXX_FLAGS, "-o", cli + ".tmp", os.path.join(CSRC, "vfray_main.cpp"),
          f"-L{PKG}", "-lvfr", "-Wl,-rpath,$ORIGIN"])
    os.replace(cli + ".tmp", cli)
    ctf = os.path.join(PKG, "cleanTF")
    _run(["g++", *CXX_FLAGS, "-o", ctf + ".tmp", os.path.join(CSRC, "cleantf_main.cpp"),
          f"-L{PKG}", "-lvfr", "-Wl,-rpath,$ORIGIN"])
    os.replace(ctf + ".tmp", ctf)
    if verbose:
        print(f"built {lib} and {cli}")
    return lib


def build_oracle(verbose: bool = True) -> str:
    """CPU oracle (test infrastructure), and the reference CPU path under
    oracle/_ref when the reference sources are present."""
    odir = os.path.join(ROOT, "oracle")
    r = subprocess.run(["make", "-s", "-C", odir, "oracle"], stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed\n{r.stdout}")
    if os.path.isdir("/root/reference"):
        r = subprocess.run(["make", "-s", "-C", odir, "ref"], stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"reference build failed\n{r.stdout}")
    if verbose:
        print("built oracle")
    return os.path.join(odir, "liboracle.so")


def build(verbose: bool = True) -> None:
    build_lib(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build()
