"""Build libzgl.so (HIP kernels for gfx950 + C ABI + host producers) in-tree.

Usage: python -m zillumgl_amd.build   (or __graft_entry__.build())

Compile flags that matter for results:
  -ffp-contract=off   no fused multiply-add contraction: the kernels evaluate the
                      reference's expressions operation by operation so that the
                      CPU oracle reproduces them bit for bit;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions;
  correctly rounded FP32 division and square root (hipcc's default).
"""
import os
import shutil
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libzgl.so")
HOST = os.path.join(HERE, "host")
HOST_OUT = os.path.join(HERE, "libzgl_host.so")
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("ZGL_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-Wall"]


def _run(cmd):
    t0 = time.time()
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return time.time() - t0


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    jobs = []
    for src in ("zgl_kernels.hip", "zgl_persistent.hip", "zgl_adjoint.hip", "zgl_bvh.hip", "zgl_queue.hip", "zgl_wavefront.hip", "zgl_lightpath.hip", "zgl_reference.hip", "zgl_capi.hip"):
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in ("zgl_build.cpp",):
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    procs = [subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True) for j in jobs]
    failed = False
    for j, p in zip(jobs, procs):
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            failed = True
        elif verbose and out.strip():
            sys.stderr.write(out)
    if failed:
        raise RuntimeError("libzgl build failed")
    tmp = OUT + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, OUT)
    if verbose:
        print("built", OUT)
    build_host(verbose)
    return OUT


def build_host(verbose=True):
    """libzgl_host.so: the C++ host layer (zillumgl_amd/host), linked against libzgl.so
    and found next to it at run time."""
    objs = []
    for src in ("zgl_host.cpp", "zgl_host_capi.cpp"):
        obj = os.path.join(BUILD, src + ".o")
        _run(["g++", *CXX_FLAGS, "-c", os.path.join(HOST, src), "-o", obj])
        objs.append(obj)
    tmp = HOST_OUT + ".tmp"
    _run(["g++", "-shared", *objs, "-o", tmp, f"-L{HERE}", "-l:libzgl.so", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_OUT)
    if verbose:
        print("built", HOST_OUT)
    return HOST_OUT


if __name__ == "__main__":
    build()
