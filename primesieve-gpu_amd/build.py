"""Build the gfx950 HIP counting backend in-tree.

  1. hipcc --genco --offload-arch=gfx950 compiles each kernel file under
     csrc/kernels/ into a code object (build/*.hsaco);
  2. g++ compiles the host library (csrc/host/*.cpp) and links the code
     objects in through csrc/embed.S (.incbin);
  3. the result is primesieve-gpu_amd/libprimesieve_gpu.so, which resolves
     libamdhip64 at run time (dlopen), so it loads on machines without ROCm.

Usage: python primesieve-gpu_amd/build.py
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "build")
LIB = os.path.join(HERE, "libprimesieve_gpu.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CXX = os.environ.get("CXX", "g++")
ROCM_INC = "/opt/rocm/include"

# every HIP source is compiled with packed-FP32 instructions disabled
HIP_FLAGS = ["--genco", "--offload-arch=gfx950", "-O3", "-std=c++17",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
CXX_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-Wextra", "-Wno-unused-parameter",
             "-D__HIP_PLATFORM_AMD__=1", "-I" + os.path.join(ROOT, "include"),
             "-I" + os.path.join(CSRC, "boundary"), "-I" + ROCM_INC]

KERNELS = {"psg_sieve": "PSG_SIEVE_HSACO", "psg_bench": "PSG_BENCH_HSACO",
           "psg_mlp": "PSG_MLP_HSACO"}
# The segment kernel's queue ticket is an LDS atomic issued by one lane.  With
# LLVM's atomic optimizer the result is read back (and waited for) right after
# the atomic, which also waits for every LDS strike the wave issued before it;
# without it the wait moves to where the ticket is used, after the item's own
# strikes have been issued (0.07 ms per launch on [1e13, 1e13+1e11]).
KERNEL_FLAGS = {"psg_sieve": ["-mllvm", "-amdgpu-atomic-optimizer-strategy=None"]}
HOST_SOURCES = ["hip_api.cpp", "tables.cpp", "HipBackend.cpp", "GpuSieve.cpp", "capi.cpp",
                "roofline.cpp", "costmodel.cpp"]


def _run_all(cmds):
    procs = [(c, subprocess.Popen(c, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)) for c in cmds]
    failed = []
    for c, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        if p.returncode != 0:
            failed.append((c, text))
        elif text.strip():
            sys.stderr.write(text)
    if failed:
        for c, text in failed:
            sys.stderr.write("FAILED: %s\n%s\n" % (" ".join(c), text))
        raise RuntimeError("build failed")


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    if not os.path.exists(HIPCC) and shutil.which("hipcc") is None:
        raise RuntimeError("hipcc not found")
    hipcc = HIPCC if os.path.exists(HIPCC) else shutil.which("hipcc")

    cmds = []
    for name in KERNELS:
        src = os.path.join(CSRC, "kernels", name + ".hip")
        out = os.path.join(BUILD, name + ".hsaco")
        cmds.append([hipcc] + HIP_FLAGS + KERNEL_FLAGS.get(name, []) + ["-I" + CSRC, "-o", out, src])
    objs = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
        objs.append(obj)
        cmds.append([CXX] + CXX_FLAGS + ["-c", os.path.join(CSRC, "host", src), "-o", obj])
    _run_all(cmds)

    embed_obj = os.path.join(BUILD, "embed.o")
    defs = ['-D%s="%s"' % (macro, os.path.join(BUILD, name + ".hsaco"))
            for name, macro in KERNELS.items()]
    _run_all([[CXX, "-c", "-fPIC"] + defs + [os.path.join(CSRC, "embed.S"), "-o", embed_obj]])

    tmp = LIB + ".tmp"
    _run_all([[CXX, "-shared", "-o", tmp] + objs + [embed_obj, "-ldl", "-pthread",
              "-Wl,--no-undefined"]])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
