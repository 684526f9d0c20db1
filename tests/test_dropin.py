"""The drop-in check of the boundary, on the CPU: the reference's own count
API -- src/api.cpp and src/api-c.cpp built with PRIMESIEVE_ENABLE_GPU, and
src/gpu/GpuSieve.cpp with its backend dispatch given one HIP branch -- is
linked with this repository's HIP backend (HipBackend.cpp, hip_api.cpp,
tables.cpp, the gfx950 code object) compiled against the reference's OWN
src/gpu/GpuBackend.hpp, and with the reference CPU engine (oracle/_ref).  The
only edits to the reference are the ones a maintainer would make (shown in
INTEGRATION form 1), applied to copies in a temporary directory:

  * GpuBackend.hpp: declare listHipDevices / createHipBackend under
    PRIMESIEVE_ENABLE_HIP, beside the CUDA and OpenCL declarations;
  * GpuSieve.cpp: call listHipDevices in listDevices() and createHipBackend
    for a "HIP" device in createBackend(), under the same macro.

Under the stand-in HIP runtime (tests/fake_hip) the program must count
through primesieve_count_primes / _twins with the reference's policy:
gpu_try_count fails over to the CPU sieve (counts equal to the oracle's),
gpu_last_error says why, gpu_count_primes throws ("GPU is not available"
without a usable backend, the device error itself when a launch fails).
Skipped where the reference sources are absent (they never reach the GPU
machine)."""
import concurrent.futures
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = "/root/reference"
PKG = os.path.join(ROOT, "primesieve-gpu_amd")
HOST = os.path.join(PKG, "csrc", "host")
REF_OBJ = os.path.join(ROOT, "oracle", "_ref", "obj")

# the reference's CPU engine, as oracle/Makefile compiles it (api and api-c
# are rebuilt here with the GPU hooks on)
CPU_UNITS = ["CountPrintPrimes", "CpuInfo", "Erat", "EratSmall", "EratMedium", "EratBig",
             "iterator-c", "iterator", "IteratorHelper", "LookupTables", "MemoryPool",
             "PrimeGenerator", "nthPrime", "ParallelSieve", "popcount", "PreSieve",
             "PrimeSieveClass", "RiemannR", "SievingPrimes", "util", "cpuid"]

HIP_DECL = """#if defined(PRIMESIEVE_ENABLE_HIP)
void listHipDevices(std::vector<DeviceInfo>& out);
std::unique_ptr<GpuBackend> createHipBackend(const DeviceInfo& dev,
                                             const GpuConfig& config,
                                             std::string& error);
#endif

"""


def _patch(text, anchor, insert, before=True):
    assert text.count(anchor) == 1, "anchor not found once: %r" % anchor
    return text.replace(anchor, insert + anchor if before else anchor + insert)


@pytest.fixture(scope="module")
def dropin(tmp_path_factory, fake_hip_lib):
    if not os.path.isdir(os.path.join(REF, "src", "gpu")):
        pytest.skip("reference sources absent")
    embed = os.path.join(PKG, "build", "embed.o")
    if not os.path.exists(embed) or not os.path.isdir(REF_OBJ):
        pytest.skip("build first: __graft_entry__.build() (library and oracle/_ref)")
    d = tmp_path_factory.mktemp("dropin")
    gdir = d / "gpu"
    gdir.mkdir()
    hdr = open(os.path.join(REF, "src", "gpu", "GpuBackend.hpp")).read()
    hdr = _patch(hdr, "#if defined(PRIMESIEVE_ENABLE_OPENCL)\nvoid listOpenCLDevices", HIP_DECL)
    (gdir / "GpuBackend.hpp").write_text(hdr)
    src = open(os.path.join(REF, "src", "gpu", "GpuSieve.cpp")).read()
    src = _patch(src, "#if defined(PRIMESIEVE_ENABLE_CUDA)\n  listCudaDevices(out);",
                 "#if defined(PRIMESIEVE_ENABLE_HIP)\n  listHipDevices(out);\n#endif\n")
    src = _patch(src, "#if defined(PRIMESIEVE_ENABLE_CUDA)\n  if (dev.backend == \"CUDA\")",
                 "#if defined(PRIMESIEVE_ENABLE_HIP)\n  if (dev.backend == \"HIP\")\n"
                 "    return createHipBackend(dev, config, error);\n#endif\n")
    (gdir / "GpuSieve.cpp").write_text(src)

    flags = ["g++", "-O1", "-fPIC", "-std=c++17", "-DPRIMESIEVE_ENABLE_GPU", "-DPRIMESIEVE_ENABLE_HIP",
             "-D__HIP_PLATFORM_AMD__=1",
             # the reference's headers first: its GpuBackend.hpp, gpu.hpp and primesieve_error.hpp
             "-I" + str(gdir), "-I" + os.path.join(REF, "include"), "-I" + os.path.join(REF, "src"),
             # then this repository's: the C ABI header and the kernel ABI
             "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(PKG, "csrc"),
             "-I/opt/rocm/include"]
    units = [(str(gdir / "GpuSieve.cpp"), "GpuSieve.o"),
             (os.path.join(REF, "src", "api.cpp"), "api.o"),
             (os.path.join(REF, "src", "api-c.cpp"), "api-c.o"),
             (os.path.join(HOST, "HipBackend.cpp"), "HipBackend.o"),
             (os.path.join(HOST, "hip_api.cpp"), "hip_api.o"),
             (os.path.join(HOST, "tables.cpp"), "tables.o"),
             (os.path.join(ROOT, "tests", "dropin", "dropin_main.cpp"), "main.o")]
    cmds = [flags + ["-c", s, "-o", str(d / o)] for s, o in units]
    with concurrent.futures.ThreadPoolExecutor(8) as ex:
        for r in ex.map(lambda c: subprocess.run(c, capture_output=True, text=True), cmds):
            assert r.returncode == 0, r.stderr[-3000:]
    exe = str(d / "dropin")
    objs = [str(d / o) for _, o in units] + [os.path.join(REF_OBJ, u + ".o") for u in CPU_UNITS] + [embed]
    r = subprocess.run(["g++", "-o", exe] + objs + ["-ldl", "-pthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


@pytest.fixture(scope="module")
def fake_hip_lib(tmp_path_factory):
    d = tmp_path_factory.mktemp("fakehip_dropin")
    so = str(d / "libfakehip.so")
    subprocess.check_call(["gcc", "-shared", "-fPIC", "-O1", "-o", so,
                           os.path.join(ROOT, "tests", "fake_hip", "fake_hip.c")])
    return so


def _run(exe, fake, mode, a, b):
    env = dict(os.environ, PSG_HIP_LIBRARY=fake, FAKE_HIP_MODE=mode)
    r = subprocess.run([exe, str(a), str(b)], env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        out[k] = v
    return out


@pytest.mark.parametrize("mode,why", [("nodevice", "no GPU device found"),
                                      ("noload", "code object"),
                                      ("nolaunch", "hipModuleLaunchKernel(psg_")])
def test_reference_api_over_hip_backend_falls_back(dropin, fake_hip_lib, oracle, mode, why):
    a, b = 10 ** 9, 10 ** 9 + 2 * 10 ** 8          # above the 1e8 worthwhile threshold
    want = oracle.count6(a, b)
    out = _run(dropin, fake_hip_lib, mode, a, b)
    assert int(out["primes"]) == want[0]
    assert int(out["twins"]) == want[1]
    assert why in out["last_error"], out
    if mode == "nolaunch":
        # the backend exists and fails mid-count: gpu_count_primes rethrows
        # the device error itself (reference src/gpu/GpuSieve.cpp:279-286)
        assert why in out.get("gpu_count_threw", ""), out
    else:
        assert "GPU is not available" in out.get("gpu_count_threw", ""), out
    if mode == "nodevice":
        assert out["available"] == "0" and out["devices"] == "0"
    else:
        # the reference's listDevices() reached listHipDevices: the fake device is listed
        assert out["available"] == "1" and out["devices"] == "1"
