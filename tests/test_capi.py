"""The C ABI on a machine without a GPU: the library loads, exports every
symbol include/primesieve_gpu.h declares, and its failure handling follows
the reference (src/gpu/GpuSieve.cpp:279-365, src/api-c.cpp:143-225) --
forced from the host with a stand-in HIP runtime, so no kernel runs."""
import os
import subprocess
import sys
import textwrap

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "primesieve-gpu_amd")
LIB = os.path.join(PKG, "libprimesieve_gpu.so")


def test_exports_every_declared_symbol(psgpu):
    declared = psgpu.declared_symbols()
    assert len(declared) >= 25
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [s for s in declared if s not in exported]
    assert not missing, missing


def test_library_does_not_link_hip(psgpu):
    """libamdhip64 is resolved at run time, never a link-time dependency."""
    out = subprocess.run(["readelf", "-d", LIB], capture_output=True, text=True).stdout
    assert "amdhip" not in out


def test_algorithmic_work_matches_oracle(psgpu, oracle):
    for a, b in [(0, 10 ** 6), (10 ** 9, 10 ** 9 + 10 ** 6), (12345, 98765)]:
        w = psgpu.algorithmic_work(a, b, 65536)
        lo = max(a, 7)
        assert w.crossings == oracle.crossings(lo, b, 53, 2 ** 32 - 1)
        assert w.crossings_big == oracle.crossings(lo, b, 8 * 65536 + 1, 2 ** 32 - 1)
    w = psgpu.algorithmic_work(0, 10 ** 10, 65536)
    assert w.words32 == 83333334          # ceil(bytes / 4) of [7, 1e10]
    assert w.crossings == 2661834318      # crossings by primes >= 53 (oracle)


@pytest.fixture(scope="module")
def fake_hip(tmp_path_factory):
    d = tmp_path_factory.mktemp("fakehip")
    so = str(d / "libfakehip.so")
    src = os.path.join(ROOT, "tests", "fake_hip", "fake_hip.c")
    subprocess.check_call(["gcc", "-shared", "-fPIC", "-O1", "-o", so, src])
    return so


def _run_with_fake(fake_hip, mode, body):
    code = "import sys; sys.path.insert(0, %r)\nimport psgpu as m\n" % PKG + textwrap.dedent(body)
    env = dict(os.environ, PSG_HIP_LIBRARY=fake_hip, FAKE_HIP_MODE=mode)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout + r.stderr


def test_no_device(fake_hip):
    out = _run_with_fake(fake_hip, "nodevice", """
        import ctypes
        assert not m.available()
        assert m.devices() == []
        assert m.active_device() == ""
        try:
            m.count(0, 10**9)
            raise SystemExit("count must fail without a GPU")
        except m.PrimesieveError as e:
            assert "no GPU device found" in str(e), e
        assert m.lib().psg_gpu_count_primes(0, 10**9) == m.PSG_ERROR
        m.set_enabled(True)
        assert m.try_count(0, 10**9) is None       # falls back, never raises
        assert "no GPU device found" in m.last_error()
        print("OK")
    """)
    assert "OK" in out


def test_code_object_load_failure(fake_hip):
    out = _run_with_fake(fake_hip, "noload", """
        assert m.available()                        # a device is listed ...
        assert m.active_device() == ""              # ... but no backend can be made
        assert "code object" in m.last_error(), m.last_error()
        m.set_enabled(True)
        assert m.try_count(0, 10**9) is None
        print("OK")
    """)
    assert "OK" in out


def test_device_failure_falls_back(fake_hip):
    """hipMalloc fails inside the backend: gpu_try_count reports 'not used',
    remembers the error and drops the backend; gpu_count_* raise."""
    out = _run_with_fake(fake_hip, "nomalloc", """
        m.set_enabled(True)
        assert m.try_count(0, 10**9) is None
        assert "fake out of memory" in m.last_error(), m.last_error()
        try:
            m.count_all(0, 10**9)
            raise SystemExit("count_all must fail")
        except m.PrimesieveError as e:
            assert "out of memory" in str(e)
        r = m.lib().psg_gpu_count_primes(0, 10**9)
        assert r == m.PSG_ERROR
        print("OK")
    """)
    assert "OK" in out


def test_worthwhile_threshold(psgpu):
    # reference src/gpu/GpuSieve.cpp:34-36, 272-277
    assert not psgpu.worthwhile(10, 5)
    assert not psgpu.worthwhile(0, 10 ** 8 - 1)
    assert psgpu.worthwhile(0, 10 ** 8)


def test_launch_failure_mid_count(fake_hip):
    """The backend is created, then a kernel launch fails: gpu_try_count
    returns 'not used' (CPU fallback) and records the error; the explicit
    gpu_count_* raise; PSG_ERROR from the C entry points."""
    out = _run_with_fake(fake_hip, "nolaunch", """
        assert m.active_device() != ""
        m.set_enabled(True)
        assert m.try_count(0, 10**9) is None
        assert "hipModuleLaunchKernel(psg_" in m.last_error(), m.last_error()
        try:
            m.count(0, 10**9)
            raise SystemExit("count must fail")
        except m.PrimesieveError as e:
            assert "hipModuleLaunchKernel(psg_" in str(e)
        assert m.lib().psg_gpu_count_twins(0, 10**9) == m.PSG_ERROR
        print("OK")
    """)
    assert "OK" in out
