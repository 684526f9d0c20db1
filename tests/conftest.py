"""Shared fixtures.  GPU tests are marked ``gpu``; everything else runs on a
machine without a GPU.  The oracle under oracle/ is the checker of every
parity test and is never used to produce a tested value."""
import ctypes
import json
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "primesieve-gpu_amd")
GOLDEN = os.path.join(ROOT, "tests", "golden")
ORACLE_LIB = os.path.join(ROOT, "oracle", "libpsg_oracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libprimesieve_ref.so")

if PKG not in sys.path:
    sys.path.insert(0, PKG)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


class Oracle:
    """ctypes view of oracle/libpsg_oracle.so (test infrastructure)."""

    def __init__(self, path=ORACLE_LIB):
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: make -C oracle")
        L = ctypes.CDLL(path)
        u64 = ctypes.c_uint64
        L.psg_oracle_count6.argtypes = [u64, u64, ctypes.POINTER(u64), ctypes.c_int]
        L.psg_oracle_count6.restype = None
        L.psg_oracle_brute_count.argtypes = [u64, u64, ctypes.c_int]
        L.psg_oracle_brute_count.restype = u64
        L.psg_oracle_ktable.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)]
        L.psg_oracle_ktable.restype = None
        L.psg_oracle_division_magic.argtypes = [ctypes.c_uint32]
        L.psg_oracle_division_magic.restype = u64
        L.psg_oracle_crossings.argtypes = [u64, u64, ctypes.c_uint32, ctypes.c_uint32]
        L.psg_oracle_crossings.restype = u64
        L.psg_oracle_crossings_brute.argtypes = [u64, u64, ctypes.c_uint32, ctypes.c_uint32]
        L.psg_oracle_crossings_brute.restype = u64
        L.psg_oracle_isqrt.argtypes = [u64]
        L.psg_oracle_isqrt.restype = u64
        self.L = L

    def count6(self, start, stop, threads=8):
        out = (ctypes.c_uint64 * 6)()
        self.L.psg_oracle_count6(start, stop, out, threads)
        return list(out)

    def brute(self, start, stop, k):
        return self.L.psg_oracle_brute_count(start, stop, k)

    def ktable(self, t):
        out = (ctypes.c_uint8 * 256)()
        self.L.psg_oracle_ktable(t, out)
        return list(out)

    def magic(self, p):
        return self.L.psg_oracle_division_magic(p)

    def crossings(self, lo, hi, pmin, pmax):
        return self.L.psg_oracle_crossings(lo, hi, pmin, pmax)

    def crossings_brute(self, lo, hi, pmin, pmax):
        return self.L.psg_oracle_crossings_brute(lo, hi, pmin, pmax)


class Reference:
    """The reference CPU library built from its sources (oracle/_ref)."""

    NAMES = ["primes", "twins", "triplets", "quadruplets", "quintuplets", "sextuplets"]

    def __init__(self, path=REF_LIB):
        L = ctypes.CDLL(path)
        self.fns = []
        for n in self.NAMES:
            f = getattr(L, "primesieve_count_" + n)
            f.restype = ctypes.c_uint64
            f.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
            self.fns.append(f)
        L.primesieve_set_num_threads.argtypes = [ctypes.c_int]
        self.L = L

    def count(self, a, b, t):
        return self.fns[t](a, b)


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


@pytest.fixture(scope="session")
def reference():
    if not os.path.exists(REF_LIB):
        pytest.skip("oracle/_ref not built (reference sources absent)")
    return Reference()


@pytest.fixture(scope="session")
def golden():
    def load(name):
        with open(os.path.join(GOLDEN, name)) as f:
            return json.load(f)
    return load


@pytest.fixture(scope="session")
def psgpu():
    import psgpu as m
    m.lib()          # raises if the library is not built
    return m


@pytest.fixture(scope="session")
def gpu(psgpu):
    """The HIP backend on cuda:0 -- fails loudly if the device or the
    gfx950 code object cannot be used."""
    if not psgpu.available():
        pytest.fail("no usable GPU: " + psgpu.last_error())
    desc = psgpu.active_device()
    if not desc:
        pytest.fail("GPU backend could not be created: " + psgpu.last_error())
    print("device:", desc.splitlines()[0], flush=True)
    return psgpu
