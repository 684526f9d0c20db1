"""ctypes binding of libprimesieve_gpu.so (include/primesieve_gpu.h).

The Python-facing mirror of the reference's GPU API
(include/primesieve/gpu.hpp:32-81): same names without the ``gpu_`` prefix
duplication, same argument meaning, and the same error behaviour --
``count_*`` raise ``PrimesieveError`` when no GPU is usable (the reference
throws primesieve_error), ``try_count`` returns None so that the caller can
fall back to a CPU sieve.

Nothing here computes a count: every count comes from the gfx950 kernels
through the C ABI.  A missing library is an error, never a silent fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libprimesieve_gpu.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "primesieve_gpu.h")

PSG_ERROR = (1 << 64) - 1
COUNT_NAMES = ["primes", "twins", "triplets", "quadruplets", "quintuplets", "sextuplets"]


class PrimesieveError(RuntimeError):
    """Raised where the reference throws primesieve_error."""


class Stats(ctypes.Structure):
    _fields_ = [
        ("kernel_ms", ctypes.c_double),
        ("fill_ms", ctypes.c_double),
        ("device_ms", ctypes.c_double),
        ("wall_ms", ctypes.c_double),
        ("segments", ctypes.c_uint64),
        ("bucket_entries", ctypes.c_uint64),
        ("launches", ctypes.c_uint32),
        ("sieve_bytes", ctypes.c_uint32),
        ("work_group_size", ctypes.c_uint32),
        ("sieving_primes", ctypes.c_uint32),
        ("bucket_retries", ctypes.c_uint32),
        ("graph", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Work(ctypes.Structure):
    _fields_ = [
        ("crossings", ctypes.c_uint64),
        ("words32", ctypes.c_uint64),
        ("crossings_big", ctypes.c_uint64),
        ("sieving_primes", ctypes.c_uint64),
    ]


class Roofline(ctypes.Structure):
    _fields_ = [
        ("and_conflict_free_gbs", ctypes.c_double),
        ("and_random_gbs", ctypes.c_double),
        ("and_conflict_free_gops", ctypes.c_double),
        ("and_random_gops", ctypes.c_double),
        ("lds_read_gbs", ctypes.c_double),
        ("hbm_copy_gbs", ctypes.c_double),
        ("compute_units", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
        ("or_conflict_free_gbs", ctypes.c_double),
        ("or_random_gbs", ctypes.c_double),
    ]


class Query(ctypes.Structure):
    _fields_ = [("start", ctypes.c_uint64), ("stop", ctypes.c_uint64),
                ("k", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


BATCH_PERSISTENT = 0
BATCH_PLAIN = 1
BATCH_GRAPH = 2


class Tuning(ctypes.Structure):
    _fields_ = [
        ("sieve_bytes", ctypes.c_uint32),
        ("work_group_size", ctypes.c_uint32),
        ("segments_per_launch", ctypes.c_uint32),
        ("slab_segments", ctypes.c_uint32),
        ("bucket_factor", ctypes.c_uint32),
        ("presieve", ctypes.c_uint32),
        ("bucket_margin", ctypes.c_uint32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_lib = None


def _declare(lib):
    u64, i32, sz = ctypes.c_uint64, ctypes.c_int, ctypes.c_size_t
    cp = ctypes.c_char_p
    sig = {
        "psg_gpu_available": (i32, []),
        "psg_gpu_device_count": (i32, []),
        "psg_gpu_device_describe": (i32, [i32, cp, sz]),
        "psg_set_gpu_device": (None, [i32]),
        "psg_get_gpu_device": (i32, []),
        "psg_set_gpu_enabled": (None, [i32]),
        "psg_get_gpu_enabled": (i32, []),
        "psg_set_gpu_sieve_size": (None, [i32]),
        "psg_set_gpu_work_group_size": (None, [i32]),
        "psg_set_gpu_segments_per_launch": (None, [i32]),
        "psg_set_gpu_presieve": (None, [i32]),
        "psg_gpu_active_device": (i32, [cp, sz]),
        "psg_gpu_last_error": (i32, [cp, sz]),
        "psg_gpu_worthwhile": (i32, [u64, u64]),
        "psg_gpu_try_count": (i32, [u64, u64, i32, ctypes.POINTER(u64)]),
        "psg_gpu_count_all": (i32, [u64, u64, ctypes.POINTER(u64)]),
        "psg_gpu_count_all_noalloc": (i32, [u64, u64, ctypes.POINTER(u64)]),
        "psg_gpu_last_stats": (i32, [ctypes.POINTER(Stats)]),
        "psg_algorithmic_work": (i32, [u64, u64, ctypes.c_uint32, ctypes.POINTER(Work)]),
        "psg_measure_roofline": (i32, [ctypes.POINTER(Roofline)]),
        "psg_set_gpu_tuning": (i32, [ctypes.POINTER(Tuning)]),
        "psg_gpu_count_batch": (i32, [ctypes.POINTER(Query), sz, ctypes.POINTER(u64), i32]),
        "psg_mlp_param_count": (i32, []),
        "psg_mlp_create": (ctypes.c_void_p, []),
        "psg_mlp_destroy": (None, [ctypes.c_void_p]),
        "psg_mlp_set_params": (i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
        "psg_mlp_get_params": (i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
        "psg_mlp_forward": (i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), sz,
                                  ctypes.POINTER(ctypes.c_float)]),
        "psg_mlp_gradients": (i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_float), sz,
                                    ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
        "psg_mlp_adam_step": (i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float), ctypes.c_float]),
        "psg_mlp_train": (i32, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                                ctypes.POINTER(ctypes.c_float), sz, ctypes.c_uint32, ctypes.c_uint32,
                                ctypes.c_float, ctypes.c_uint32, ctypes.POINTER(ctypes.c_float)]),
        "psg_get_gpu_tuning": (i32, [ctypes.POINTER(Tuning)]),
        "psg_gpu_sieving_primes": (u64, [ctypes.c_uint32, ctypes.c_void_p, ctypes.c_void_p, u64]),
    }
    for n in COUNT_NAMES:
        sig["psg_gpu_count_" + n] = (u64, [u64, u64])
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


def lib(path=None):
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None or path is not None:
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise PrimesieveError("libprimesieve_gpu.so is not built (run __graft_entry__.build())")
        handle = _declare(ctypes.CDLL(p))
        if path is not None:
            return handle
        _lib = handle
    return _lib


def _string(fn, *args):
    buf = ctypes.create_string_buffer(4096)
    fn(*args, buf, len(buf))
    return buf.value.decode(errors="replace")


def available():
    return bool(lib().psg_gpu_available())


def devices():
    L = lib()
    return [_string(L.psg_gpu_device_describe, i) for i in range(L.psg_gpu_device_count())]


def set_device(index):
    lib().psg_set_gpu_device(int(index))


def set_enabled(enable):
    lib().psg_set_gpu_enabled(1 if enable else 0)


def set_sieve_size(kilobytes):
    lib().psg_set_gpu_sieve_size(int(kilobytes))


def set_work_group_size(size):
    lib().psg_set_gpu_work_group_size(int(size))


def set_segments_per_launch(n):
    lib().psg_set_gpu_segments_per_launch(int(n))


def set_presieve(enable):
    lib().psg_set_gpu_presieve(1 if enable else 0)


def set_tuning(sieve_bytes=0, work_group_size=0, segments_per_launch=0, slab_segments=0,
               bucket_factor=0, presieve=True, bucket_margin=0):
    """Set the whole tuning vector (0 = auto); recreates the backend."""
    t = Tuning(sieve_bytes, work_group_size, segments_per_launch, slab_segments,
               bucket_factor, 1 if presieve else 0, bucket_margin)
    if lib().psg_set_gpu_tuning(ctypes.byref(t)) != 0:
        raise PrimesieveError("psg_set_gpu_tuning failed")


def get_tuning():
    t = Tuning()
    if lib().psg_get_gpu_tuning(ctypes.byref(t)) != 0:
        raise PrimesieveError(last_error())
    return t.as_dict()


def active_device():
    return _string(lib().psg_gpu_active_device)


def last_error():
    return _string(lib().psg_gpu_last_error)


def worthwhile(start, stop):
    return bool(lib().psg_gpu_worthwhile(start, stop))


def count(start, stop, count_type=0):
    """gpu_count_<type>(start, stop); count_type 0..5 = primes..sextuplets."""
    fn = getattr(lib(), "psg_gpu_count_" + COUNT_NAMES[count_type])
    r = fn(start, stop)
    if r == PSG_ERROR:
        raise PrimesieveError(last_error() or "GPU count failed")
    return r


def count_all(start, stop):
    """All six counts of [start, stop] from one sieve pass."""
    out = (ctypes.c_uint64 * 6)()
    if lib().psg_gpu_count_all(start, stop, out) != 0:
        raise PrimesieveError(last_error() or "GPU count failed")
    return list(out)


def count_all_noalloc(start, stop):
    """count_all with no device allocation (psg_gpu_count_all_noalloc): the
    backend and its sieving-prime cache must exist already."""
    out = (ctypes.c_uint64 * 6)()
    if lib().psg_gpu_count_all_noalloc(start, stop, out) != 0:
        raise PrimesieveError(last_error() or "GPU count failed")
    return list(out)


def count_batch(queries, mode=BATCH_PERSISTENT):
    """queries: sequence of (start, stop, k) with k = 1..6, or an (n, 3)
    uint64 array.  Returns a list of counts."""
    n = len(queries)
    arr = (Query * n)()
    for i, (a, b, k) in enumerate(queries):
        arr[i].start, arr[i].stop, arr[i].k = int(a), int(b), int(k)
    out = (ctypes.c_uint64 * n)()
    if lib().psg_gpu_count_batch(arr, n, out, mode) != 0:
        raise PrimesieveError(last_error() or "batch count failed")
    return list(out)


def sieving_primes(limit):
    """The device cache's sieving primes <= limit (from 53; from 7 without
    the pre-sieve) and their division magics, as two numpy arrays (uint32,
    uint64) read back from the device.  Large caches are generated on the
    GPU."""
    import numpy as np
    L = lib()
    n = L.psg_gpu_sieving_primes(int(limit), None, None, 0)
    if n == PSG_ERROR:
        raise PrimesieveError(last_error() or "sieving primes failed")
    pr = np.zeros(n, dtype=np.uint32)
    mg = np.zeros(n, dtype=np.uint64)
    if n and L.psg_gpu_sieving_primes(int(limit), pr.ctypes.data, mg.ctypes.data, n) != n:
        raise PrimesieveError(last_error() or "sieving primes failed")
    return pr, mg


def try_count(start, stop, count_type=0):
    """gpu_try_count: the count, or None when the GPU was not used."""
    r = ctypes.c_uint64(0)
    if lib().psg_gpu_try_count(start, stop, count_type, ctypes.byref(r)):
        return r.value
    return None


def last_stats():
    s = Stats()
    if lib().psg_gpu_last_stats(ctypes.byref(s)) != 0:
        raise PrimesieveError("no backend")
    return s


def algorithmic_work(start, stop, seg_bytes):
    w = Work()
    if lib().psg_algorithmic_work(start, stop, seg_bytes, ctypes.byref(w)) != 0:
        raise PrimesieveError(last_error())
    return w


def lds_roofline():
    """Measured LDS / HBM ceilings of the active device (dict)."""
    r = Roofline()
    if lib().psg_measure_roofline(ctypes.byref(r)) != 0:
        raise PrimesieveError(last_error())
    return {name: getattr(r, name) for name, _ in r._fields_ if name != "reserved"}


def declared_symbols(header=HEADER):
    """Function names declared in include/primesieve_gpu.h."""
    import re
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psg_[a-z0-9_]+)\s*\(", text)))
