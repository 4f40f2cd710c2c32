"""Seeded synthetic inputs of the benchmark workloads (libcucm_datagen.so)."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dg = None


def _lib():
    global _dg
    if _dg is None:
        path = os.path.join(HERE, "libcucm_datagen.so")
        if not os.path.exists(path):
            raise RuntimeError("libcucm_datagen.so is not built (python -m cucm_amd.build)")
        L = ctypes.CDLL(path)
        vp, u64, u32, d = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_double
        L.cucm_gen_markov2.argtypes = [vp, u64, u64, d]
        L.cucm_gen_markov2_files.argtypes = [vp, vp, vp, u32, u64, d, ctypes.c_int]
        L.cucm_gen_mixed.argtypes = [vp, u64, u64, ctypes.c_int]
        L.cucm_gen_structured.argtypes = [vp, u64, u64, ctypes.c_int]
        L.cucm_gen_lognormal_sizes.argtypes = [vp, u32, d, d, u64, u64, u64]
        L.cucm_gen_poisson_arrivals.argtypes = [vp, u32, d, u64]
        _dg = L
    return _dg


def markov2(n, seed=1, s=1.1) -> np.ndarray:
    """Order-2 Markov bytes over 64 symbols, Zipf(s) rows (1 MiB / seed 1 = the small config)."""
    out = np.empty(n, dtype=np.uint8)
    _lib().cucm_gen_markov2(out.ctypes.data, n, seed, s)
    return out


def markov2_files(sizes, seed0=1000, s=1.1, threads=0):
    """Concatenated markov2 files, file i seeded seed0 + i; returns (arena, offsets)."""
    sizes = np.asarray(sizes, dtype=np.uint64)
    offs = np.zeros(len(sizes), dtype=np.uint64)
    if len(sizes) > 1:
        offs[1:] = np.cumsum(sizes)[:-1]
    out = np.empty(int(sizes.sum()), dtype=np.uint8)
    _lib().cucm_gen_markov2_files(out.ctypes.data, offs.ctypes.data, sizes.ctypes.data, len(sizes),
                                  seed0, s, threads)
    return out, offs


def mixed(n, seed=2, threads=0) -> np.ndarray:
    """60% order-3 Markov text, 25% LZ copies, 15% random 4 KiB blocks."""
    out = np.empty(n, dtype=np.uint8)
    _lib().cucm_gen_mixed(out.ctypes.data, n, seed, threads)
    return out


def structured(n, seed=4, threads=0) -> np.ndarray:
    """4 MiB regions: int32 random walks, float32 smooth fields, 16-byte Zipf records."""
    out = np.empty(n, dtype=np.uint8)
    _lib().cucm_gen_structured(out.ctypes.data, n, seed, threads)
    return out


def lognormal_sizes(count, median, sigma, lo, hi, seed=3) -> np.ndarray:
    out = np.empty(count, dtype=np.uint64)
    _lib().cucm_gen_lognormal_sizes(out.ctypes.data, count, median, sigma, lo, hi, seed)
    return out


def poisson_arrivals(count, rate, seed=5) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    _lib().cucm_gen_poisson_arrivals(out.ctypes.data, count, rate, seed)
    return out
