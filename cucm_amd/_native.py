"""ctypes binding of libcucm.so (include/cucm.h).

The library is required: importing the codec entry points without the built
HIP library raises CucmError; there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcucm.so")

OK = 0
E_PARAM, E_CORRUPT, E_SMALL, E_NOMEM, E_BUDGET, E_DEVICE, E_UNSUPPORTED, E_NOTREADY = range(-1, -9, -1)
FLAG_NO_TAIL_PAD = 0x1
FLAG_FROZEN = 0x2  # frozen model: the initial model predicts every bit, never updated
OP_COMPRESS, OP_DECOMPRESS = 0, 1

# Symbols declared in include/cucm.h (checked by the test suite).
EXPORTS = [
    "cucm_default_params", "cucm_strerror", "cucm_compress_bound", "cucm_peek_header",
    "cucm_compress", "cucm_decompress", "cucm_ctx_create", "cucm_ctx_destroy",
    "cucm_ctx_compress", "cucm_ctx_decompress", "cucm_compress_batch", "cucm_decompress_batch",
    "cucm_compress_many", "cucm_decompress_many", "cucm_compress_many_device",
    "cucm_decompress_many_device", "cucm_compress_plain_device", "cucm_decompress_plain_device",
    "cucm_ctx_malloc", "cucm_ctx_free", "cucm_ctx_memcpy_h2d", "cucm_ctx_memcpy_d2h",
    "cucm_ctx_synchronize", "cucm_ctx_last_stats", "cucm_ctx_set_max_waves",
    "cucm_ctx_set_phase_profile", "cucm_ctx_phase_cycles",
    "cucm_serve_start", "cucm_serve_start_ex", "cucm_serve_stop", "cucm_submit", "cucm_poll", "cucm_serve_launches",
    "cucm_ctx_set_waves_per_block", "cucm_ctx_set_team", "cucm_container_head_len", "cucm_stream_compress_begin", "cucm_stream_compress_next",
    "cucm_stream_decompress_begin", "cucm_stream_decompress_next", "cucm_stream_end",
    "cucm_compress_device_noalloc",
]


class CucmError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        msg = "%s: %s (%d)" % (what, _strerror(code), code) if what else "%s (%d)" % (_strerror(code), code)
        super().__init__(msg)


class Params(ctypes.Structure):
    _fields_ = [
        ("chunk_size", ctypes.c_uint32),
        ("unit_bits", ctypes.c_uint16),
        ("batch_chunks", ctypes.c_uint16),
        ("prelearn_size", ctypes.c_uint32),
        ("table_bits", ctypes.c_uint8),
        ("spec_bits", ctypes.c_uint8),
        ("reserved", ctypes.c_uint16),
        ("flags", ctypes.c_uint32),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("wall_ms", ctypes.c_double),
        ("codec_ms", ctypes.c_double),
        ("assemble_ms", ctypes.c_double),
        ("h2d_ms", ctypes.c_double),
        ("d2h_ms", ctypes.c_double),
        ("tasks", ctypes.c_uint64),
        ("waves", ctypes.c_uint64),
        ("launches", ctypes.c_uint64),
        ("in_bytes", ctypes.c_uint64),
        ("out_bytes", ctypes.c_uint64),
        ("model_bytes", ctypes.c_uint64),
        ("teams", ctypes.c_uint64),
        ("spec_bits", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Job(ctypes.Structure):
    _fields_ = [
        ("op", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("inp", ctypes.c_void_p),
        ("in_len", ctypes.c_uint64),
        ("params", Params),
        ("out", ctypes.c_void_p),
        ("out_cap", ctypes.c_uint64),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int32),
        ("batch_jobs", ctypes.c_int32),
        ("out_len", ctypes.c_uint64),
        ("queue_ms", ctypes.c_double),
        ("latency_ms", ctypes.c_double),
    ]


_lib = None


def lib():
    """Loads libcucm.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CucmError(E_UNSUPPORTED, "libcucm.so is not built (python -m cucm_amd.build)")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, u32, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_int32
        P = ctypes.POINTER
        L.cucm_default_params.argtypes = [P(Params)]
        L.cucm_default_params.restype = None
        L.cucm_strerror.argtypes = [ctypes.c_int]
        L.cucm_strerror.restype = ctypes.c_char_p
        L.cucm_compress_bound.argtypes = [u64, P(Params)]
        L.cucm_compress_bound.restype = u64
        L.cucm_peek_header.argtypes = [vp, u64, P(u64), P(Params)]
        L.cucm_compress.argtypes = [vp, u64, P(Params), vp, u64, P(u64)]
        L.cucm_decompress.argtypes = [vp, u64, vp, u64, P(u64)]
        L.cucm_ctx_create.argtypes = [P(vp), ctypes.c_int, u64]
        L.cucm_ctx_destroy.argtypes = [vp]
        L.cucm_ctx_destroy.restype = None
        L.cucm_ctx_compress.argtypes = [vp, vp, u64, P(Params), vp, u64, P(u64)]
        L.cucm_ctx_decompress.argtypes = [vp, vp, u64, vp, u64, P(u64)]
        L.cucm_compress_batch.argtypes = [vp, vp, u64, u32, u32, u32, vp, u64, P(u64)]
        L.cucm_decompress_batch.argtypes = [vp, vp, u64, vp, u64, u32, u32, u32, u32]
        many = [vp, u32, P(vp), P(u64), P(Params), P(vp), P(u64), P(u64), P(i32)]
        L.cucm_compress_many.argtypes = many
        L.cucm_compress_many_device.argtypes = many
        dmany = [vp, u32, P(vp), P(u64), P(vp), P(u64), P(u64), P(i32), u32]
        L.cucm_decompress_many.argtypes = dmany
        L.cucm_decompress_many_device.argtypes = dmany
        L.cucm_compress_plain_device.argtypes = [vp, u32, P(vp), P(u64), P(Params), P(vp), P(u64), P(u64)]
        L.cucm_decompress_plain_device.argtypes = [vp, u32, P(vp), P(u64), P(vp), P(u64), P(u64), u32]
        L.cucm_ctx_malloc.argtypes = [vp, P(vp), u64]
        L.cucm_ctx_free.argtypes = [vp, vp]
        L.cucm_ctx_memcpy_h2d.argtypes = [vp, vp, vp, u64]
        L.cucm_ctx_memcpy_d2h.argtypes = [vp, vp, vp, u64]
        L.cucm_ctx_synchronize.argtypes = [vp]
        L.cucm_ctx_last_stats.argtypes = [vp, P(Stats)]
        L.cucm_ctx_set_max_waves.argtypes = [vp, u32]
        L.cucm_ctx_set_phase_profile.argtypes = [vp, ctypes.c_int]
        L.cucm_ctx_phase_cycles.argtypes = [vp, P(u64)]
        L.cucm_serve_start.argtypes = [vp, u32, u32]
        L.cucm_serve_start_ex.argtypes = [vp, u32, u32, u32]
        L.cucm_serve_launches.argtypes = [vp, P(u64)]
        L.cucm_serve_stop.argtypes = [vp]
        L.cucm_submit.argtypes = [vp, P(Job), P(u64)]
        L.cucm_poll.argtypes = [vp, u64, ctypes.c_int, P(Result)]
        L.cucm_container_head_len.argtypes = [vp, u64, P(u64)]
        L.cucm_ctx_set_waves_per_block.argtypes = [vp, u32]
        L.cucm_ctx_set_team.argtypes = [vp, u32]
        L.cucm_stream_compress_begin.argtypes = [vp, u64, P(Params), vp, vp, u64, P(u64), P(vp)]
        L.cucm_stream_compress_next.argtypes = [vp, vp, u64, vp, u64, P(u64)]
        L.cucm_stream_decompress_begin.argtypes = [vp, vp, u64, u32, P(u64), P(Params), vp, u64,
                                                   P(vp)]
        L.cucm_stream_decompress_next.argtypes = [vp, vp, u64, vp, u64, P(u64)]
        L.cucm_compress_device_noalloc.argtypes = [vp, vp, u64, P(Params), vp, u64, P(u64)]
        L.cucm_stream_end.argtypes = [vp]
        L.cucm_stream_end.restype = None
        _lib = L
    return _lib


PROBE_PATH = os.path.join(HERE, "libcucm_probe.so")
_probe = None


def probe_lib():
    """Loads libcucm_probe.so, the measurement kernels bench.py compares the
    codec with (not part of the codec's C ABI)."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_PATH):
            raise CucmError(E_UNSUPPORTED, "libcucm_probe.so is not built (python -m cucm_amd.build)")
        L = ctypes.CDLL(PROBE_PATH)
        vp, u64, u32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32
        P = ctypes.POINTER
        L.cucm_probe_gather.argtypes = [vp, u64, u32, u32, u32, u32, u32, P(ctypes.c_double), P(u64)]
        L.cucm_probe_copy.argtypes = [vp, vp, u64, P(ctypes.c_double)]
        L.cucm_probe_replay_decode.argtypes = [vp, vp, u32, u32, u32, u32, u32, u32,
                                               P(ctypes.c_double)]
        L.cucm_probe_replay_encode.argtypes = [vp, vp, u32, u32, u32, u32, u32,
                                               P(ctypes.c_double)]
        _probe = L
    return _probe


def _strerror(code):
    try:
        return lib().cucm_strerror(code).decode()
    except Exception:
        return "error"


def check(rc, what=""):
    if rc != 0:
        raise CucmError(rc, what)
    return rc
