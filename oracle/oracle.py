"""ctypes binding of the CPU oracle (oracle/cucm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the cucm_amd package.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "libcucm_oracle.so")

_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.c_char_p
        vp = ctypes.c_void_p
        u64 = ctypes.c_uint64
        u32 = ctypes.c_uint32
        i64 = ctypes.c_int64
        ci = ctypes.c_int
        L.orc_get_tables.argtypes = [vp, vp, vp, vp]
        L.orc_fnv1a32.argtypes = [u8p, ctypes.c_size_t]
        L.orc_fnv1a32.restype = u32
        L.orc_slots.argtypes = [u8p, u64, ci, vp]
        L.orc_coder_step.argtypes = [vp, vp, ci, ci, vp, vp]
        L.orc_model_new.argtypes = [ci]
        L.orc_model_new.restype = vp
        L.orc_model_free.argtypes = [vp]
        L.orc_model_copy.argtypes = [vp, vp]
        L.orc_model_state.argtypes = [vp, ci, u32]
        L.orc_model_state.restype = ctypes.c_uint8
        L.orc_model_statep.argtypes = [vp, ci, ci]
        L.orc_model_statep.restype = u32
        L.orc_model_weight.argtypes = [vp, ci]
        L.orc_model_weight.restype = ctypes.c_int32
        L.orc_model_predict.argtypes = [vp, ci, u32]
        L.orc_statemap_update.argtypes = [vp, ci, u32, ci]
        L.orc_mixer_predict.argtypes = [vp, ci, vp]
        L.orc_mixer_update.argtypes = [vp, ci, ci, ci, vp]
        L.orc_model_digest.argtypes = [vp]
        L.orc_model_digest.restype = u64
        L.orc_encode_chunk.argtypes = [u8p, u64, u64, u32, vp, ci, vp, u64]
        L.orc_encode_chunk.restype = i64
        L.orc_decode_chunk.argtypes = [u8p, u64, u64, u64, u32, vp, vp]
        L.orc_blob_bound.argtypes = [u64, u64]
        L.orc_blob_bound.restype = u64
        L.orc_encode_batch.argtypes = [u8p, u64, u64, u32, vp, ci, ci, vp, u64, vp]
        L.orc_encode_batch_literal.argtypes = [u8p, u64, u64, u32, vp, vp, u64, vp]
        L.orc_decode_batch.argtypes = [u8p, u64, u64, u64, u32, vp, ci, vp]
        L.orc_decode_batch_literal.argtypes = [u8p, u64, u64, u64, u32, vp, ci, vp, vp]
        L.orc_prelearn_encode.argtypes = [u8p, u64, vp, vp, u64]
        L.orc_prelearn_encode.restype = i64
        L.orc_prelearn_train.argtypes = [u8p, u64, vp]
        L.orc_prelearn_decode.argtypes = [u8p, u64, u64, vp, vp]
        L.orc_compress_bound.argtypes = [u64, u32, u32, u32]
        L.orc_compress_bound.restype = u64
        L.orc_compress.argtypes = [u8p, u64, u32, u32, u32, u32, ci, ci, ci, vp, u64, vp]
        L.orc_decompress.argtypes = [u8p, u64, ci, vp, u64, vp]
        L.orc_encode_chunk_frozen.argtypes = [u8p, u64, vp, vp, u64]
        L.orc_encode_chunk_frozen.restype = i64
        L.orc_decode_chunk_frozen.argtypes = [u8p, u64, u64, vp, vp]
        L.orc_encode_batch_frozen.argtypes = [u8p, u64, u64, vp, ci, vp, u64, vp]
        L.orc_decode_batch_frozen.argtypes = [u8p, u64, u64, u64, vp, ci, vp]
        L.orc_compress_frozen.argtypes = [u8p, u64, u32, u32, u32, u32, ci, ci, vp, u64, vp]
        L.orc_compress_files.argtypes = [vp, vp, vp, u32, u32, u32, u32, ci, ci, vp, vp, vp, vp]
        L.orc_decompress_files.argtypes = [vp, vp, vp, u32, ci, vp, vp, vp, vp]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        raise OracleError("%s failed with %d" % (what, rc))


def tables():
    L = lib()
    sq = (ctypes.c_int32 * 4096)()
    st = (ctypes.c_int32 * 4096)()
    dt = (ctypes.c_int32 * 1024)()
    nex = (ctypes.c_uint8 * 512)()
    L.orc_get_tables(sq, st, dt, nex)
    return list(sq), list(st), list(dt), bytes(nex)


def fnv1a32(data: bytes) -> int:
    return lib().orc_fnv1a32(data, len(data))


def slots(chunk: bytes, b: int, tb: int):
    out = (ctypes.c_uint32 * 8)()
    lib().orc_slots(chunk, b, tb, out)
    return list(out)


class Model:
    """A CuCM model: 8 StateMaps of 2^tb one-byte states, 8x256 statep, 80x8 weights."""

    def __init__(self, tb: int):
        self.tb = tb
        self.ptr = lib().orc_model_new(tb)
        if not self.ptr:
            raise OracleError("model alloc failed")

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().orc_model_free(self.ptr)
            self.ptr = None

    def copy(self):
        m = Model(self.tb)
        lib().orc_model_copy(m.ptr, self.ptr)
        return m

    def state(self, k, slot):
        return lib().orc_model_state(self.ptr, k, slot)

    def statep(self, k, s):
        return lib().orc_model_statep(self.ptr, k, s)

    def weight(self, i):
        return lib().orc_model_weight(self.ptr, i)

    def predict(self, k, slot):
        return lib().orc_model_predict(self.ptr, k, slot)

    def sm_update(self, k, slot, y):
        lib().orc_statemap_update(self.ptr, k, slot, y)

    def mix(self, cxt, smp):
        arr = (ctypes.c_int32 * 8)(*smp)
        return lib().orc_mixer_predict(self.ptr, cxt, arr)

    def mix_update(self, y, pr, cxt, smp):
        arr = (ctypes.c_int32 * 8)(*smp)
        lib().orc_mixer_update(self.ptr, y, pr, cxt, arr)

    def digest(self):
        return lib().orc_model_digest(self.ptr)


def coder_step(x1, x2, pr, y):
    a = ctypes.c_uint32(x1)
    b = ctypes.c_uint32(x2)
    buf = (ctypes.c_uint8 * 8)()
    n = ctypes.c_uint64(0)
    lib().orc_coder_step(ctypes.byref(a), ctypes.byref(b), pr, y, buf, ctypes.byref(n))
    return a.value, b.value, bytes(buf[: n.value])


def encode_chunk(chunk: bytes, S: int, U: int, model: Model, tail_pad=True, L=None):
    L_ = len(chunk) if L is None else L
    cap = 12 * S + 16
    out = ctypes.create_string_buffer(cap)
    n = lib().orc_encode_chunk(chunk, L_, S, U, model.ptr, 1 if tail_pad else 0, out, cap)
    if n < 0:
        raise OracleError("encode_chunk %d" % n)
    return out.raw[:n]


def decode_chunk(payload: bytes, L: int, S: int, U: int, model: Model) -> bytes:
    out = ctypes.create_string_buffer(max(L, 1))
    _check(lib().orc_decode_chunk(payload, len(payload), L, S, U, model.ptr, out), "decode_chunk")
    return out.raw[:L]


def encode_batch(data: bytes, S: int, U: int, model: Model, tail_pad=True, threads=0, literal=False):
    L = lib()
    cap = L.orc_blob_bound(len(data), S)
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_uint64(0)
    if literal:
        rc = L.orc_encode_batch_literal(data, len(data), S, U, model.ptr, out, cap, ctypes.byref(n))
    else:
        rc = L.orc_encode_batch(data, len(data), S, U, model.ptr, 1 if tail_pad else 0, threads, out,
                                cap, ctypes.byref(n))
    _check(rc, "encode_batch")
    return out.raw[: n.value]


def decode_batch(blob: bytes, N: int, S: int, U: int, model: Model, threads=0, literal_m=None):
    L = lib()
    out = ctypes.create_string_buffer(max(N, 1))
    if literal_m is None:
        _check(L.orc_decode_batch(blob, len(blob), N, S, U, model.ptr, threads, out), "decode_batch")
        return out.raw[:N]
    flags = ctypes.c_uint32(0)
    _check(L.orc_decode_batch_literal(blob, len(blob), N, S, U, model.ptr, literal_m, out,
                                      ctypes.byref(flags)), "decode_batch_literal")
    return out.raw[:N], flags.value


def compress(data: bytes, S: int, U: int = 128, batch_num: int = 100, P: int = 0, tb: int = 25,
             tail_pad=True, threads=0, frozen=False) -> bytes:
    """Container of `data`; with frozen=True every chunk is coded with the
    initial (prelearned or zero) model and no updates (orc_compress_frozen)."""
    L = lib()
    cap = L.orc_compress_bound(len(data), S, P, batch_num)
    if cap == 0:
        raise OracleError("bad params")
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_uint64(0)
    if frozen:
        rc = L.orc_compress_frozen(data, len(data), S, U, batch_num, P, tb, threads, out, cap,
                                   ctypes.byref(n))
    else:
        rc = L.orc_compress(data, len(data), S, U, batch_num, P, tb, 1 if tail_pad else 0,
                            threads, out, cap, ctypes.byref(n))
    _check(rc, "compress")
    return out.raw[: n.value]


def encode_chunk_frozen(chunk: bytes, model: Model) -> bytes:
    cap = 12 * len(chunk) + 16
    out = ctypes.create_string_buffer(cap)
    n = lib().orc_encode_chunk_frozen(chunk, len(chunk), model.ptr, out, cap)
    if n < 0:
        raise OracleError("encode_chunk_frozen %d" % n)
    return out.raw[:n]


def decode_chunk_frozen(payload: bytes, L: int, model: Model) -> bytes:
    out = ctypes.create_string_buffer(max(L, 1))
    _check(lib().orc_decode_chunk_frozen(payload, len(payload), L, model.ptr, out),
           "decode_chunk_frozen")
    return out.raw[:L]


def encode_batch_frozen(data: bytes, S: int, model: Model, threads=0) -> bytes:
    L = lib()
    cap = L.orc_blob_bound(len(data), S)
    out = ctypes.create_string_buffer(cap)
    n = ctypes.c_uint64(0)
    _check(L.orc_encode_batch_frozen(data, len(data), S, model.ptr, threads, out, cap,
                                     ctypes.byref(n)), "encode_batch_frozen")
    return out.raw[: n.value]


def decode_batch_frozen(blob: bytes, N: int, S: int, model: Model, threads=0) -> bytes:
    out = ctypes.create_string_buffer(max(N, 1))
    _check(lib().orc_decode_batch_frozen(blob, len(blob), N, S, model.ptr, threads, out),
           "decode_batch_frozen")
    return out.raw[:N]


def decompress(blob: bytes, threads=0) -> bytes:
    L = lib()
    N = int.from_bytes(blob[:4], "big") if len(blob) >= 4 else 0
    out = ctypes.create_string_buffer(max(N, 1))
    n = ctypes.c_uint64(0)
    _check(L.orc_decompress(blob, len(blob), threads, out, max(N, 1), ctypes.byref(n)), "decompress")
    return out.raw[: n.value]


def prelearn_encode(data: bytes, model: Model) -> bytes:
    cap = 12 * len(data) + 16
    out = ctypes.create_string_buffer(cap)
    n = lib().orc_prelearn_encode(data, len(data), model.ptr, out, cap)
    if n < 0:
        raise OracleError("prelearn %d" % n)
    return out.raw[:n]


def prelearn_train(data: bytes, model: Model):
    lib().orc_prelearn_train(data, len(data), model.ptr)


def prelearn_decode(payload: bytes, P: int, model: Model) -> bytes:
    out = ctypes.create_string_buffer(max(P, 1))
    _check(lib().orc_prelearn_decode(payload, len(payload), P, model.ptr, out), "prelearn_decode")
    return out.raw[:P]


def compress_files(arena, offs, lens, chunk_max, U=128, batch_num=100, tb=18, threads=0):
    """One container per file (chunk size min(len, chunk_max)), files coded in
    parallel, one file per task; returns (containers, offsets, lengths) with
    the containers packed in one numpy arena."""
    import numpy as np
    arena = np.ascontiguousarray(arena, dtype=np.uint8)
    n = len(lens)
    offs = np.ascontiguousarray(offs, dtype=np.uint64)
    lens = np.ascontiguousarray(lens, dtype=np.uint64)
    caps = np.array([compress_bound_one(int(l), chunk_max, batch_num) for l in lens], dtype=np.uint64)
    ooffs = np.concatenate([[0], np.cumsum(caps)])[:-1].astype(np.uint64)
    out = np.empty(int(caps.sum()) or 1, dtype=np.uint8)
    olens = np.zeros(n, dtype=np.uint64)
    _check(lib().orc_compress_files(arena.ctypes.data, offs.ctypes.data, lens.ctypes.data, n,
                                    chunk_max, U, batch_num, tb, threads, out.ctypes.data,
                                    ooffs.ctypes.data, caps.ctypes.data, olens.ctypes.data),
           "compress_files")
    return out, ooffs, olens


def decompress_files(blobs, ioffs, ilens, sizes, threads=0):
    """Inverse of compress_files: returns the files packed back to back."""
    import numpy as np
    n = len(ilens)
    sizes = np.ascontiguousarray(sizes, dtype=np.uint64)
    ooffs = np.concatenate([[0], np.cumsum(sizes)])[:-1].astype(np.uint64)
    out = np.empty(int(sizes.sum()) or 1, dtype=np.uint8)
    olens = np.zeros(n, dtype=np.uint64)
    ioffs = np.ascontiguousarray(ioffs, dtype=np.uint64)
    ilens = np.ascontiguousarray(ilens, dtype=np.uint64)
    _check(lib().orc_decompress_files(blobs.ctypes.data, ioffs.ctypes.data, ilens.ctypes.data, n,
                                      threads, out.ctypes.data, ooffs.ctypes.data,
                                      sizes.ctypes.data, olens.ctypes.data), "decompress_files")
    return out[: int(sizes.sum())]


def compress_bound_one(n, chunk_max, batch_num=100):
    S = max(1, min(n, chunk_max))
    return lib().orc_compress_bound(n, S, 0, batch_num)
