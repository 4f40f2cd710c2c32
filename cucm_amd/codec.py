"""Host-side mirror of CuCM's compress/decompress interface on the HIP codec.

Reference interface (host C++): compress_file(ifstream, ofstream, n, &out_n,
chunksize, unitlength, prelearn_size=0, batch_num=100)
(cucm_compress_v2.cu:1151-1154); decompress(ifstream, n, ofstream, out_n,
maxassumed_length=3) (cucm_decompress_v2.cu:1185; here maxassumed_length
defaults to an automatic width); compress_batch /
decompressbatch (cucm_compress_v2.cu:990-992, cucm_decompress_v2.cu:898-903).
Every call goes through libcucm.so's C ABI; there is no CPU fallback.
"""
import ctypes

import numpy as np

from ._native import (FLAG_FROZEN, FLAG_NO_TAIL_PAD, CucmError, Job, Params, Result, Stats, check,
                      lib)

MiB = 1 << 20


def make_params(chunk_size=MiB, unit_bits=128, batch_chunks=100, prelearn_size=0, table_bits=25,
                spec_bits=0, tail_pad=True, frozen=False):
    """cucm_params; frozen=True selects the frozen model (CUCM_FLAG_FROZEN)."""
    p = Params()
    lib().cucm_default_params(ctypes.byref(p))
    p.chunk_size = chunk_size
    p.unit_bits = unit_bits
    p.batch_chunks = batch_chunks
    p.prelearn_size = prelearn_size
    p.table_bits = table_bits
    p.spec_bits = spec_bits
    p.flags = (0 if tail_pad else FLAG_NO_TAIL_PAD) | (FLAG_FROZEN if frozen else 0)
    return p


def _buf(data):
    """(address, length, keepalive) of a bytes-like or numpy object."""
    if isinstance(data, np.ndarray):
        a = np.ascontiguousarray(data).view(np.uint8).reshape(-1)
        return a.ctypes.data, a.size, a
    b = bytes(data)
    cb = ctypes.create_string_buffer(b, len(b) or 1)
    return ctypes.addressof(cb), len(b), cb


def compress_bound(n, params):
    return lib().cucm_compress_bound(n, ctypes.byref(params))


def output_cap(n, params):
    """A container buffer size that real inputs fit: 1.5x the chunk-padded
    input (the last chunk is coded up to the full chunk size unless tail
    padding is off), or the strict bound when that is smaller."""
    S = max(params.chunk_size, 1)
    chunks = (max(n - params.prelearn_size, 0) + S - 1) // S
    padded = params.prelearn_size + chunks * S
    cap = padded + padded // 2 + 4096 * (2 + chunks)
    return min(cap, int(compress_bound(n, params)))


def peek_header(blob):
    addr, n, keep = _buf(blob)
    orig = ctypes.c_uint64(0)
    p = Params()
    check(lib().cucm_peek_header(addr, n, ctypes.byref(orig), ctypes.byref(p)), "peek_header")
    return orig.value, p


class DeviceBuffer:
    """Device memory on a context's GPU (cucm_ctx_malloc)."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        self.ptr = ctypes.c_void_p()
        check(lib().cucm_ctx_malloc(ctx.handle, ctypes.byref(self.ptr), max(self.nbytes, 1)), "malloc")

    @property
    def addr(self):
        return self.ptr.value

    def upload(self, data, offset=0):
        addr, n, keep = _buf(data)
        check(lib().cucm_ctx_memcpy_h2d(self.ctx.handle, self.addr + offset, addr, n), "h2d")

    def download(self, n=None, offset=0) -> bytes:
        n = self.nbytes - offset if n is None else n
        out = np.empty(n, dtype=np.uint8)
        check(lib().cucm_ctx_memcpy_d2h(self.ctx.handle, out.ctypes.data, self.addr + offset, n), "d2h")
        return out.tobytes()

    def free(self):
        if self.ptr.value:
            lib().cucm_ctx_free(self.ctx.handle, self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """A device context: one GPU, its stream, model slots and arenas."""

    def __init__(self, device=0, hbm_budget=0):
        self.handle = ctypes.c_void_p()
        check(lib().cucm_ctx_create(ctypes.byref(self.handle), device, hbm_budget), "ctx_create")

    def close(self):
        if self.handle and self.handle.value:
            lib().cucm_ctx_destroy(self.handle)
            self.handle = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_max_waves(self, waves):
        check(lib().cucm_ctx_set_max_waves(self.handle, waves))

    def set_waves_per_block(self, waves):
        check(lib().cucm_ctx_set_waves_per_block(self.handle, waves))

    def set_team(self, mode):
        """Team execution (one chunk per workgroup of four waves): 0 automatic,
        1 never, 2 always."""
        check(lib().cucm_ctx_set_team(self.handle, mode))

    PHASES = ["init", "stage", "hash+gather", "tables+mix", "coder", "mixer+collide", "resolve",
              "statep", "writeback", "dec_paths", "dec_mix+decode", "dec_records"]

    def set_phase_profile(self, on=True):
        check(lib().cucm_ctx_set_phase_profile(self.handle, 1 if on else 0))

    def phase_cycles(self) -> dict:
        arr = (ctypes.c_uint64 * 12)()
        check(lib().cucm_ctx_phase_cycles(self.handle, arr))
        return dict(zip(self.PHASES, list(arr)))

    def stats(self) -> dict:
        st = Stats()
        check(lib().cucm_ctx_last_stats(self.handle, ctypes.byref(st)))
        return st.as_dict()

    def synchronize(self):
        check(lib().cucm_ctx_synchronize(self.handle))

    def alloc(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    # ---- host buffers -------------------------------------------------
    def compress(self, data, params=None, **kw) -> bytes:
        return self.compress_many([data], params, **kw)[0]

    def decompress(self, blob, spec_bits=0) -> bytes:
        return self.decompress_many([blob], spec_bits)[0]

    def compress_many(self, datas, params=None, **kw):
        n = len(datas)
        if params is None:
            params = make_params(**kw)
        plist = params if isinstance(params, (list, tuple)) else [params] * n
        bufs = [_buf(d) for d in datas]
        ins = (ctypes.c_void_p * n)(*[b[0] for b in bufs])
        lens = (ctypes.c_uint64 * n)(*[b[1] for b in bufs])
        ps = (Params * n)(*plist)
        # realistic output fits; the library reports E_SMALL otherwise
        caps = [output_cap(b[1], p) for b, p in zip(bufs, plist)]
        outs = [np.empty(max(c, 1), dtype=np.uint8) for c in caps]
        outp = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
        capa = (ctypes.c_uint64 * n)(*caps)
        olen = (ctypes.c_uint64 * n)()
        st = (ctypes.c_int32 * n)()
        check(lib().cucm_compress_many(self.handle, n, ins, lens, ps, outp, capa, olen, st), "compress_many")
        res = []
        for i in range(n):
            if st[i] == -3:  # retry this one with the strict bound
                full = np.empty(int(compress_bound(bufs[i][1], plist[i])), dtype=np.uint8)
                one = self._compress_one(bufs[i], plist[i], full)
                res.append(one)
                continue
            check(st[i], "compress[%d]" % i)
            res.append(outs[i][: olen[i]].tobytes())
        return res

    def _compress_one(self, b, p, out):
        olen = ctypes.c_uint64(0)
        check(lib().cucm_ctx_compress(self.handle, b[0], b[1], ctypes.byref(p), out.ctypes.data,
                                      out.size, ctypes.byref(olen)), "compress")
        return out[: olen.value].tobytes()

    def decompress_many(self, blobs, spec_bits=0):
        n = len(blobs)
        bufs = [_buf(b) for b in blobs]
        ins = (ctypes.c_void_p * n)(*[b[0] for b in bufs])
        lens = (ctypes.c_uint64 * n)(*[b[1] for b in bufs])
        sizes = [int.from_bytes(bytes(b)[:4], "big") if len(b) >= 4 else 0 for b in blobs]
        outs = [np.empty(max(s, 1), dtype=np.uint8) for s in sizes]
        outp = (ctypes.c_void_p * n)(*[o.ctypes.data for o in outs])
        capa = (ctypes.c_uint64 * n)(*sizes)
        olen = (ctypes.c_uint64 * n)()
        st = (ctypes.c_int32 * n)()
        check(lib().cucm_decompress_many(self.handle, n, ins, lens, outp, capa, olen, st, spec_bits),
              "decompress_many")
        res = []
        for i in range(n):
            check(st[i], "decompress[%d]" % i)
            res.append(outs[i][: olen[i]].tobytes())
        return res

    # ---- batch records (compress_batch / decompressbatch) --------------
    def compress_batch(self, data, chunk_size, unit_bits=128, table_bits=25) -> bytes:
        addr, n, keep = _buf(data)
        C = (n + chunk_size - 1) // chunk_size
        strict = 4 * C + C * (32 * chunk_size + 1)
        # the last chunk is coded up to the full chunk size (tail padding), so
        # size from the padded length; a chunk that still overflows is
        # retried once with the strict bound
        padded = C * chunk_size
        cap = min(strict, padded + padded // 2 + 4096 * (C + 1))
        for attempt in (cap, strict):
            out = np.empty(attempt, dtype=np.uint8)
            olen = ctypes.c_uint64(0)
            rc = lib().cucm_compress_batch(self.handle, addr, n, chunk_size, unit_bits, table_bits,
                                           out.ctypes.data, attempt, ctypes.byref(olen))
            if rc == -3 and attempt < strict:
                continue
            check(rc, "compress_batch")
            return out[: olen.value].tobytes()

    def decompress_batch(self, blob, n, chunk_size, unit_bits=128, table_bits=25, spec_bits=0) -> bytes:
        addr, bl, keep = _buf(blob)
        out = np.empty(n, dtype=np.uint8)
        check(lib().cucm_decompress_batch(self.handle, addr, bl, out.ctypes.data, n, chunk_size,
                                          unit_bits, table_bits, spec_bits), "decompress_batch")
        return out.tobytes()

    # ---- device-resident -------------------------------------------------
    def compress_many_device(self, ins, lens, params, outs, caps, plain=False):
        """ins/outs: device addresses; returns (out_len list, status list)."""
        n = len(ins)
        plist = params if isinstance(params, (list, tuple)) else [params] * n
        a = (ctypes.c_void_p * n)(*ins)
        l = (ctypes.c_uint64 * n)(*lens)
        ps = (Params * n)(*plist)
        o = (ctypes.c_void_p * n)(*outs)
        c = (ctypes.c_uint64 * n)(*caps)
        olen = (ctypes.c_uint64 * n)()
        st = (ctypes.c_int32 * n)()
        if plain:
            check(lib().cucm_compress_plain_device(self.handle, n, a, l, ps, o, c, olen), "compress_plain")
        else:
            check(lib().cucm_compress_many_device(self.handle, n, a, l, ps, o, c, olen, st), "compress_device")
        return list(olen), list(st)

    def compress_device_noalloc(self, d_in, n, params, d_out, cap):
        """One device-resident input through the allocation-free entry point;
        returns the container length."""
        olen = ctypes.c_uint64(0)
        check(lib().cucm_compress_device_noalloc(self.handle, d_in, n, ctypes.byref(params), d_out,
                                                 cap, ctypes.byref(olen)), "compress_noalloc")
        return olen.value

    def decompress_many_device(self, ins, lens, outs, caps, spec_bits=0, plain=False):
        n = len(ins)
        a = (ctypes.c_void_p * n)(*ins)
        l = (ctypes.c_uint64 * n)(*lens)
        o = (ctypes.c_void_p * n)(*outs)
        c = (ctypes.c_uint64 * n)(*caps)
        olen = (ctypes.c_uint64 * n)()
        st = (ctypes.c_int32 * n)()
        if plain:
            check(lib().cucm_decompress_plain_device(self.handle, n, a, l, o, c, olen, spec_bits),
                  "decompress_plain")
        else:
            check(lib().cucm_decompress_many_device(self.handle, n, a, l, o, c, olen, st, spec_bits),
                  "decompress_device")
        return list(olen), list(st)

    # ---- streaming container I/O ----------------------------------------
    def compress_stream(self, data, params, segment_batches=1):
        """Compresses through the streaming C ABI (header, then segments of
        `segment_batches` batch records each); returns the container, which
        equals compress()'s."""
        addr, N, keep = _buf(data)
        P = params.prelearn_size
        head = np.empty(21 + P, dtype=np.uint8)
        hl = ctypes.c_uint64()
        s = ctypes.c_void_p()
        check(lib().cucm_stream_compress_begin(self.handle, N, ctypes.byref(params), addr,
                                               head.ctypes.data, head.size, ctypes.byref(hl),
                                               ctypes.byref(s)), "stream_compress_begin")
        parts = [head[: hl.value].tobytes()]
        q = make_params(params.chunk_size, params.unit_bits, params.batch_chunks, 0,
                        params.table_bits, tail_pad=not (params.flags & FLAG_NO_TAIL_PAD),
                        frozen=bool(params.flags & FLAG_FROZEN))
        seg = params.batch_chunks * params.chunk_size * max(1, segment_batches)
        try:
            pos = P
            while pos < N:
                n = min(seg, N - pos)
                out = np.empty(output_cap(n, q), dtype=np.uint8)
                olen = ctypes.c_uint64()
                check(lib().cucm_stream_compress_next(s, addr + pos, n, out.ctypes.data, out.size,
                                                      ctypes.byref(olen)), "stream_compress_next")
                parts.append(out[: olen.value].tobytes())
                pos += n
        finally:
            lib().cucm_stream_end(s)
        return b"".join(parts)

    def decompress_stream(self, blob, spec_bits=0, records_per_call=1):
        """Decompresses through the streaming C ABI, `records_per_call` batch
        records per call."""
        blob = bytes(blob)
        hl = ctypes.c_uint64()
        check(lib().cucm_container_head_len(blob, len(blob), ctypes.byref(hl)), "head_len")
        N, hp = peek_header(blob)
        pre = np.empty(max(hp.prelearn_size, 1), dtype=np.uint8)
        n0 = ctypes.c_uint64()
        p = Params()
        s = ctypes.c_void_p()
        check(lib().cucm_stream_decompress_begin(self.handle, blob, hl.value, spec_bits,
                                                 ctypes.byref(n0), ctypes.byref(p), pre.ctypes.data,
                                                 hp.prelearn_size, ctypes.byref(s)),
              "stream_decompress_begin")
        parts = [pre[: hp.prelearn_size].tobytes()]
        try:
            pos = hl.value
            per = p.chunk_size * p.batch_chunks
            while pos < len(blob):
                start, k = pos, 0
                while pos < len(blob) and k < records_per_call:
                    pos += 4 + int.from_bytes(blob[pos:pos + 4], "little")
                    k += 1
                recs = blob[start:pos]
                out = np.empty(max(per * k, 1), dtype=np.uint8)
                olen = ctypes.c_uint64()
                check(lib().cucm_stream_decompress_next(s, recs, len(recs), out.ctypes.data,
                                                        out.size, ctypes.byref(olen)),
                      "stream_decompress_next")
                parts.append(out[: olen.value].tobytes())
        finally:
            lib().cucm_stream_end(s)
        return b"".join(parts)

    # ---- serving loop (continuous batching) ------------------------------
    def serve_start(self, lanes=0, max_inflight=0, static_lanes=False):
        """Starts the request loop: the serving kernel takes requests as job
        slots (`max_inflight`, 0 = 256) free up, with `lanes` static-batch
        lanes for the requests it does not take; `static_lanes` runs every
        request on the lanes instead (0 = 3 lanes)."""
        check(lib().cucm_serve_start_ex(self.handle, lanes, max_inflight, 1 if static_lanes else 0),
              "serve_start")
        self._inflight = {}

    def serve_stop(self):
        """Waits for every submitted request, then stops the loop."""
        check(lib().cucm_serve_stop(self.handle), "serve_stop")

    def serve_launches(self) -> int:
        v = ctypes.c_uint64()
        check(lib().cucm_serve_launches(self.handle, ctypes.byref(v)), "serve_launches")
        return v.value

    def _submit(self, op, data, params, cap):
        addr, n, keep = _buf(data)
        out = np.empty(max(cap, 1), dtype=np.uint8)
        job = Job(op=op, inp=addr, in_len=n, params=params, out=out.ctypes.data, out_cap=cap)
        jid = ctypes.c_uint64()
        check(lib().cucm_submit(self.handle, ctypes.byref(job), ctypes.byref(jid)), "submit")
        self._inflight[jid.value] = (keep, out)
        return jid.value

    def submit_compress(self, data, params) -> int:
        return self._submit(0, data, params, output_cap(len(data), params))

    def submit_decompress(self, blob, spec_bits=0) -> int:
        size = int.from_bytes(bytes(blob[:4]), "big") if len(blob) >= 4 else 0
        return self._submit(1, blob, make_params(spec_bits=spec_bits), size)

    def poll(self, job_id, wait=True):
        """Returns (output bytes, result dict) once the job is done, or
        None while it runs (wait=False)."""
        r = Result()
        rc = lib().cucm_poll(self.handle, job_id, 1 if wait else 0, ctypes.byref(r))
        if rc == -8:
            return None
        check(rc, "poll")
        keep, out = self._inflight.pop(job_id)
        info = {"status": r.status, "batch_jobs": r.batch_jobs, "queue_ms": r.queue_ms,
                "latency_ms": r.latency_ms, "out_len": r.out_len}
        check(r.status, "job %d" % job_id)
        return out[: r.out_len].tobytes(), info


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context()
    return _default


def compress(data, chunk_size=MiB, unit_bits=128, batch_chunks=100, prelearn_size=0, table_bits=25,
             tail_pad=True, frozen=False) -> bytes:
    p = make_params(chunk_size, unit_bits, batch_chunks, prelearn_size, table_bits, 4, tail_pad,
                    frozen)
    return default_context().compress(data, p)


def decompress(blob, spec_bits=0) -> bytes:
    return default_context().decompress(blob, spec_bits)


def compress_file(infile, outfile, chunksize=MiB, unitlength=128, prelearn_size=0, batch_num=100,
                  table_bits=25):
    """Mirrors compress_file (cucm_compress_v2.cu:1151); returns the container size."""
    with open(infile, "rb") as f:
        data = f.read()
    blob = compress(data, chunksize, unitlength, batch_num, prelearn_size, table_bits)
    with open(outfile, "wb") as f:
        f.write(blob)
    return len(blob)


def decompress_file(infile, outfile, maxassumed_length=None):
    """Mirrors decompress (cucm_decompress_v2.cu:1185).  The reference decides
    maxassumed_length+1 bits per speculative round (3 + 1 by default); here
    that width is rounded down to 1, 2, 4 or 8 bits, and None (the default)
    lets the library choose it with the launch shape.  The width never
    changes the output."""
    if maxassumed_length is None:
        spec = 0
    else:
        width = maxassumed_length + 1
        spec = max(b for b in (1, 2, 4, 8) if b <= max(width, 1))
    with open(infile, "rb") as f:
        blob = f.read()
    data = decompress(blob, spec)
    with open(outfile, "wb") as f:
        f.write(data)
    return len(data)
