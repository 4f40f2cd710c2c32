/*
 * cucm.h -- C ABI of the MI355X-native CuCM chunked context-mixing codec.
 *
 * Drop-in boundary for CuCM's compress/decompress path.  The reference has
 * no library interface, only host C++ functions inside two CUDA programs;
 * each entry point below names the reference function whose behaviour it
 * replaces.  Containers produced with table_bits == 25 are byte-identical to
 * the reference's (cucm_compress_v2.cu:1151-1250, format comment
 * :1420-1443) and decode with the reference decoder.
 *
 * Conventions
 *   - The caller owns every buffer.  No entry point prints or exit()s.
 *   - Return 0 (CUCM_OK) or a negative CUCM_E_* code.
 *   - A cucm_ctx is bound to one device and is safe to use from one host
 *     thread at a time; distinct contexts may be used concurrently.
 *   - The one-shot functions (cucm_compress, cucm_decompress) use a
 *     process-wide context on device 0, serialised by a mutex.
 */
#ifndef CUCM_H
#define CUCM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CUCM_OK 0
#define CUCM_E_PARAM (-1)       /* invalid parameter */
#define CUCM_E_CORRUPT (-2)     /* malformed container */
#define CUCM_E_SMALL (-3)       /* output buffer too small */
#define CUCM_E_NOMEM (-4)       /* host or device allocation failed */
#define CUCM_E_BUDGET (-5)      /* job does not fit the context's HBM budget */
#define CUCM_E_DEVICE (-6)      /* HIP runtime error */
#define CUCM_E_UNSUPPORTED (-7) /* valid for the format, not for this build */
#define CUCM_E_NOTREADY (-8)    /* cucm_poll: job still running */

/* Encoder option: code the file's last chunk only up to its real end instead
 * of padding it to the nominal chunk size with stale bits as the reference
 * does (cucm_compress_v2.cu:850-851).  The decoder ignores padding, so both
 * forms decode identically; only the padded form is byte-identical to the
 * reference encoder. */
#define CUCM_FLAG_NO_TAIL_PAD 0x1u
/* Encoder option, the paper's "frozen model" (the reference ships no code
 * for it): every bit of every chunk is predicted from the initial model --
 * zero, or trained on the first prelearn_size bytes -- which is never
 * updated, and each chunk codes exactly its own bits (no tail padding).
 * The per-bit model and coder are the normal path's.  The container marks
 * it with bit 1 (0x2) of the header's flag byte (offset 20), which the
 * decoder reads back; the reference's reader does not know that bit.  With
 * no updates the model is read-only: one copy serves every chunk. */
#define CUCM_FLAG_FROZEN 0x2u

typedef struct cucm_params {
  uint32_t chunk_size;    /* S, bytes per chunk (reference -s, default 1 MiB) */
  uint16_t unit_bits;     /* U, bits per frozen-model unit, a multiple of 32 up
                             to 65,504 (reference -u, default 128); units above
                             256 bits run on a slower single-wave codec that
                             keeps the unit's records in device memory */
  uint16_t batch_chunks;  /* chunks per batch record (reference -b, default 100) */
  uint32_t prelearn_size; /* P, leading bytes that train the initial model (-p) */
  uint8_t table_bits;     /* log2 slots per order table, 8..25; 25 = reference */
  uint8_t spec_bits;      /* decoder bits per speculative round: 1, 2, 4 or 8;
                             0 (default) = automatic: 4-bit rounds as teams while
                             every chunk of a launch fits resident as a team,
                             else one wave per chunk with 2-bit rounds; never
                             changes the stream */
  uint16_t reserved;
  uint32_t flags;         /* CUCM_FLAG_*; other bits are rejected (CUCM_E_PARAM) */
} cucm_params;

/* Reference defaults: S = 1 MiB, U = 128, batch 100, P = 0, table_bits 25. */
void cucm_default_params(cucm_params *p);
const char *cucm_strerror(int code);

/* Largest container cucm_compress can produce for in_len bytes (strict:
 * the carry-less coder can emit up to 4 bytes per coded bit). */
uint64_t cucm_compress_bound(uint64_t in_len, const cucm_params *p);

/* Reads the 21-byte header (cucm_decompress_v2.cu:1189-1203,
 * main's 4-byte size read :1297-1301); out->flags gets CUCM_FLAG_FROZEN
 * when the container was coded with the frozen model. */
int cucm_peek_header(const uint8_t *in, uint64_t in_len, uint64_t *orig_len, cucm_params *out);

/* ---- one-shot, host buffers ------------------------------------------ */
/* Replaces compress_file (cucm_compress_v2.cu:1151-1154) on a byte buffer. */
int cucm_compress(const uint8_t *in, uint64_t in_len, const cucm_params *p, uint8_t *out,
                  uint64_t out_cap, uint64_t *out_len);
/* Replaces decompress (cucm_decompress_v2.cu:1185) on a byte buffer. */
int cucm_decompress(const uint8_t *in, uint64_t in_len, uint8_t *out, uint64_t out_cap,
                    uint64_t *out_len);

/* ---- contexts ---------------------------------------------------------- */
typedef struct cucm_ctx cucm_ctx;
/* hbm_budget_bytes == 0 means 90% of the device's free memory. */
int cucm_ctx_create(cucm_ctx **ctx, int device, uint64_t hbm_budget_bytes);
void cucm_ctx_destroy(cucm_ctx *ctx);
int cucm_ctx_compress(cucm_ctx *ctx, const uint8_t *in, uint64_t in_len, const cucm_params *p,
                      uint8_t *out, uint64_t out_cap, uint64_t *out_len);
int cucm_ctx_decompress(cucm_ctx *ctx, const uint8_t *in, uint64_t in_len, uint8_t *out,
                        uint64_t out_cap, uint64_t *out_len);

/* ---- batch records (no file header) ---------------------------------- */
/* Replaces compress_batch (cucm_compress_v2.cu:990-992): encodes n bytes as
 * ceil(n/S) chunks into one blob = 4C big-endian end offsets + payloads,
 * starting from a fresh model (P = 0). */
int cucm_compress_batch(cucm_ctx *ctx, const uint8_t *in, uint64_t n, uint32_t chunk_size,
                        uint32_t unit_bits, uint32_t table_bits, uint8_t *blob, uint64_t cap,
                        uint64_t *blob_len);
/* Replaces decompressbatch (cucm_decompress_v2.cu:898-903); unlike the
 * reference it does not modify any caller array. */
int cucm_decompress_batch(cucm_ctx *ctx, const uint8_t *blob, uint64_t blob_len, uint8_t *out,
                          uint64_t out_len, uint32_t chunk_size, uint32_t unit_bits,
                          uint32_t table_bits, uint32_t spec_bits);

/* ---- many files in one persistent launch ------------------------------ */
/* Host buffers.  out_len[i] receives each result; the call fails as a whole
 * on parameter errors, per-file container errors are reported in status[i]
 * (status may be NULL). */
int cucm_compress_many(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[],
                       const uint64_t in_len[], const cucm_params params[], uint8_t *const out[],
                       const uint64_t out_cap[], uint64_t out_len[], int32_t status[]);
int cucm_decompress_many(cucm_ctx *ctx, uint32_t n, const uint8_t *const in[],
                         const uint64_t in_len[], uint8_t *const out[], const uint64_t out_cap[],
                         uint64_t out_len[], int32_t status[], uint32_t spec_bits);
/* Device-resident buffers (device pointers on the context's device); data
 * never crosses PCIe except headers and sizes. */
int cucm_compress_many_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                              const uint64_t in_len[], const cucm_params params[],
                              uint8_t *const d_out[], const uint64_t out_cap[],
                              uint64_t out_len[], int32_t status[]);
int cucm_decompress_many_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                                const uint64_t in_len[], uint8_t *const d_out[],
                                const uint64_t out_cap[], uint64_t out_len[], int32_t status[],
                                uint32_t spec_bits);
/* "Plain" path: the same work issued one file per launch (the reference's
 * structure), for comparison with the persistent batched launch. */
int cucm_compress_plain_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                               const uint64_t in_len[], const cucm_params params[],
                               uint8_t *const d_out[], const uint64_t out_cap[],
                               uint64_t out_len[]);
int cucm_decompress_plain_device(cucm_ctx *ctx, uint32_t n, const uint8_t *const d_in[],
                                 const uint64_t in_len[], uint8_t *const d_out[],
                                 const uint64_t out_cap[], uint64_t out_len[],
                                 uint32_t spec_bits);

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (serving loops with a fixed footprint, code that must not call the
 * allocator).  Codes one device-resident input into a device-resident
 * container exactly as cucm_compress_many_device does, taking all its
 * scratch (tasks, results, payload staging, offsets, model slots) from one
 * 64 MiB device array that each GPU holds from library load until the
 * process ends; it calls no hipMalloc/hipFree and no allocating routine.
 * Limits: prelearn_size must be 0 (CUCM_E_UNSUPPORTED otherwise); the chunk
 * staging (1.5 S + 256 B per chunk) plus one model slot (7 x 2^table_bits B,
 * none for the frozen model) must fit the array (CUCM_E_BUDGET otherwise --
 * e.g. up to ~24 MiB of input at table_bits 18, ~1 MiB at 25); a chunk whose
 * payload exceeds its staging (adversarial input only) gives
 * CUCM_E_UNSUPPORTED.  On CUCM_E_SMALL *out_len holds the container size.
 * Calls on one GPU share the array and are serialised in this process. */
int cucm_compress_device_noalloc(cucm_ctx *ctx, const uint8_t *d_in, uint64_t in_len,
                                 const cucm_params *p, uint8_t *d_out, uint64_t out_cap,
                                 uint64_t *out_len);

/* ---- streaming container I/O ----------------------------------------- */
/* The reference's compress_file / decompress (cucm_compress_v2.cu:1151-1250,
 * cucm_decompress_v2.cu:1184-1248) hold the whole file in host memory; a
 * stream produces or consumes the same container piece by piece, so the
 * device footprint is one segment of batch records however large the input
 * (up to the container's 32-bit size fields: CUCM_E_PARAM beyond them). */
typedef struct cucm_stream cucm_stream;
/* Bytes of header + prelearn segment at the front of a container, from its
 * first 21 bytes. */
int cucm_container_head_len(const uint8_t *in, uint64_t in_len, uint64_t *head_len);
/* Starts compressing an input of total_len bytes; prelearn = its first
 * p->prelearn_size bytes.  Writes the header and prelearn segment to head
 * (CUCM_E_SMALL, with *head_len set, if head_cap is too small). */
int cucm_stream_compress_begin(cucm_ctx *ctx, uint64_t total_len, const cucm_params *p,
                               const uint8_t *prelearn, uint8_t *head, uint64_t head_cap,
                               uint64_t *head_len, cucm_stream **s);
/* Codes the next len input bytes after the prelearn bytes -- a multiple of
 * batch_chunks * chunk_size, or all that remains -- and writes their batch
 * records (LE u32 length + blob each) to out. */
int cucm_stream_compress_next(cucm_stream *s, const uint8_t *in, uint64_t len, uint8_t *out,
                              uint64_t out_cap, uint64_t *out_len);
/* Starts decompressing: head holds the container's first
 * cucm_container_head_len bytes.  Returns the original size and parameters
 * and writes the prelearn bytes to prelearn_out. */
int cucm_stream_decompress_begin(cucm_ctx *ctx, const uint8_t *head, uint64_t head_len,
                                 uint32_t spec_bits, uint64_t *orig_len, cucm_params *p,
                                 uint8_t *prelearn_out, uint64_t prelearn_cap, cucm_stream **s);
/* Decodes whole batch records (the next ones in container order) to out. */
int cucm_stream_decompress_next(cucm_stream *s, const uint8_t *records, uint64_t len,
                                uint8_t *out, uint64_t out_cap, uint64_t *out_len);
void cucm_stream_end(cucm_stream *s);

/* ---- device memory helpers (for callers without their own allocator) -- */
int cucm_ctx_malloc(cucm_ctx *ctx, void **dptr, uint64_t bytes);
int cucm_ctx_free(cucm_ctx *ctx, void *dptr);
int cucm_ctx_memcpy_h2d(cucm_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int cucm_ctx_memcpy_d2h(cucm_ctx *ctx, void *dst, const void *src, uint64_t bytes);
int cucm_ctx_synchronize(cucm_ctx *ctx);

/* ---- statistics of the last call on a context ------------------------- */
typedef struct cucm_stats {
  double wall_ms;          /* host wall time of the call */
  double codec_ms;         /* encode/decode kernel, HIP events on the ctx stream */
  double assemble_ms;      /* container layout + payload copy kernels */
  double h2d_ms, d2h_ms;   /* transfers (host-buffer entry points only) */
  uint64_t tasks;          /* chunks processed */
  uint64_t waves;          /* resident worker waves of the persistent launch */
  uint64_t launches;       /* codec kernel launches */
  uint64_t in_bytes, out_bytes;
  uint64_t model_bytes;    /* HBM held by model slots */
  uint64_t teams;          /* workgroups that ran as teams of 4 waves (0: one wave per chunk) */
  uint64_t spec_bits;      /* decoder bits per speculative round of the last decode launch */
} cucm_stats;
int cucm_ctx_last_stats(cucm_ctx *ctx, cucm_stats *st);

/* Per-phase shader-clock totals, summed over all waves, of the last codec
 * launch on ctx (phase profiling must be enabled first; it adds a clock read
 * per phase).  The counterpart of the reference's host phase timers
 * g_time_mem/predict/mix/encode (cucm_compress_v2.cu:983-988).  Phases: 0 model
 * init, 1 unit staging, 2 hash + gather, 3 slot tables + mix, 4 arithmetic
 * coder, 5 mixer deltas + collision scan, 6 state resolution, 7 statep
 * recurrences, 8 state write-back, 9 decoder path predictions, 10 decoder mix +
 * decode, 11 decoder records. */
#define CUCM_PHASES 12
int cucm_ctx_set_phase_profile(cucm_ctx *ctx, int enable);
int cucm_ctx_phase_cycles(cucm_ctx *ctx, uint64_t cycles[CUCM_PHASES]);

/* Upper bound on concurrently resident worker waves (0 = automatic). */
int cucm_ctx_set_max_waves(cucm_ctx *ctx, uint32_t waves);
/* Waves per workgroup, 1..8 (0 = automatic): the waves of a workgroup share
 * one LDS copy of the lookup tables; fewer per workgroup spreads a small
 * launch over more CUs.  Never changes the output. */
int cucm_ctx_set_waves_per_block(cucm_ctx *ctx, uint32_t waves);
/* Team execution: one chunk per workgroup of 4 waves.  Encoder: two waves
 * hash, gather and predict a unit while one walks the previous unit's
 * StateMap recurrences and one codes its bits.  Decoder: the leader runs the
 * speculative round chain; the others record rounds, resolve states, walk
 * the StateMap recurrences and hash the next round's paths.  mode 0
 * (default): teams when every chunk of a launch can be resident as a team at
 * once (decoder: spec_bits 1, 2 or 4; with spec_bits 0 only then); 1: never;
 * 2: whenever allowed.  Never changes the output.  Automatic mode also uses
 * teams when the chunks fill at least two full passes of resident teams
 * (encoder, and decoder with a fixed spec_bits). */
int cucm_ctx_set_team(cucm_ctx *ctx, uint32_t mode);

/* ---- serving: continuous batching ------------------------------------ */
/* No counterpart in the reference (one file per process run there).
 * Requests of mixed kinds and sizes are queued as they arrive and admitted,
 * oldest first, into one running kernel as soon as a job slot and memory
 * are free: their chunk tasks join a task ring that the kernel's resident
 * waves take as they free up, compress and decompress mixed (see
 * cucm_serve_start).  A request's result never depends on what it ran
 * beside. */
#define CUCM_OP_COMPRESS 0
#define CUCM_OP_DECOMPRESS 1
typedef struct cucm_job {
  int32_t op;              /* CUCM_OP_* */
  int32_t reserved;
  const uint8_t *in;       /* host buffer, valid until cucm_poll reports the job */
  uint64_t in_len;
  cucm_params params;      /* compress: all fields; decompress: spec_bits only */
  uint8_t *out;            /* host buffer */
  uint64_t out_cap;
} cucm_job;
typedef struct cucm_result {
  int32_t status;          /* CUCM_OK or CUCM_E_* */
  int32_t batch_jobs;      /* requests on the GPU with this one when it was admitted
                              (serving kernel), or in its batch (lanes) */
  uint64_t out_len;
  double queue_ms;         /* submit -> admitted (tasks published, or taken by a lane) */
  double latency_ms;       /* submit -> result in the caller's buffer */
} cucm_result;
/* Starts the serving loop on ctx.  One persistent launch of the serving
 * kernel takes the chunk tasks of every admitted request, compress and
 * decompress mixed, as its waves free up; a request is admitted as soon as
 * one of max_inflight (0 = 256, rounded up to a power of 2, at most 4096) job
 * slots and its memory are free.  Requests the kernel does not take (frozen
 * model, U above 128, tables above 2^20 slots, or larger than the request
 * arena) run as static batches on `lanes` (0 = 1, at most 16) concurrent
 * lanes.  No counterpart in the reference, which has no serving loop. */
int cucm_serve_start(cucm_ctx *ctx, uint32_t lanes, uint32_t max_inflight);
/* cucm_serve_start with flags: CUCM_SERVE_STATIC_LANES runs every request as
 * static batches on the lanes (lanes 0 = 3), without the serving kernel --
 * the loop's earlier design, kept for comparison. */
#define CUCM_SERVE_STATIC_LANES 1u
int cucm_serve_start_ex(cucm_ctx *ctx, uint32_t lanes, uint32_t max_inflight, uint32_t flags);
/* Waits until every submitted request has completed, then stops the lanes.
 * Must not run concurrently with cucm_submit/cucm_poll on the same ctx. */
int cucm_serve_stop(cucm_ctx *ctx);
int cucm_submit(cucm_ctx *ctx, const cucm_job *job, uint64_t *job_id);
/* CUCM_E_NOTREADY while running (wait == 0), else blocks until done.  A
 * completed job is reported once; its id is then forgotten. */
int cucm_poll(cucm_ctx *ctx, uint64_t job_id, int wait, cucm_result *res);
/* Launches issued by the serving loop so far. */
int cucm_serve_launches(cucm_ctx *ctx, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
