/*
 * faf_amd.h -- additive entry points of the MI355X build.
 *
 * None of these exist in the reference (FastAndFourier 1.2.0 is
 * single-signal and single-threaded); they are new symbols, so programs
 * written against the reference are unaffected.  All buffers may be device
 * pointers (used in place) or host pointers (staged through device memory).
 */
#ifndef FAF_AMD_H
#define FAF_AMD_H
#include "fastandfourier.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- stream and synchronisation control (thread-local) ----
 * faf_execute* launch on the calling thread's stream (a hipStream_t,
 * NULL = the default stream).  With sync on (the default) every execute
 * returns after its results are complete, as in the reference. */
int faf_amd_set_stream(void *hip_stream);
void *faf_amd_get_stream(void);
int faf_amd_set_sync(int on);
int faf_amd_synchronize(void);
int faf_amd_device_ready(void);              /* 0 if a gfx950 device is usable */
/* the device's streaming copy (16-byte loads and stores, grid-stride over
 * `grid` workgroups; 0 = 8 per CU): what a read-once / write-once kernel
 * reaches on this board, the practical ceiling the transforms are measured
 * against beside the 8 TB/s peak.  Device buffers, 16-byte aligned. */
int faf_amd_copy_ceiling(void *dst, const void *src, size_t bytes, int grid);
/* the same copy; streaming != 0 stores with non-temporal hints, as the
 * transforms do for rows they write once */
int faf_amd_copy_ceiling_ex(void *dst, const void *src, size_t bytes, int grid,
                            int streaming);

/* ---- graph capture of executes (thread-local) ----
 * Between faf_graph_begin and faf_graph_end, faf_execute* and
 * faf_ragged_execute on this thread are recorded into a HIP graph instead of
 * running; faf_graph_launch then replays the whole sequence with one launch
 * on the thread's stream (synchronising when sync is on).  The capture uses
 * the thread's stream, or a private stream when that is the default stream.
 * While a capture is open:
 *   - every buffer must be device memory (host staging cannot be captured);
 *   - each plan must already have run once at its largest batch, so that its
 *     workspace is allocated (an execute that would grow it fails);
 *   - creating plans, autotuning and faf_amd_synchronize fail.
 * An execute that fails inside the capture makes faf_graph_end return NULL.
 * Replays read and write the buffers that were captured, so they see the
 * buffers' contents at replay time.  Plans must outlive their graphs. */
typedef struct faf_graph faf_graph;
int faf_graph_begin(void);
faf_graph *faf_graph_end(void);
int faf_graph_launch(const faf_graph *g);
int faf_graph_nodes(const faf_graph *g);     /* kernel and copy nodes in the graph */
void faf_graph_destroy(faf_graph *g);

/* ---- batched execute ----
 * Runs `count` independent signals of the plan in one launch.  `in` and
 * `out` describe signal 0 exactly as for faf_execute; signal i starts
 * in_dist / out_dist elements (complex elements for INTERLEAVED) later in
 * every plane.  dist 0 means packed (the buffer length). */
int faf_execute_batch(const faf_transform *t, faf_buffer *out, const faf_buffer *in,
                      size_t count, size_t in_dist, size_t out_dist);

/* ---- batched execute that allocates nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the same transform as faf_execute_batch, with the same
 * kernels and the same results bit for bit, but the workspace that plans
 * longer than one workgroup holds (four-step intermediates of FFT/RFFT
 * lengths above 16384 points, long CWT segments) comes from a static
 * device array of FAF_NOALLOC_SCRATCH_BYTES per GPU instead of a buffer
 * the plan allocates.  Four-step batches are cut into chunks that fit it; a
 * run whose single chunk needs more fails with an error naming both sizes.
 * Every buffer (and a pipeline's bound spectrum) must be device memory.
 * The call never allocates: anything that would (host staging, a plan
 * record not yet resident) fails with a message instead, so run a plan once
 * through faf_execute_batch (or create it) before calling this.
 * All calls on one GPU share that GPU's scratch, which keeps the last
 * call's contents: calls that overlap on one GPU (other threads, other
 * streams) race and must be serialised by the caller. */
#define FAF_NOALLOC_SCRATCH_BYTES ((size_t)32 << 20)
int faf_execute_batch_noalloc(const faf_transform *t, faf_buffer *out, const faf_buffer *in,
                              size_t count, size_t in_dist, size_t out_dist);

/* ---- ragged C2C batch in one launch ----
 * Signals of different lengths are concatenated in split planes at
 * offsets = exclusive prefix sum of lengths.  Every length must be legal
 * for faf_create_fft (5-smooth, codelet, or Rader prime; others need
 * FAF_FLAG_BLUESTEIN in flags).  Forward, NumPy norm (unscaled). */
typedef struct faf_ragged faf_ragged;
faf_ragged *faf_ragged_create(const size_t *lengths, size_t count, uint32_t flags);
size_t faf_ragged_total(const faf_ragged *r);
int faf_ragged_execute(faf_ragged *r, float *out_re, float *out_im, const float *in_re,
                       const float *in_im);
void faf_ragged_destroy(faf_ragged *r);
/* Autotunes the launch for this batch's lengths: for each length it times,
 * on the device and on a batch of that length alone, every way the length
 * can run (compile-time passes, or runtime radix passes in each candidate
 * radix order) with several signals-per-unit counts, median of 5 launches
 * each; then it times the whole batch (median of 9) with the per-length
 * winners, the best compile-time choices (the launch without runtime paths
 * uses a leaner kernel) and the defaults, in both launch shapes and both
 * workgroup sizes, and keeps the fastest.  The winners are recorded for the
 * process, so later faf_ragged_create calls use them for those lengths;
 * faf_ragged_create with FAF_FLAG_MEASURE in flags runs this at create.
 * Rebuilds r in place.  Returns the number of lengths timed (0 when every
 * length already had a recorded choice), or -1. */
int faf_ragged_autotune(faf_ragged *r);
/* How length `len` runs in r: kind (0 runtime radix passes, 1 Rader, 2
 * Bluestein, 3 compile-time passes, 4 Rader with compile-time inner
 * passes), signals per unit, and the radix order of its (inner) FFT. */
int faf_ragged_choice(const faf_ragged *r, size_t len, int *kind, int *spb, int *order,
                      int *norder);
/* Records for the process how length `len` runs in ragged plans created
 * afterwards, in faf_ragged_choice's terms: kind (as above), signals per
 * unit (0: as many as fit) and the radix order of its (inner) FFT (norder
 * 0: the default order of that kind).  A compile-time kind accepts only
 * the orders its kernel carries; kind -1 forgets the length's record, so
 * it runs by the planner's defaults (and the autotuner times it again).
 * Returns 0, or -1 with a message when the length cannot run that way. */
int faf_ragged_set_choice(size_t len, int kind, int spb, const int *order, int norder);
/* Launch shape of r: 0 = one workgroup per unit (the units are dispatched
 * largest first; the default), 1 = persistent workgroups, one per resident
 * slot, that take units from a device queue until it is empty.
 * faf_ragged_autotune times both on the batch and keeps the faster. */
int faf_ragged_persistent(const faf_ragged *r);
int faf_ragged_set_persistent(faf_ragged *r, int on);
/* Threads per workgroup of r's launch: 256 when every length of the batch
 * runs on that many (one 8192-point signal per workgroup), else 512.
 * faf_ragged_set_block takes 0 (that automatic size), 256 or 512 and
 * rebuilds r's units; when a length needs more threads it fails and r stays
 * as it was.  faf_ragged_autotune times both sizes on the batch and keeps
 * the faster. */
int faf_ragged_block(const faf_ragged *r);
int faf_ragged_set_block(faf_ragged *r, int threads);

/* ---- framed real STFT ----
 * faf_create_stft with cfg.layout == FAF_LAYOUT_REAL (an error in the
 * reference) builds a framed STFT: cfg.win_length (default n) samples per
 * frame, cfg.hop_length (default win/4), symmetric Hann window
 * (reference src/faf_twiddles.c:159-163), one rfft per frame, no centre
 * padding; a record shorter than the window is one zero-padded frame.
 * Output: re/im planes [records][frames][win/2+1]. */
size_t faf_stft_frames(const faf_transform *t, size_t len);
int faf_execute_stft(const faf_transform *t, float *out_re, float *out_im, const float *x,
                     size_t len, size_t records, size_t rec_dist);

/* ---- CWT with analytic complex64 rows ----
 * out: interleaved complex64 [count][n_rows][n], with Re(out row) equal to
 * the reference's real row W = irfft(rfft(x) * psi). */
int faf_execute_cwt_complex(const faf_transform *t, float *out, const float *x, size_t count,
                            size_t in_dist);

/* ---- fused spectral pipeline without the DSL ----
 * rfft(n) -> ops -> irfft(n) in one kernel.  op codes: 1 conj, 2 mul by the
 * spectrum bound with chirp_bind(t, "H", ...), 3 bandpass keeping bins
 * [a1[i], a2[i]].  Same semantics as reference src/chirp.c:421-471. */
faf_transform *faf_create_pipeline(const faf_config *cfg, int nops, const int *ops,
                                   const uint32_t *a1, const uint32_t *a2);

/* ---- autotuner ----
 * Times the candidate launch shapes of the plan on `count` synthetic
 * signals -- the specialised kernel where the length has one, and the
 * runtime radix kernels in each candidate radix order with 1..64 signals
 * per workgroup, each the median of 5 runs -- applies the fastest to the
 * plan and records it for the process under the plan's key (kind, length,
 * direction, interleaved layout): plans of that key created later start
 * from it.  FAF_FLAG_MEASURE in cfg.flags runs this at create for FFT and
 * RFFT plans whose key has no recorded shape.  Returns the chosen
 * signals-per-workgroup, or -1. */
int faf_amd_autotune(faf_transform *t, size_t count);
/* The plan's launch shape: specialised kernel or not, the radix order of
 * its (inner) FFT and signals per workgroup. */
int faf_amd_plan_shape(const faf_transform *t, int *specialised, int *order, int *norder,
                       int *spb);

/* ---- kernel selection (thread-local, read at create time) ----
 * Power-of-two lengths 256..16384 (C2C) and 512..32768 (R2C/C2R, fused
 * pipeline, STFT window) use compile-time specialised kernels by default;
 * 0 makes later creates use the runtime-generic radix kernels instead. */
int faf_amd_set_specialised(int on);

/* ---- dual-frame filter row of a CWT/ICWT bank (packed Hermitian, n/2+1
 * bins; row 0 is the lowpass when present).  Valid until destroy. ---- */
const float *faf_amd_cwt_dual_f32(const faf_transform *t, size_t row);

/* ---- device-side timing of the last execute on this thread (ms) ---- */
float faf_amd_last_kernel_ms(void);
int faf_amd_set_timing(int on);

/* ---- serving loop with continuous batching ----
 * Requests carry a device-resident input and get their result in device
 * memory that the server allocates from an arena of cfg.hbm_budget bytes.
 * A worker thread admits requests in arrival order while their result (and
 * workspace) fits the budget, gathers admitted requests into one launch of
 * the persistent serving kernel whenever a stream slot is free (at once when
 * the device is idle, else after window_us or when max_batch is reached) and
 * completes them as the slot's events fire, so requests that arrive while
 * others run are admitted and overlap with them on the device.
 *
 * Request kinds and results (interleaved complex64, unnormalised forward):
 *   RFFT  n even real samples        -> n/2+1 bins
 *   C2C   n interleaved complex      -> n bins
 *   STFT  n real samples             -> frames x (win/2+1), frames =
 *         1 + (n - win)/hop (one zero-padded frame when n < win), symmetric
 *         Hann window, no centre padding (as faf_execute_stft)
 *   CWT   n real samples (n even)    -> rows x n analytic rows (as
 *         faf_execute_cwt_complex) of a Morse bank with cwt_voices voices
 *         per octave from fmin = 16 fs/n to fmin * 2^cwt_octaves, plus the
 *         lowpass row. */
typedef enum {
    FAF_REQ_RFFT = 0, FAF_REQ_C2C = 1, FAF_REQ_STFT = 2, FAF_REQ_CWT = 3
} faf_req_kind;

#define FAF_REQ_DISCARD 0x1u   /* free the result as soon as it completes */

typedef struct faf_server_config {
    size_t hbm_budget;      /* bytes of results + workspace in flight (default 64 GiB) */
    double window_us;       /* batching window while the device is busy (default 60) */
    int streams;            /* stream slots, 1..8 (default 3) */
    int max_batch;          /* requests per launch (default 1024) */
    int plain;              /* 1: no batching, one launch + synchronize per request */
    size_t stft_win;        /* default 512 */
    size_t stft_hop;        /* default 128 */
    double cwt_fs;          /* default 16000 */
    unsigned cwt_voices;    /* default 8 */
    double cwt_octaves;     /* default 31.5 / 8 (32 scales) */
} faf_server_config;

typedef struct faf_server faf_server;

typedef struct faf_server_stats {
    uint64_t submitted, completed, failed;
    uint64_t launches;           /* batches dispatched (one stream slot each) */
    double mean_batch;           /* requests per batch */
    double p50_us, p99_us, max_us, mean_us;   /* arrival -> completion */
    double span_s;               /* first arrival -> last completion */
    double req_per_s;            /* completed / span */
    double gb_per_s;             /* algorithmic bytes (input + result) / span */
    size_t peak_hbm_bytes;       /* arena bytes held at the peak */
    uint64_t admission_waits;    /* requests that waited for budget */
    uint64_t long_requests;      /* requests run through their own plan launches */
    double host_us_per_batch;    /* worker time to build and launch a batch */
    double device_us_per_batch;  /* launch -> completion seen by the worker */
} faf_server_stats;

faf_server_config faf_server_config_init(void);
faf_server *faf_server_create(const faf_server_config *cfg);
void faf_server_destroy(faf_server *s);     /* waits for work in flight */
/* switch plain per-request service on or off (waits for work in flight) */
int faf_server_set_plain(faf_server *s, int on);
/* change the batching window (waits for work in flight) */
int faf_server_set_window(faf_server *s, double window_us);
/* build and upload the plan of (kind, n) now rather than at first use */
int faf_server_prepare(faf_server *s, faf_req_kind kind, size_t n);
/* the same for count (kinds[i], ns[i]) pairs, built on the server's builder
 * threads side by side; returns when every plan is ready, 0, or -1 with the
 * first failure's message (the other plans are still built) */
int faf_server_prepare_many(faf_server *s, const int *kinds, const size_t *ns, size_t count);
/* bytes of the result of one (kind, n) request, 0 if not servable */
size_t faf_server_result_bytes(faf_server *s, faf_req_kind kind, size_t n);
/* Returns a ticket >= 0, or -1 (faf_get_error) for a request that is wrong
 * on its face (unknown kind, length out of range, odd rfft, NULL input).
 * A (kind, n) seen for the first time has its plan built on one of the
 * server's builder threads: the call returns at once and the request waits
 * for the plan while requests with ready plans go ahead of it.  Should that
 * build fail, the ticket fails (faf_server_wait returns -1) and later
 * submissions of that (kind, n) return -1 with the build's error. */
int64_t faf_server_submit(faf_server *s, faf_req_kind kind, size_t n, const float *in,
                          uint32_t flags);
int faf_server_poll(faf_server *s, int64_t ticket);   /* 1 done, 0 pending, -1 failed */
int faf_server_wait(faf_server *s, int64_t ticket);   /* 0 done, -1 failed */
/* device pointer of a completed result (NULL if pending, failed or freed) */
const float *faf_server_result(faf_server *s, int64_t ticket, size_t *bytes);
int faf_server_release(faf_server *s, int64_t ticket);   /* returns its memory */
int faf_server_drain(faf_server *s);                      /* waits for every request */
int faf_server_stats_get(faf_server *s, faf_server_stats *st);
void faf_server_stats_reset(faf_server *s);   /* later stats cover later requests */
/* Replays a request stream in real time from the calling thread: request i
 * is submitted arrival_ns[i] after the call starts (arrival = that instant)
 * with `flags`; returns after all complete.  latency_ns, if given, receives
 * each request's arrival -> completion time. */
int faf_server_replay(faf_server *s, const int *kinds, const size_t *ns,
                      const float *const *ins, const uint64_t *arrival_ns, size_t count,
                      uint32_t flags, uint64_t *latency_ns);

#ifdef __cplusplus
}
#endif
#endif
