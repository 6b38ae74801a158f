/*
 * gpuspgemm.h -- C ABI of the MI355X (gfx950) SpGEMM library.
 *
 * Every entry point returns an spg_status and never throws.  Matrices are
 * 0-based CSR with int32 indices, as utils::CSR in the reference
 * (src/libutils/include/utils/csr.cuh:23-31); `val` is float or double per
 * spg_dtype.  Device pointers are plain HIP device addresses.
 *
 * Reference interfaces each entry point stands in for (paths relative to
 * the GPUSpGEMM source tree):
 *
 *   spg_multiply        proposal::proposal<T>(A, B) -> C
 *                       src/libproposal/include/proposal/proposal.cuh:85-124
 *   spg_multiply_async  proposal::proposal<T> on the caller's stream: the
 *                       reference runs on the per-thread default stream
 *                       (cmake/wrappers.cmake:61-63), so concurrent host
 *                       threads multiply on streams of their own
 *   spg_csr_free        utils::CSR<T,Device>::~CSR / release()
 *                       src/libutils/include/utils/csr.cuh:114-117, 241-252
 *   spg_csr_free_async  the same, ordered on a caller stream
 *   spg_set_log_level   the apps' --loglevel (spdlog level), which the
 *                       library's debug logs follow: bin tables, nnz(C), the
 *                       fail-bin size (binning.cuh:181-187, 232-233,
 *                       sym1.cuh:430)
 *   spg_get_nip         utils::get_nip(a, b) -> int64
 *                       src/libutils/include/utils/nip.cuh:52-78
 *   spg_malloc/spg_free utils::malloc<Device>/free<Device>
 *                       src/libutils/include/utils/runtime.cuh:11-20, 58-66
 *   spg_memcpy          utils::memcpy (host<->device, cudaMemcpyDefault)
 *                       src/libutils/include/utils/runtime.cuh:45-48; used by
 *                       CSR::to<Location>() csr.cuh:193-203
 *   spg_device_sync     utils::device_sync  runtime.cuh:98-101
 *   spg_status_str /    utils::handle_cuda_error message text
 *   spg_last_error      src/libutils/include/utils/errors.cuh:11-22
 *
 * Library-specific additions (no reference counterpart; needed by the
 * drop-in harness, the benchmark, the tuner and the two system forms that
 * stand beside the plain call):
 *   spg_init, spg_device_info, spg_set_tuning, spg_get_tuning,
 *   spg_get_variants, spg_set_profiling, spg_profile_count, spg_profile_get,
 *   spg_multiply_batched (many products in one pass of the kernels),
 *   spg_set_graphs, spg_graph_stats (graph replay of repeated calls),
 *   spg_server_* (a request loop with continuous batching).
 */
#ifndef GPUSPGEMM_H
#define GPUSPGEMM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum spg_status {
    SPG_OK = 0,
    SPG_ERR_INVALID_ARGUMENT = 1, /* null pointer, A.n != B.m, negative sizes */
    SPG_ERR_OUT_OF_MEMORY = 2,    /* device allocation failed */
    SPG_ERR_HIP = 3,              /* a HIP runtime call failed; see spg_last_error */
    SPG_ERR_OVERFLOW = 4,         /* nnz(C) or a row's product count >= 2^31 */
    SPG_ERR_UNSUPPORTED_DEVICE = 5, /* not a gfx950 device */
    SPG_ERR_INTERNAL = 6          /* a consistency check inside the library failed */
} spg_status;

typedef enum spg_dtype { SPG_F32 = 0, SPG_F64 = 1 } spg_dtype;

/* Same members and order as utils::CSR (csr.cuh:25-30). */
typedef struct spg_csr {
    int32_t nnz;
    int32_t m;
    int32_t n;
    int32_t *rpt; /* [m + 1] */
    int32_t *col; /* [nnz]   */
    void *val;    /* [nnz] float or double */
} spg_csr;

/* Initialise the library on the current HIP device (idempotent; called by
   every entry point on first use). */
spg_status spg_init(void);

/* C = A . B on the current device.  A and B are borrowed device matrices
   (B's rows sorted ascending, indices in range).  On success *C holds
   library-allocated device arrays (release with spg_csr_free) and the
   result is complete when the call returns, as the reference's
   synchronous proposal().  The work is ordered after what the calling
   thread queued on its default stream (hipStreamPerThread, which also
   follows the legacy null stream).  Calls from different host threads run
   concurrently, each on a lane (aux streams, scratch) of its own. */
spg_status spg_multiply(const spg_csr *A, const spg_csr *B, spg_csr *C, spg_dtype dtype);

/* C = A . B ordered on `stream` (a hipStream_t; null: the calling thread's
   default stream).  The call waits once, for the read-back of nnz(C) that
   sizes C (the method's symbolic phase), and returns once the fused pass is
   queued: C's arrays are valid at return and its contents are complete when
   `stream` reaches the point of the call.  Calls on one stream run one
   after another; calls on different streams run concurrently.  A
   consistency failure of the fused pass (only possible when B has
   duplicate columns in a row or indices out of range) is reported by the
   next call on the same stream. */
spg_status spg_multiply_async(const spg_csr *A, const spg_csr *B, spg_csr *C, spg_dtype dtype, void *stream);

/* No counterpart in the reference: C = A . B for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving
   loops, harnesses that forbid hipMalloc).  The caller passes C->rpt
   (A->m + 1 entries) and C->col / C->val (`capacity` entries each); the
   call fills them, sets C->nnz, C->m, C->n and returns once C is complete.
   Scratch comes from a fixed 256 MiB device array that the library holds
   on each device until the process ends; no device or pinned memory is
   allocated inside the call.  A product whose scratch does not fit in it
   fails with SPG_ERR_OUT_OF_MEMORY before any launch; nnz(C) > capacity
   fails with SPG_ERR_OUT_OF_MEMORY and C->nnz set to the nnz needed (C's
   arrays then hold no result).  Calls on one device share the array: they
   take turns on a lock, so calls that would overlap run one after
   another.  Work is ordered after the calling thread's default stream. */
spg_status spg_multiply_noalloc(const spg_csr *A, const spg_csr *B, spg_csr *C, int64_t capacity,
                                spg_dtype dtype);

/* Many independent products C_i = A_i . B_i in one pass of the kernels
   (rows of all instances are binned together).  Arrays have `count`
   entries; each C[i] is allocated like spg_multiply's. */
spg_status spg_multiply_batched(int32_t count, const spg_csr *A, const spg_csr *B,
                                spg_csr *C, spg_dtype dtype);

/* Release the device arrays of a library-allocated matrix and zero it.
   The library keeps released blocks and hands them out again for later
   outputs once the GPU work queued so far on the calling thread's default
   stream has completed; work on other streams that reads C must be ordered
   before that (as for hipFreeAsync). */
spg_status spg_csr_free(spg_csr *C);
/* The same, with the blocks reused once `stream` (a hipStream_t; null:
   the thread's default stream) reaches this point. */
spg_status spg_csr_free_async(spg_csr *C, void *stream);

/* Total number of intermediate products sum_i sum_{k in A(i,:)} nnz(B(k,:)). */
spg_status spg_get_nip(const spg_csr *A, const spg_csr *B, int64_t *nip);

/* Device memory helpers for hosts that do not link HIP themselves. */
spg_status spg_malloc(void **ptr, size_t bytes);
spg_status spg_free(void *ptr);
spg_status spg_memcpy(void *dst, const void *src, size_t bytes); /* any direction */
spg_status spg_device_sync(void);

const char *spg_status_str(spg_status status);
/* Message of the last failure on this thread ("" if none). */
const char *spg_last_error(void);

/* Library log level, spdlog's order: 0 trace (every launch), 1 debug (per
   multiply: rows, NIP, nnz(C), rows per bin of both phases, fail-bin size,
   host time of the phases), 2 info ... 6 off (the default).  Before the
   first call the environment variable SPG_LOG_LEVEL (a name or a number)
   sets it.  Lines go to stderr. */
spg_status spg_set_log_level(int32_t level);

/* Device facts used by the tuner and the benchmark. */
typedef struct spg_device_info {
    char arch[64];           /* e.g. "gfx950" */
    int32_t compute_units;
    int32_t lds_per_block;   /* bytes */
    int32_t clock_khz;
    int64_t global_mem_bytes;
} spg_device_info;
spg_status spg_device_info_get(spg_device_info *info);

/* Bin tables.  phase 0 = symbolic count (bins by NIP), phase 1 = the fused
   sorted-columns + numeric pass (bins by nnz and NIP).  Each bin names a
   compiled kernel variant and the largest NIP / nnz it accepts; rows go to
   the first bin that accepts them.  Text format, one bin per line:
   "<variant> <nip_cap> <nnz_cap>". */
spg_status spg_set_tuning(int32_t phase, const char *table);
spg_status spg_get_tuning(int32_t phase, char *buf, size_t buf_len);
/* Names of the compiled variants of a phase, one per line. */
spg_status spg_get_variants(int32_t phase, char *buf, size_t buf_len);

/* Per-kernel timing with HIP events on the library stream (off by
   default).  After a multiply, spg_profile_count() records are available:
   kernel name, rows handled, launch grid and duration in milliseconds. */
spg_status spg_set_profiling(int32_t enabled);
int32_t spg_profile_count(void);
spg_status spg_profile_get(int32_t index, char *name, size_t name_len, int64_t *rows,
                           int32_t *grid, double *ms);

/* Graph replay (off by default).  When on, each of a multiply's two
   phases (symbolic: NIP, binning, counts, scan, read-back; fused pass:
   every bin launch over the lane's streams and its read-back) is described
   by a signature of everything its launches depend on: the value type, the
   number of products, their total rows and tiles, the bin tables, the
   per-bin row counts of the fused pass and the addresses of the library's
   own buffers.  The second call with a signature captures the phase into a
   HIP graph; later calls with it replay the graph, one launch in place of
   one per kernel.  Operand and output addresses may change from call to
   call: the kernels read them from a device table that the graph refreshes
   from pinned host memory each time.  Results are identical either way.
   Profiling (spg_set_profiling) turns replay off while it is on.  Replay
   applies to spg_multiply / spg_multiply_batched; a server's lanes run
   eagerly.  While a call captures, HIP rejects synchronous calls on the
   legacy stream from every thread of the process: the library's own
   helpers wait for the call, other code must not run such calls beside
   a multiply with replay on. */
spg_status spg_set_graphs(int32_t enabled);
/* graphs captured and phase replays since the context was created */
spg_status spg_graph_stats(int64_t *captured, int64_t *replayed);

/* ---- served form ----------------------------------------------------
 * A server owns three lanes (stream sets with their own scratch), each
 * with a worker thread.  Submitted requests queue up; a free worker takes
 * the oldest and the next queued requests of its value type (up to
 * max_batch requests and 4 Mi nonzeros of A; a larger request runs alone),
 * runs their symbolic phase as one batch, and admits the batch once the
 * exact output memory it now knows (nnz(C_i) * (4 + sizeof(value)) summed)
 * fits the memory budget beside the batches the other lanes are running.
 * Lanes run concurrently: arrivals are taken while earlier batches are on
 * the GPU, and the phases of different batches overlap.
 * A batch that cannot fit even alone is split; a single request whose
 * output exceeds the budget fails with SPG_ERR_OUT_OF_MEMORY.
 * Operands must stay valid until the request's spg_server_wait returns. */
typedef struct spg_server spg_server;

spg_status spg_server_create(size_t memory_budget_bytes, int32_t max_batch, spg_server **out);
spg_status spg_server_submit(spg_server *server, const spg_csr *A, const spg_csr *B, spg_dtype dtype,
                             int64_t *ticket);
/* Blocks until the request is done; on success C (library-owned, free with
 * spg_csr_free) and the latency from submit to completion in milliseconds. */
spg_status spg_server_wait(spg_server *server, int64_t ticket, spg_csr *C, double *latency_ms);
/* batches run, requests served, largest batch, requests sent back to the
 * queue because their batch could not fit the budget even alone, and the
 * largest output memory of admitted batches running at once */
spg_status spg_server_stats(spg_server *server, int64_t *batches, int64_t *served, int64_t *max_batch_seen,
                            int64_t *deferred, int64_t *peak_bytes);
/* Finishes queued and running requests, then releases the server (results
 * never waited for are freed). */
spg_status spg_server_destroy(spg_server *server);

#ifdef __cplusplus
}
#endif

#endif /* GPUSPGEMM_H */
