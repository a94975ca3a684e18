/*
 * ppsk.h -- C ABI of the MI355X (gfx950) core/accessory distance library
 * (libppsk.so).  Plain C: opaque handles, pointers and sizes; no C++ types
 * and no exceptions cross this boundary.  Every call returns PPSK_OK or an
 * error code; ppsk_last_error() gives the message of the calling thread's
 * last failure.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the pp-sketchlib source tree).  The reference's C++ layer
 * (query_db_cuda / query_db_sparse_cuda, src/gpu/gpu_api.cpp:352-531) drives
 * its CUDA internals through src/gpu/gpu.hpp:36-73; a drop-in host layer
 * drives this ABI from the same place instead.
 *
 * Sketch words are passed in the reference's own layout: for one sample and
 * one k-mer length, sketchsize64 * bbits uint64 words, word w / bit layer l at
 * index w * bbits + l (src/sketch/sketch.cpp:82-92).  bbits must be 14
 * (src/reference.hpp:18, the only value the reference sketcher writes).
 */
#ifndef PPSK_H
#define PPSK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PPSK_OK 0
#define PPSK_ERR_ARG 1     /* invalid argument / incompatible sketches      */
#define PPSK_ERR_HIP 2     /* HIP runtime or kernel failure                  */
#define PPSK_ERR_MEMORY 3  /* device memory exhausted                        */
#define PPSK_ERR_FIT 4     /* k-mer regression had < 2 valid points          */
#define PPSK_ERR_STATE 5   /* operation not valid in the current state       */
#define PPSK_ERR_CANCELLED 6 /* stopped by ppsk_cancel; outputs undefined     */

/* Random-match correction modes (src/random/random_match.hpp:25). */
#define PPSK_RANDOM_NONE 0       /* RandomMC(): r = 0                          */
#define PPSK_RANDOM_EQUAL_BASE 1 /* RandomMC(use_rc): csrs closed form         */
#define PPSK_RANDOM_MC 2         /* Monte Carlo table, per-sample clusters     */

/* Dense output kinds. */
#define PPSK_OUT_CORE_ACC 0 /* float32 [rows][2] (core, accessory)            */
#define PPSK_OUT_JACCARD 1  /* float32 [rows][n_k] random-corrected Jaccards  */
#define PPSK_OUT_COUNTS 2   /* uint16  [rows][n_k] raw bin-match counts        */

typedef struct ppsk_set ppsk_set; /* resident sketch set on one device */

/* Message of this thread's last failed call ("" if none). */
const char *ppsk_last_error(void);

/* Library version, e.g. "1.0.0". */
const char *ppsk_version(void);

/*
 * Select and initialise a device; report free/total HBM and LDS per
 * workgroup.  Replaces initialise_device (src/gpu/dist.cu:373-406,
 * declared src/gpu/gpu.hpp:36).
 */
int ppsk_device_init(int device_id, size_t *mem_free, size_t *mem_total,
                     size_t *shared_size);

/*
 * Progress and cancellation of the distance calls on a device (ppsk_dense,
 * ppsk_dense_noalloc, ppsk_knn, ppsk_threshold, ppsk_self_pass,
 * ppsk_batch_run), as the reference's GPU progress poll and kill flag
 * (src/gpu/dist.cu:351-370, src/gpu/cuda.cuh:77-87) and its CPU path's
 * Ctrl-C check (src/api.cpp:217-226, 285-287).  Calls on one device run one
 * at a time; these three may be called from any thread at any time, also
 * while such a call runs, and need no HIP device (device_id in [0, 64)).
 *
 * ppsk_progress: (tile, k-mer length) work items finished and planned by the
 * call running on the device (*running = 1), or by the last one; *pending = 1
 * while a cancellation request stands.  Any output pointer may be NULL.
 * ppsk_cancel: asks the call running on the device to stop, and every call
 * that starts on the device to stop at once, until ppsk_cancel_reset.  The
 * kernels read the request between work items and finish normally (no
 * trap); the call waits for them and returns PPSK_ERR_CANCELLED, its outputs
 * undefined.
 */
int ppsk_progress(int device_id, uint64_t *done, uint64_t *total,
                  int32_t *running, int32_t *pending);
int ppsk_cancel(int device_id);
int ppsk_cancel_reset(int device_id);

/*
 * Create a resident sketch set of n_samples samples with the given k-mer
 * lengths (ascending) and sketch size.  Replaces the flattening and
 * host->device copies of flatten_by_bins / flatten_by_samples and
 * DeviceMemory (src/gpu/gpu.hpp:38-49, src/gpu/device_memory.cu:6-65): the
 * set is kept in HBM in the library's own stripe layout and reused across
 * queries.
 */
int ppsk_set_create(int device_id, int64_t n_samples, int32_t n_k,
                    const int32_t *klist, int32_t sketchsize64, int32_t bbits,
                    ppsk_set **out);

/* Upload samples [first, first + count): raw is [count][n_k][s64 * bbits]. */
int ppsk_set_upload(ppsk_set *set, int64_t first, int64_t count,
                    const uint64_t *raw);

/*
 * Per-sample random-match inputs.  lengths (may be NULL when the equal-base
 * mode is unused): sequence lengths, from which the library precomputes the
 * csrs terms j(len) = 1 - (1 - 2^-((k-1)*2+rc))^len per k
 * (src/random/random_match.cpp:318-336; k >= 33 is rejected, the reference
 * shift is undefined there).  clusters (may be NULL): Monte Carlo cluster id
 * per sample (RandomMC::lookup_array, random_match.cpp:254-273).
 */
int ppsk_set_meta(ppsk_set *set, const uint64_t *lengths, int32_t use_rc,
                  const uint16_t *clusters);

/*
 * A read-only view of the first count samples of a set (no copy): a drop-in
 * ppsk_set for every call, sharing the parent's device buffers and
 * metadata.  Free it with ppsk_set_free (which then frees only the view)
 * before the parent.  Lets many jobs query prefixes of one resident
 * database, as the reference's sample lists select sub-databases
 * (src/api.cpp:436-516).
 */
int ppsk_set_prefix(ppsk_set *parent, int64_t count, ppsk_set **view);

int ppsk_set_free(ppsk_set *set);

int64_t ppsk_set_size(const ppsk_set *set);

/*
 * Dense distances.  query == NULL (or == ref): self mode, rows are the
 * condensed upper triangle, row = square_to_condensed(i, j, n) for i < j.
 * Otherwise rows are query-major, row = q * n_ref + r.  The Jaccard
 * denominator comes from the row sample (r1 = i in self mode, r1 = query in
 * query mode, src/api.cpp:195-268).  Random correction: mode and, for
 * PPSK_RANDOM_MC, a [n_k][C][C] float table indexed
 * [cluster(r1 of random_match)][cluster(other)] with zeros where
 * k >= max_k (random_match.cpp:185-206).  tolerance: the regression validity
 * threshold (src/dist/linear_regression.cpp:22).
 * out: host memory (out_on_device = 0) or device memory from
 * ppsk_device_alloc (out_on_device = 1).
 * Pairs whose regression window has < 2 points get (0, 0), are counted in
 * *n_fail (may be NULL) with the smallest such row in *first_fail_row, and
 * the call returns PPSK_ERR_FIT after writing every other row.
 * Replaces dispatchDists (src/gpu/dist.cu:408-489, src/gpu/gpu.hpp:51-61).
 */
int ppsk_dense(ppsk_set *ref, ppsk_set *query, int32_t random_mode,
               int32_t n_clusters, const float *mc_table, double tolerance,
               int32_t out_kind, void *out, int32_t out_on_device,
               uint64_t *n_fail, int64_t *first_fail_row);

/*
 * Size of the fixed device scratch array of ppsk_dense_noalloc (64 MiB; one
 * copy per GPU the library runs on, held until the process ends).
 */
#define PPSK_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)

/*
 * ppsk_dense with device output that allocates no device memory.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, modes, output order and results as ppsk_dense
 * with out_on_device = 1 (out_device: caller-owned device memory of the
 * output size).  Its working memory (fit counters, the Monte Carlo table, the
 * tile list and the per-tile match counts of each launch) is a fixed array of
 * PPSK_NOALLOC_SCRATCH_BYTES in device memory: any sample count is accepted
 * (the pass is cut into launches that fit), a Monte Carlo table larger than
 * the array is rejected with PPSK_ERR_ARG.  The device must have been
 * initialised (ppsk_device_init or any earlier call on it), else
 * PPSK_ERR_STATE.  Every call on one GPU shares that GPU's array: calls are
 * serialised by the library within one process; calls that overlap on one GPU
 * from outside that lock would race.
 */
int ppsk_dense_noalloc(ppsk_set *ref, ppsk_set *query, int32_t random_mode,
                       int32_t n_clusters, const float *mc_table, double tolerance,
                       int32_t out_kind, void *out_device, uint64_t *n_fail,
                       int64_t *first_fail_row);

/*
 * Self kNN, as query_db_sparse: for every sample i the kNN smallest of
 * column dist_col (0 core, 1 accessory) over all j (i == j counts as +inf),
 * ascending per row, i_vec/j_vec/dists of n * kNN entries.  With jaccard
 * set, the ranked distance is 1 - J at k-mer length index dist_col and the
 * returned values are J again (src/api.cpp:368-371, 423-429).  Selection and
 * tie order reproduce the reference's std::priority_queue scan over j
 * (src/api.cpp:319-432).  Requires 1 <= kNN <= n; a row's heap lives in LDS
 * up to kNN = 2048 and in device memory above.
 * Replaces sparseDists (src/gpu/dist.cu:495-791, src/gpu/gpu.hpp:63-73).
 */
int ppsk_knn(ppsk_set *ref, int32_t random_mode, int32_t n_clusters,
             const float *mc_table, double tolerance, int32_t jaccard,
             int32_t kNN, int32_t dist_col, int64_t *i_vec, int64_t *j_vec,
             float *dists, uint64_t *n_fail);

/*
 * Self threshold sparsification, as querySelfSparse with dist_cutoff > 0:
 * upper triangle (i < j) entries of column dist_col with d < cutoff, in
 * row-major order (src/sketchlib_bindings.cpp:159-169,
 * src/dist/matrix_ops.cpp:38-69).  The library allocates *i_vec, *j_vec and
 * *dists (n_out entries each); release them with ppsk_free_host.
 */
int ppsk_threshold(ppsk_set *ref, int32_t random_mode, int32_t n_clusters,
                   const float *mc_table, double tolerance, int32_t dist_col,
                   float cutoff, int64_t **i_vec, int64_t **j_vec,
                   float **dists, int64_t *n_out, uint64_t *n_fail);

/*
 * One self pass serving all three outputs of a large all-vs-all: the
 * condensed core/accessory column pair (n(n-1)/2 float2 rows, streamed to the
 * host buffer dense_out in row blocks while later blocks compute; NULL skips
 * it), the kNN lists of ppsk_knn (kNN = 0 skips them) and the threshold
 * lists of ppsk_threshold (cutoff <= 0 skips them), all from one distance
 * computation kept in device memory.  Results equal the three separate
 * calls.  Needs a symmetric random-match table (always true except for an
 * asymmetric Monte Carlo table).  dense_out from ppsk_host_alloc makes the
 * copies fully asynchronous.  Replaces the reference's self queryDatabase
 * followed by querySelfSparse over the same sketches
 * (src/sketchlib_bindings.cpp:89-183, src/gpu/dist.cu:408-489).
 */
int ppsk_self_pass(ppsk_set *ref, int32_t random_mode, int32_t n_clusters,
                   const float *mc_table, double tolerance, float *dense_out,
                   int32_t kNN, int32_t knn_col, int64_t *knn_i,
                   int64_t *knn_j, float *knn_d, float cutoff,
                   int32_t thr_col, int64_t **thr_i, int64_t **thr_j,
                   float **thr_d, int64_t *thr_n, uint64_t *n_fail);

void ppsk_free_host(void *ptr);

/*
 * Page-locked host memory (hipHostMalloc) for dense outputs, so that the
 * row-block copies overlap the distance kernels.
 */
int ppsk_host_alloc(size_t bytes, void **ptr);
int ppsk_host_free(void *ptr);

/* Device memory for out_on_device outputs. */
int ppsk_device_alloc(int device_id, size_t bytes, void **ptr);
int ppsk_device_free(void *ptr);
int ppsk_copy_to_host(void *dst, const void *device_src, size_t bytes);

/*
 * Batched query-vs-ref jobs in one persistent launch.  Each job is a
 * ppsk_dense call in query mode (rows = query samples, query-major output,
 * r1 = query) with its own sets, k-mer lengths, sketch size, random
 * correction, tolerance and device output buffer (ppsk_device_alloc, sized
 * as for ppsk_dense); a job whose query is its ref (the same handle) is a
 * self job (condensed upper triangle, as ppsk_dense with query == NULL).  The reference runs such jobs one queryDatabase /
 * query_db_cuda call at a time (src/gpu/gpu_api.cpp:352-452); results here
 * are identical to those single calls.
 */
typedef struct ppsk_job {
  ppsk_set *ref;        /* columns */
  ppsk_set *query;      /* rows */
  int32_t random_mode;  /* PPSK_RANDOM_* */
  int32_t n_clusters;   /* PPSK_RANDOM_MC: table [n_k][C][C] */
  const float *mc_table;
  double tolerance;
  int32_t out_kind;     /* PPSK_OUT_CORE_ACC / JACCARD / COUNTS */
  void *out;            /* device memory */
} ppsk_job;

typedef struct ppsk_batch ppsk_batch;

/*
 * Plan a batch: validates every job, uploads per-job parameters and tables.
 * The sets must stay alive (and unchanged) while the batch exists.
 */
int ppsk_batch_create(int device_id, const ppsk_job *jobs, int32_t n_jobs,
                      ppsk_batch **out);

/*
 * Run every job: one graph launch (queue reset + persistent kernel), replayed
 * on later runs.  variant: the plan -- every job in one tile shape (2:
 * 64 x 256, 4: 32 x 256, 5: 16 x 256), 6: each job's rows split over the
 * three tile heights (row tiles laid from its first row, each at a multiple
 * of its own height; one persistent launch running every shape), or 0 for
 * the autotuner.  On the first run of a batch with an unseen job-shape
 * signature the autotuner times plans 2, 4 and 5 on the batch itself,
 * derives each shape's throughput on padded work, splits every job's rows
 * for the least estimated time, times that plan 6 too and keeps the fastest
 * plan for the process.  Before tuning, plan 6 splits each job's rows for
 * the least padding.  n_fail (may be NULL): per-job count of pairs whose
 * regression failed (PPSK_ERR_FIT is returned after all output is written).
 * used_variant / kernel_ms (may be NULL): plan used and its launch time.
 */
int ppsk_batch_run(ppsk_batch *batch, int32_t variant, uint64_t *n_fail,
                   int32_t *used_variant, double *kernel_ms);

/* Autotuner result: chosen plan and per-plan time (index = plan, 0..6;
 * 0 where not timed); row_tiles (may be NULL, 3 x n_jobs entries): for each
 * job, the row tiles of 64, 32 and 16 rows of its plan-6 split (after
 * tuning: the split of least estimated time; plan 6 is not timed when that
 * is one shape for every job). */
int ppsk_batch_tuning(const ppsk_batch *batch, int32_t *variant,
                      double *ms_per_variant, int32_t *row_tiles);

int ppsk_batch_free(ppsk_batch *batch);

/*
 * Serving: a resident reference set answers query batches that arrive at any
 * time.  Each request returns, per query, the kNN nearest references by
 * column knn_col (0 core, 1 accessory; the heap order of ppsk_knn, no sample
 * excluded) and every reference with column thr_col below cutoff (row-major,
 * as ppsk_threshold), with r1 = the query (query mode of ppsk_dense).  The
 * reference's closest interface is queryDatabase followed by a sparse
 * extraction, one call per query batch (src/sketchlib_bindings.cpp:89-183).
 *
 * Continuous batching: n_slots workers, each with its own stream, fill steps
 * of at most max_rows query samples with queued requests, oldest first, while
 * the step's device memory fits the serving arena (arena_bytes; 0 = free HBM
 * less 4 GiB).  A request larger than the room left in a step is split: the
 * step takes its next rows and the rest stays at the head of the queue, so
 * requests of any size are accepted and a large one runs on several steps at
 * once.  Requests arriving while other steps run join the next step that has
 * room, and steps of different workers overlap on the GPU.  max_rows bounds
 * a step's duration and with it the latency a step adds to the requests
 * queued behind it.  A request's results do not depend on the steps it
 * joins.
 * random_mode: PPSK_RANDOM_NONE, PPSK_RANDOM_EQUAL_BASE (query lengths passed
 * at submit) or PPSK_RANDOM_MC: the [n_k][C][C] table of ppsk_dense (zeros
 * where k >= max_k), indexed [cluster of the reference][cluster of the query]
 * as the reference's query path does (random_match(r1, q_cluster, q_len, k),
 * src/random/random_match.cpp:185-197, src/api.cpp:237-268), with the
 * references' clusters from ppsk_set_meta and each query's cluster (its
 * nearest centroid, RandomMC::closest_cluster, random_match.cpp:230-237)
 * passed at submit.  The reference set must outlive the server.
 */
typedef struct ppsk_server ppsk_server;

typedef struct ppsk_result {
  int64_t n_queries;
  int64_t knn_n; /* n_queries * kNN; i = query index within the request */
  int64_t *knn_i, *knn_j;
  float *knn_d;
  int64_t thr_n;
  int64_t *thr_i, *thr_j;
  float *thr_d;
  uint64_t n_fail;  /* pairs whose regression failed (distances 0, 0) */
  double t_submit, t_start, t_done; /* seconds since server creation */
  int64_t step_id;   /* step that served the request's first rows */
  int64_t step_rows; /* query samples in that step */
  int64_t n_steps;   /* steps that carried the request's rows */
  void *owner;       /* library-internal */
} ppsk_result;

int ppsk_server_create(ppsk_set *ref, int32_t random_mode, int32_t n_clusters,
                       const float *mc_table, double tolerance, int32_t kNN,
                       int32_t knn_col, float cutoff, int32_t thr_col,
                       int32_t max_rows, int32_t n_slots, size_t arena_bytes,
                       ppsk_server **out);

/* Queue a request: raw [n_queries][n_k][s64 * 14] in the reference layout,
 * lengths [n_queries] (equal-base mode, else may be NULL), clusters
 * [n_queries] (Monte Carlo mode, else may be NULL).  Copies the inputs;
 * never blocks on the GPU. */
int ppsk_server_submit(ppsk_server *server, int64_t n_queries,
                       const uint64_t *raw, const uint64_t *lengths,
                       const uint16_t *clusters, int64_t *ticket);

/* Block until the request is served; release the result with
 * ppsk_result_free.  Each ticket can be collected once. */
int ppsk_server_wait(ppsk_server *server, int64_t ticket, ppsk_result **out);
void ppsk_result_free(ppsk_result *result);

int ppsk_server_stats(ppsk_server *server, int64_t *steps, int64_t *rows,
                      int64_t *requests, int32_t *peak_inflight,
                      size_t *peak_arena_bytes, double *gpu_ms);

/* Stops the workers.  Requests not yet served fail: a ppsk_server_wait in
 * progress returns PPSK_ERR_STATE, and free returns only after every waiter
 * has left.  No call on the server may start once free has begun. */
int ppsk_server_free(ppsk_server *server);

/*
 * Timing of the last dense/kNN/threshold call on this thread, measured with
 * HIP events on the launch stream: summed duration of the distance-tile
 * kernel launches and their number.
 */
int ppsk_last_timing(double *kernel_ms, int64_t *n_launches, int64_t *n_items);

/*
 * Sketch sequences on the device: the bindash sketch over canonical
 * (use_rc) or forward ntHash values of every k-mer free of non-ACGTU
 * characters, densification and bit slicing, as sketch()
 * (src/sketch/sketch.cpp:118-176) with dense (not codon-phased) seeds and no
 * read filtering.  seq holds every sample's contigs back to back; contig c
 * is seq[contig_off[c] .. contig_off[c+1]); sample s owns contigs
 * sample_contig[s] .. sample_contig[s+1]-1 (contigs shorter than k yield no
 * k-mers; the reference's reader already drops contigs shorter than the
 * largest k, src/api.cpp:87).  raw_out: [n_samples][n_k][s64 * 14] words in
 * the reference layout; inv_minhash (may be NULL): per sample the sum over k
 * of bin 0 / (2^61 - 1) before densification (the reads size estimate is
 * n_k / that sum, src/reference.cpp:49-55); densified (may be NULL): per
 * sample 1 if any k needed densification.  klist: ascending, 3 <= k <= 255;
 * 1 <= sketchsize64 <= 1023.  Replaces create_sketches / sketch_gpu
 * (src/api.cpp:39-138, src/sketch/sketch.cpp:178-205).
 */
int ppsk_sketch(int device_id, int64_t n_samples, const char *seq,
                int64_t seq_bytes, const int64_t *contig_off,
                int64_t n_contigs, const int64_t *sample_contig, int32_t n_k,
                const int32_t *klist, int32_t sketchsize64, int32_t use_rc,
                uint64_t *raw_out, double *inv_minhash, int32_t *densified);

/*
 * ppsk_sketch with the reference's two sketching options.
 * codon_phased: the codon-phased seeds of generate_seeds
 * (src/sketch/sketch.cpp:43-54): a span of 3k - 2 bases of which every
 * third (offsets 0, 3, ..., 3(k-1)) is hashed, the span free of non-ACGTU
 * characters (src/sketch/nthash.hpp:634-722); 3k <= 255.
 * min_count > 0 (1..255): the reads count filter
 * (src/sketch/sketch.cpp:128-149, src/sketch/countmin.cpp:37-85): every
 * valid k-mer occurrence is counted in sequence order and binned only if its
 * count so far reaches min_count.  exact = 0: the 4-row count-min sketch of
 * 2^27 - 1 cells per row, with the reference's second hash (dense seeds: that
 * of the first k-mer of the current run of valid k-mers; codon-phased: that
 * of the k-mer itself) and its flat cell layout; exact = 1: counts per 64-bit
 * hash (HashCounter).  At most 2^29 - 1 windows per sample and k when
 * counting.  min_count = 0 and codon_phased = 0 is ppsk_sketch.  Replaces
 * sketch() with codon-phased seeds or a read counter and the GPU count-min
 * sketcher (src/gpu/sketch.cu:140-176, src/gpu/gpu_countmin.cu).
 */
int ppsk_sketch_ex(int device_id, int64_t n_samples, const char *seq,
                   int64_t seq_bytes, const int64_t *contig_off,
                   int64_t n_contigs, const int64_t *sample_contig, int32_t n_k,
                   const int32_t *klist, int32_t sketchsize64, int32_t use_rc,
                   int32_t codon_phased, int32_t min_count, int32_t exact,
                   uint64_t *raw_out, double *inv_minhash, int32_t *densified);

/*
 * Seeded synthetic sketches (no reference counterpart: a generator for tests
 * and benchmarks, writing the reference word layout directly).  Bins evolve
 * down a random binary tree: a child keeps its parent's 14-bit bin value with
 * probability (1 - dc)^k (1 - da), else draws a fresh uniform value; edge
 * rates dc ~ U[0, dcore_max / D], da ~ U[0, dacc_max / D], D = ceil(log2 n).
 * raw: [n][n_k][s64 * 14] uint64; lengths: [n] (may be NULL); base_freq:
 * [n][4] ACGT (may be NULL).  Deterministic for any n_threads (0 = all).
 */
int ppsk_synth_sketches(int64_t n, int32_t n_k, const int32_t *klist,
                        int32_t sketchsize64, uint64_t seed, double dcore_max,
                        double dacc_max, double length_mean, uint64_t *raw,
                        uint64_t *lengths, double *base_freq, int32_t n_threads);

#ifdef __cplusplus
}
#endif

#endif /* PPSK_H */
