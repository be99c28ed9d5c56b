/*
 * libodba_hip: MI355X-native DTW / DBA hot path of OpenDBA behind a C ABI.
 *
 * Plain pointers and sizes only. Host memory in, host memory out unless a
 * function says otherwise; the library owns device residency. Every function
 * returns 0 (ODBA_OK), a reference exit code (exit_codes.hpp:4-32) or an
 * ODBA_E_* code >= 100 (odba_types.h); odba_last_error() gives the message.
 * One odba_ctx per host thread; a ctx owns one HIP stream on one device.
 *
 * Value types: ODBA_I32 / ODBA_U32 / ODBA_U64 / ODBA_F32 / ODBA_F64, the
 * template instantiations of the reference (openDBA.cu:121-143). Sequences
 * are arrays of that type; lengths are element counts.
 */
#ifndef ODBA_H
#define ODBA_H

#include <stddef.h>
#include <stdint.h>

#include "odba_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct odba_ctx odba_ctx;
typedef struct odba_seqset odba_seqset;
typedef struct odba_dba odba_dba;

/* Context on HIP device `device` (replaces the implicit cudaSetDevice/stream
 * pairs of dba.hpp:54-63 and the device queries of cuda_utils.hpp:42-57). */
int odba_ctx_create(int device, odba_ctx** out);
int odba_ctx_destroy(odba_ctx* ctx);
const char* odba_last_error(const odba_ctx* ctx);
/* Compute units, peak engine clock (kHz), device memory (bytes), name. */
int odba_device_info(odba_ctx* ctx, int* compute_units, int* clock_khz, uint64_t* mem_bytes, char* name, int name_len);
/* Device memory this context plans its DBA working set against (default:
 * all of it). No counterpart in the reference, which plans against the
 * free memory of the whole device (dba.hpp:523-548); contexts that share a
 * device (serving slots) each get a share. */
int odba_ctx_set_mem_cap(odba_ctx* ctx, uint64_t bytes);

/*
 * All-vs-all DTW distances in the given order: pair (r, c > r) aligns
 * Y = seqs[r] (rows) against X = seqs[c] (columns) and dist_out[k] receives
 * the pair's distance at the row-major upper-triangle index
 * k = n(n-1)/2 - (n-r)(n-r-1)/2 + (c-r-1).
 * Replaces the DTWDistance<T> launch loop of approximateMedoidIndices
 * (dba.hpp:67-158, kernel dtw.hpp:39-259). ref_w is the reference swath
 * width (maxThreadsPerBlock, 1024); only 1024 is accepted.
 * The distance is sqrtf(cost), divided by |Y| when exactly one of
 * open_start / open_end is set (dtw.hpp:250-255), computed in float and
 * converted to the value type.
 */
int odba_dtw_all_pairs(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                       int use_open_start, int use_open_end, int ref_w, void* dist_out);

/* Device-resident sequence set: the sequences are copied to HBM once (in
 * the given order, padded) and reused by the calls below. */
int odba_seqset_create(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                       odba_seqset** out);
int odba_seqset_destroy(odba_seqset* set);
/* odba_dtw_all_pairs on a resident set. dist_out may be NULL (results stay
 * on the device); kernel_ms (optional) receives the kernel time measured
 * with HIP events on the ctx stream; flags_out (optional, n(n-1)/2 ints)
 * receives bit 0 = open-end search abrogated (bits 8.. = index of the swath
 * where it stopped), bit 1 = the reference would
 * read an unwritten LDS slot for this |Y| == 1 pair (dtw.hpp:155). */
int odba_dtw_all_pairs_resident(odba_ctx* ctx, odba_seqset* set, int use_open_start, int use_open_end,
                                void* dist_out, int* flags_out, double* kernel_ms);

/*
 * Bytes of the fixed per-device array odba_dtw_all_pairs_noalloc works in:
 * 256 for its pair counter, the rest for two edge columns of
 * (longest length + 256) values per resident wave, four waves per
 * workgroup. Every device the library runs on holds it until the process
 * ends.
 */
#define ODBA_NOALLOC_SCRATCH_BYTES (64u << 20)

/*
 * odba_dtw_all_pairs_resident with no device allocation. No counterpart in
 * the reference: it is for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). dist_dev (n(n-1)/2 values of the set's type, same index as
 * odba_dtw_all_pairs) and flags_dev (optional, n(n-1)/2 ints, bits as for
 * odba_dtw_all_pairs_resident) are DEVICE pointers on the context's device;
 * every element is written. The pairs run one wave each, with as many waves
 * resident as ODBA_NOALLOC_SCRATCH_BYTES holds; a set whose longest sequence
 * needs more than that for one workgroup (above about 2M samples of float,
 * 1M of double) is refused with ODBA_E_INVALID before anything runs.
 * Results equal odba_dtw_all_pairs_resident's. The array is shared by every
 * call on one device and keeps the last call's contents: calls that overlap
 * on one device (two contexts or threads at once) race and must not be
 * made. Returns when the distances are written.
 */
int odba_dtw_all_pairs_noalloc(odba_ctx* ctx, odba_seqset* set, int use_open_start, int use_open_end, void* dist_dev,
                               int* flags_dev);

/*
 * DBA rounds of this context are replayed as hipGraphs after a few eager
 * rounds (default on; no counterpart in the reference). Measured 1-3%
 * faster alone on a device, but slower when several contexts keep the same
 * device busy (the serving loop turns it off for its worker contexts).
 */
int odba_ctx_set_graphs(odba_ctx* ctx, int on);

/*
 * All-vs-all autotuner (no counterpart in the reference). Times every
 * candidate schedule on the resident set in this mode: one wave per pair in
 * both packing orders, and, when the set is eligible (float, or integers
 * whose costs cannot wrap; every sequence >= 64 samples) and has fewer than
 * 16 pairs per compute unit, bands of every strip width and band width that
 * the longest sequence admits. The fastest (median of `reps` kernel times)
 * is remembered in the context for sets of the same shape (value type,
 * mode, count, lengths) and used by later all-vs-all calls on them; the
 * results are the same under every schedule. best / best_ms / tried
 * (optional) receive the choice, its time and the number of candidates.
 */
int odba_tune_all_pairs(odba_ctx* ctx, odba_seqset* set, int use_open_start, int use_open_end, int reps,
                        odba_schedule* best, double* best_ms, int* tried);
/* The schedule the next all-vs-all on this set and mode will use. */
int odba_get_schedule(odba_ctx* ctx, odba_seqset* set, int use_open_start, int use_open_end, odba_schedule* out);
/* Pins a schedule for sets of this shape and mode (null: forget it). */
int odba_set_schedule(odba_ctx* ctx, odba_seqset* set, int use_open_start, int use_open_end,
                      const odba_schedule* sched);

/*
 * One pair with its step matrix: moves_out (optional, ly*lx bytes,
 * row-major) receives the reference step codes of every cell
 * (dtw.hpp:10-16; 0 where the reference leaves a cell unwritten after an
 * open-end abrogation), dist_out the pair distance as above. This is the
 * path-storing DTWDistance launch sequence of DBAUpdate (dba.hpp:589-630).
 */
int odba_dtw_path(odba_ctx* ctx, int dtype, const void* y, uint64_t ly, const void* x, uint64_t lx,
                  int use_open_start, int use_open_end, uint8_t* moves_out, void* dist_out);

/*
 * One DBA round (DBAUpdate<T>, dba.hpp:432-848) on the device: every
 * sequence aligned to the centroid (flipped when open_end and the centroid
 * is shorter, dba.hpp:497-501), traceback sums accumulated in the
 * reference's order, new_centroid[t] = sum / count (dba.hpp:820-822),
 * max_delta = max |old - new| (dba.hpp:830-836). counts_out is optional.
 * path_policy: ODBA_PATH_FULL, ODBA_PATH_STRIPE or ODBA_PATH_AUTO with the
 * free-memory figure ref_free_mem the reference's rule compares against
 * (dba.hpp:523-548); with ODBA_PATH_AUTO and ref_free_mem == 0 the figure is
 * the device's free memory once the sequences are resident (capped at the
 * context's memory cap), as the reference reads it with cudaMemGetInfo.
 */
int odba_dba_round(odba_ctx* ctx, int dtype, const void* centroid, uint64_t lc, const void* const* seqs,
                   const uint64_t* lens, int n, int use_open_start, int use_open_end, int path_policy,
                   uint64_t ref_free_mem, void* new_centroid, uint32_t* counts_out, double* max_delta);

/*
 * Device-resident DBA session for the round loop of performDBA
 * (dba.hpp:1084-1110): the cluster's sequences are uploaded once;
 * odba_dba_set_centroid loads a centroid (the medoid or a checkpoint),
 * odba_dba_step runs one DBAUpdate (as odba_dba_round) and keeps the new
 * centroid on the device for the next step; new_centroid / counts_out are
 * optional host copies.
 */
int odba_dba_create(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                    int use_open_start, int use_open_end, int path_policy, uint64_t ref_free_mem, odba_dba** out);
int odba_dba_destroy(odba_dba* dba);
int odba_dba_set_centroid(odba_dba* dba, const void* centroid, uint64_t lc);
int odba_dba_step(odba_dba* dba, void* new_centroid, uint32_t* counts_out, double* max_delta);

/*
 * Traceback recording for the path files DBAUpdate writes per sequence
 * (writeDTWPath, io_utils.hpp:285-323, called at dba.hpp:756,793): with
 * recording on, each odba_dba_step keeps the cells every traceback visits,
 * from the top-right corner to the (0,0)-row anchor, with their step codes
 * (the stripe-mode codes when a sequence took the stripe path, dba.hpp:756).
 * odba_dba_get_path returns them for sequence `seq` (session order) packed
 * as (i << 33) | (j << 3) | code, i = row, j = column of the aligned matrix
 * (the centroid is on the rows when the sequence was flipped), code 0 for
 * NIL and the reference code otherwise. steps == NULL only reports
 * *n_steps.
 */
int odba_dba_record_paths(odba_dba* dba, int on);
int odba_dba_get_path(odba_dba* dba, int seq, uint64_t* steps, uint64_t capacity, uint64_t* n_steps);

/*
 * performDBA's per-cluster consensus (dba.hpp:990-1110) for all clusters of
 * a clustering at once: sequences (already sorted and normalised as
 * performDBA has them), memberships[n] in 0..num_clusters-1, medoids[k] the
 * medoid index of cluster k (odba_cluster). Each round packs the path jobs
 * of every cluster that has not stopped into one launch per kernel; each
 * cluster stops on its own rule (delta == 0, the open-mode flip-flop test,
 * max_rounds). Singletons come back as they are (dba.hpp:1004-1041).
 * mu / sigma (optional, per sequence) rescale each centroid by its
 * medoid's (dba.hpp:1117-1125). centroids_out holds the centroids
 * back to back at centroid_offs[k] (num_clusters + 1 offsets, lengths =
 * the medoids' lengths); rounds_out[k] = rounds run for cluster k.
 */
int odba_dba_clusters(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                      const int* memberships, const int* medoids, int num_clusters, int use_open_start,
                      int use_open_end, int max_rounds, const double* mu, const double* sigma, void* centroids_out,
                      uint64_t* centroid_offs, int* rounds_out);

/*
 * Clustering and medoids of approximateMedoidIndices (dba.hpp:219-351) on
 * the all-vs-all distances (row-major upper triangle, value type dtype):
 * distances divided by their maximum, complete-linkage hierarchical
 * clustering (hclust_fast), then cdist > 1: K clusters with K raised until
 * cdist of them have two or more members (cutree_k); cdist == 1: one
 * cluster; 0 <= cdist < 1: cut at that dendrogram height (cutree_cdist).
 * memberships[n]; medoids[] (capacity n) per cluster with the reference's
 * rules; *num_clusters. Runs on the host (no device needed).
 */
int odba_cluster(int dtype, const void* dists, int n, const uint64_t* lens, double cdist, int* memberships,
                 int* medoids, int* num_clusters);
/* The dendrogram alone: complete linkage of a condensed double distance
 * matrix, R-convention merge ((n-1) x 2, column-major) and heights, and the
 * cut into nclust clusters. */
int odba_hclust_complete(int n, const double* condensed, int* merge, double* height);
int odba_cutree_k(int n, const int* merge, int nclust, int* labels);

/*
 * Z-normalise sequences in place (normalizeSequences, gpu_utils.hpp:131-170):
 * population sigma, sigma == 0 treated as 1; mu / sigma (optional) receive
 * each sequence's mean and standard deviation before that substitution.
 */
int odba_znorm(odba_ctx* ctx, int dtype, void* const* seqs, const uint64_t* lens, int n, double* mu, double* sigma);

/*
 * performDBA<T> for one cluster (cdist == 1, CLUSTER_AND_CONSENSUS;
 * dba.hpp:863-1189) as one call: sanitize, stable length sort, optional
 * z-norm, all-vs-all medoid search, DBA rounds until delta == 0 (or the
 * open-mode flip-flop stop, or max_rounds), rescale by the medoid's mu and
 * sigma. centroid_out must hold max(lens) elements; perm_out (optional, n
 * ints) receives the sort permutation; dists_out (optional, n(n-1)/2
 * values) the pair distances in sorted order.
 */
int odba_perform_dba(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                     int use_open_start, int use_open_end, int norm, int max_rounds, int path_policy,
                     uint64_t ref_free_mem, void* centroid_out, uint64_t* centroid_len, int* medoid_sorted,
                     int* perm_out, int* rounds_out, void* dists_out);

/*
 * Prefix location (chopPrefixFromSequences<T>, dba.hpp:1192-1395): each
 * sequence is aligned against the prefix (prefix on Y, open end, swath width
 * 1024, path kept), and the path is walked back from the top-right corner.
 * chop_pos[i] = the column where the walk leaves the top row's open-right
 * run; hist[i * prefix_len + r] = the number of path cells in prefix row r
 * from there down to (0,0) (the reference's leader-path histogram). With
 * norm, the sequences are z-normalised in place first and the prefix is
 * z-normalised in a private copy, as the reference does. The caller cuts
 * each sequence at chop_pos (the reference keeps seq[chop_pos:]).
 * prefix_len >= 2.
 */
int odba_prefix_chop(odba_ctx* ctx, int dtype, const void* prefix, uint64_t prefix_len, void* const* seqs,
                     const uint64_t* lens, int n, int norm, uint64_t* chop_pos, int32_t* hist);

/*
 * Adaptive k-segmentation (adaptive_segmentation<T>, segmentation.hpp:339-583;
 * kernel adaptive_device_segmentation<T>, segmentation.hpp:48-337): each
 * sequence is cut into chunks of 256*ceil(minseg/3) samples, every chunk is
 * segmented by the Bellman DP with the largest non-noisy segment count, and
 * the segment medians, after the reference's epsilon merge of chunk borders
 * and the prefix skip, are written to out (concatenated; out_offs / out_lens
 * per sequence). lds_bytes is the shared-memory budget the reference sizes
 * its tables with (the CUDA device's opt-in maximum per block, 48 KiB when
 * unknown); it sets the longest segment the DP considers.
 * out_capacity (elements) must be at least odba_kseg_capacity(...).
 */
int odba_kseg(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n, int min_segment_length,
              int prefix_skip, int lds_bytes, void* out, uint64_t out_capacity, uint64_t* out_lens,
              uint64_t* out_offs);
/* Upper bound on the total number of segments odba_kseg can produce. */
uint64_t odba_kseg_capacity(const uint64_t* lens, int n, int min_segment_length);

/*
 * Packed DBA rounds across jobs (no counterpart in the reference; for
 * servers that batch the rounds of concurrent jobs). odba_job_prepare runs
 * the first half of odba_perform_dba for one job on ctx's stream -
 * sanitize, stable length sort, upload, z-norm (norm != 0), all-vs-all,
 * medoid - and keeps its sequences resident on ctx's device. A rounds
 * engine holds prepared jobs of one value type and one (open start, open
 * end) mode; each odba_rounds_step runs one DBA round (full path matrices)
 * of every job it holds, packed into shared launches on the engine
 * context's stream. Each job keeps its own stop rule (delta 0, the open-mode
 * repeat test, max_rounds) and its result (centroid rescaled by the
 * medoid's mu/sigma when normalised, medoid index in sorted order, rounds)
 * equals odba_perform_dba's with ODBA_PATH_FULL. Jobs that ended are
 * returned by odba_rounds_step (up to `capacity` per call, the rest on
 * later calls); n_active counts jobs held or not yet returned. An engine
 * is used by one thread at a time; jobs may be prepared on other contexts
 * of the same device. odba_job_destroy frees a job's device memory (also
 * the engine's buffers for it) and must not be called while an engine
 * holds the job.
 * A job's step words (2 bits per path-matrix cell, every swath of every
 * member, and the open-end top-row words) lie in a buffer of its own that
 * lives as long as the job: odba_job_words gives its size in bytes and,
 * with allocate != 0, allocates it on the calling thread (a server does
 * this on its preparing thread once the bytes fit its memory budget);
 * odba_rounds_add allocates it when that has not been done.
 * odba_job_shape gives a prepared job's centroid length (its medoid's) and
 * the largest number of 1024-column swaths of its path matrices.
 */
typedef struct odba_job odba_job;
typedef struct odba_rounds odba_rounds;
int odba_job_prepare(odba_ctx* ctx, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                     int use_open_start, int use_open_end, int norm, int max_rounds, odba_job** out);
int odba_job_result(const odba_job* job, void* centroid_out, uint64_t capacity, uint64_t* centroid_len,
                    int* medoid_sorted, int* rounds, int* job_rc);
int odba_job_words(odba_job* job, int allocate, uint64_t* word_bytes);
int odba_job_shape(const odba_job* job, uint64_t* centroid_len, int* max_swaths);
int odba_job_destroy(odba_job* job);
int odba_rounds_create(odba_ctx* ctx, int dtype, int use_open_start, int use_open_end, odba_rounds** out);
int odba_rounds_add(odba_rounds* engine, odba_job* job);
int odba_rounds_step(odba_rounds* engine, odba_job** finished, int capacity, int* n_finished, int* n_active);
int odba_rounds_destroy(odba_rounds* engine);

/*
 * Serving (no counterpart in the reference, which runs one job per process):
 * DBA jobs - the performDBA pipeline for one cluster, as odba_perform_dba
 * with the full path matrix - are submitted at any time, admitted in
 * arrival order when one of `slots` worker contexts is free and the job's
 * estimated device working set fits in what is left of hbm_budget (0: 80%
 * of device memory), and run concurrently, each on its own HIP stream.
 * A job's result is the same alone and in a mix. Times are seconds since
 * the server was created: times[0] submit, [1] admission, [2] completion.
 * slots: 1..32. The slots' streams overlap only if the process has a
 * hardware queue per stream: set GPU_MAX_HW_QUEUES >= slots (HIP's default
 * is 4) in the environment before the process first uses HIP.
 */
typedef struct odba_server odba_server;
int odba_server_create(int device, uint64_t hbm_budget, int slots, odba_server** out);
int odba_server_destroy(odba_server* server);  /* waits for running jobs; queued ones fail */
int odba_server_submit(odba_server* server, int dtype, const void* const* seqs, const uint64_t* lens, int n,
                       int use_open_start, int use_open_end, int norm, int max_rounds, uint64_t* job_id);
/* Blocks until the job is done or failed (timeout_s < 0: no limit). */
int odba_server_wait(odba_server* server, uint64_t job_id, double timeout_s, int* state);
int odba_server_result(odba_server* server, uint64_t job_id, void* centroid_out, uint64_t capacity,
                       uint64_t* centroid_len, int* medoid_sorted, int* rounds, double* times, int* job_rc);
int odba_server_release(odba_server* server, uint64_t job_id);
int odba_server_stats(odba_server* server, uint64_t* bytes_in_use, uint64_t* budget, int* running, int* queued);

#ifdef __cplusplus
}
#endif

#endif /* ODBA_H */
