/*
 * gamma_smc_amd -- C ABI of the MI355X (gfx950) Gamma-SMC decoder.
 *
 * Plain pointers and 64-bit sizes; no C++ or framework types cross this
 * boundary. Every entry point returns 0 on success or a negative status
 * (GSMC_E*) and leaves a message for gsmc_last_error() (per thread). No
 * exception crosses the ABI. A context is bound to the device that was
 * current when it was created; calls on one context are serialised by the
 * context itself, calls on different contexts may run concurrently.
 *
 * Each entry point names the reference interface it replaces, in the
 * reference tree of kevinkorfmann/gamma_smc_cu (file:line).
 */
#ifndef GAMMA_SMC_AMD_H
#define GAMMA_SMC_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GSMC_OK 0
#define GSMC_EINVAL -1     /* invalid argument (reference: std::invalid_argument -> ValueError) */
#define GSMC_EHIP -2       /* HIP runtime error (reference: CUDA_CHECK -> RuntimeError)        */
#define GSMC_ENOMEM -3     /* device or pinned host allocation failed                          */
#define GSMC_EIO -4        /* flow-field file missing or malformed                             */

/* run flags */
#define GSMC_CI 1          /* also write Wilson-Hilferty lower / upper (mean_only=False)     */
#define GSMC_POSTERIOR 2   /* also write posterior alpha / beta (return_posterior=True)       */
#define GSMC_SUMMARY 4     /* batch jobs: write per-site mean / min / max over the pairs
                              (site_mean, site_min, site_max) instead of per-pair rows       */
#define GSMC_Q8 8          /* mean receives uint8 log-quantized codes [rows][P] instead of
                              floats: code = clamp(floor(ln(max(mean, 1)) / ln(20 Ne) * 255
                              + 0.5), 0, 255) (gamma_smc.cu:23-39, bindings.cpp:1680-1810)  */
#define GSMC_Q4 16         /* uint4 codes (levels 15), two rows per byte, even row in the
                              low nibble: [ceil(rows / 2)][P]; block cores must be even     */

/* table precision */
#define GSMC_TABLES_F32 0  /* reference-identical FP32 table construction                    */
#define GSMC_TABLES_F64 1  /* FP64 construction, rounded to FP32 when stored                  */

typedef struct gsmc_context gsmc_context;
typedef struct gsmc_batch gsmc_batch;
typedef struct gsmc_graph gsmc_graph;

/* One job of a batched launch: pairs decoded over sites [site_lo, site_hi)
 * of the context as one sequence (its first site's gap measured from
 * position -1, as for a cohort holding only those sites), split into block
 * windows of core_block_sites with flank_sites flanks (core_block_sites <= 0:
 * one window). Pointers are device memory; outputs are
 * [site_hi - site_lo][n_pairs]. Pair ids are read on the device without a
 * host check: ids outside [0, n_haps) are clamped and give undefined rows. */
typedef struct gsmc_job {
    int64_t site_lo, site_hi;
    int64_t core_block_sites, flank_sites;
    const int32_t* pair_i;
    const int32_t* pair_j;
    int64_t n_pairs;
    float* mean;
    float* lower;
    float* upper;
    float* alpha;
    float* beta;
    /* GSMC_SUMMARY launches: [site_hi - site_lo] per-site reductions over the
     * job's pairs (site_min / site_max may be NULL); per-pair outputs unused. */
    float* site_mean;
    float* site_min;
    float* site_max;
} gsmc_job;

/* Message of the last failed call on this thread ("" if none). */
const char* gsmc_last_error(void);

/* Build the transition tables on the host, without a device: fwd and bwd
 * receive n_steps x 2 x 2550 (m, c) float pairs each ([step][hom, het][node]),
 * stretch receives the 2550 (m, c) pairs of the n_steps-bp hom stretch.
 * Scaled rates as gsmc_context_create forms them. Replaces
 * build_flow_field_cache (src/flow_field.cpp:205-388) and load_flow_field
 * (:8-54). */
int gsmc_build_tables(const char* flow_field_path, int n_steps, double Ne, double mu, double rho,
                      int table_precision, float* fwd, float* bwd, float* stretch);

/* Number of HIP devices, select / report the current device. Replace
 * _core.get_device_count, set_device and get_device
 * (src/bindings.cpp:3630-3645). */
int gsmc_device_count(int* count);
int gsmc_set_device(int device);
int gsmc_get_device(int* device);

/* Free and total bytes of the current device.
 * Replaces _core.cuda_mem_info (src/bindings.cpp:3646-3652). */
int gsmc_mem_info(uint64_t* free_bytes, uint64_t* total_bytes);

/* Bitpack a host uint8 matrix G[n_haps][n_sites] on the device into
 * packed[n_haps][ceil(n_sites/64)]: bit b of word w = G[h][64w+b] & 1, tail
 * bits zero. Replaces _core.bitpack / bitpack_genotypes_gpu
 * (src/kernels/bitpack.cu:11-38). */
int gsmc_bitpack(const uint8_t* G, int64_t n_haps, int64_t n_sites, uint64_t* packed);

/* Inverse of gsmc_bitpack (src/kernels/bitpack.cu:43-68, _core.unpack). */
int gsmc_unpack(const uint64_t* packed, int64_t n_haps, int64_t n_sites, uint8_t* G);

/* Create a decoder context: upload and bitpack G (host uint8, row-major
 * [n_haps][n_sites]), upload the per-site gap descriptors derived from
 * positions (host float64 [n_sites]), and build or reuse the transition
 * tables for (Ne, mu, rho, cache_steps) from the flow field at
 * flow_field_path. cache_steps <= 0 means 1000. Scaled rates are
 * 4.0f*Ne*rho and 4.0f*Ne*mu in FP32 with Ne, mu, rho cast to float.
 * Contexts with equal tables share them; the tables a device used last
 * stay on it after their contexts are destroyed (about 370 MB at 1000
 * steps), as the reference's process-wide cache does, and are replaced
 * when a context needs a different set.
 * Replaces FlowContext.__init__ (src/bindings.cpp:2552-2670, binding
 * :3687-3699) and its table set-up ensure_cache (:2016-2157). */
int gsmc_context_create(const uint8_t* G, int64_t n_haps, int64_t n_sites, const double* positions,
                        double Ne, double mu, double rho, const char* flow_field_path, int cache_steps,
                        int table_precision, gsmc_context** out);

/* Same, from host rows already bitpacked as gsmc_bitpack produces them
 * (resident cohorts too large to hold as uint8). Same reference interface
 * as gsmc_context_create, with the H2D + bitpack step (bindings.cpp:2658-2665)
 * done by the caller. */
int gsmc_context_create_packed(const uint64_t* packed, int64_t n_haps, int64_t n_sites,
                               const double* positions, double Ne, double mu, double rho,
                               const char* flow_field_path, int cache_steps, int table_precision,
                               gsmc_context** out);

/* Release a context (FlowContext destructor, src/bindings.cpp:2672-2677). */
void gsmc_context_destroy(gsmc_context* ctx);

/* n_haps, n_sites, device_id properties (src/bindings.cpp:3727-3729), and
 * the FP32 Ne used to rescale posteriors; any pointer may be NULL. */
int gsmc_context_info(const gsmc_context* ctx, int64_t* n_haps, int64_t* n_sites, int* device_id, float* Ne);

/* Decoder tuning; results do not depend on any of it. checkpoint_k: sites
 * per checkpoint window, 8 or 16 (default 8), or 1: the forward kernel
 * stores its state at every core site and the backward kernel reads it
 * back instead of recomputing it (mean, code and summary launches; CI /
 * posterior launches use 16). With the static schedule that takes 8 B per
 * pair and core site of device memory, in launches split into phases that
 * fit; with the persistent schedule one fused kernel reuses a state slot
 * per resident wave. pairs_per_lane: 1, 2 or 4 (4 with checkpoint_k 8)
 * (default 1), i.e. 64, 128 or 256 pairs per 64-lane workgroup, for the
 * mean-only kernel (CI / posterior launches use 1, summaries at most 2);
 * schedule: 0 static grid
 * of work items, 1 persistent waves draining a device queue (default 0;
 * summary launches at checkpoint_k 8 with one pair per lane always use the
 * static grid, whose kernel runs more waves per SIMD); xcd_order: static grid ordered so each XCD walks a contiguous run of
 * items (default 0); min_waves: work items the host-output path puts in
 * one launch at least (default 8192). No reference counterpart (the
 * reference has fixed 256-thread launches). */
int gsmc_context_set_tuning(gsmc_context* ctx, int checkpoint_k, int pairs_per_lane, int schedule, int xcd_order,
                            int64_t min_waves_per_launch);

/* Cap on the device memory one launch may use for checkpoints / stored
 * forward states (0, the default: free device memory less 2 GiB). Launches
 * needing more run as consecutive phases of work items that each fit.
 * Results do not depend on it. No reference counterpart. */
int gsmc_context_set_state_budget(gsmc_context* ctx, uint64_t max_state_bytes);

/* Free the device workspaces the context keeps between calls (metadata,
 * checkpoints / stored states, flank states, summary partials, output
 * slots), after waiting for the context's work. A workspace keeps its
 * largest launch's buffers; this returns that memory, for instance after a
 * launch-shape sweep. The next call allocates what it needs again. No
 * reference counterpart. */
int gsmc_context_release_buffers(gsmc_context* ctx);

/* Full-sequence forward-backward of n_pairs pairs (pair_i[p], pair_j[p]) into
 * host arrays laid out site-major [n_sites][n_pairs]: mean always; lower and
 * upper when flags & GSMC_CI; alpha and beta when flags & GSMC_POSTERIOR
 * (the others may be NULL). Replaces FlowContext.run_fb
 * (src/bindings.cpp:2897-2989) and gamma_smc_flow_cached_fb_gpu
 * (src/kernels/gamma_smc_flow.cu:844-903). */
int gsmc_run_fb(gsmc_context* ctx, const int32_t* pair_i, const int32_t* pair_j, int64_t n_pairs, int flags,
                float* mean, float* lower, float* upper, float* alpha, float* beta);

/* Block windows (core_start, core_stop, padded_start, padded_stop) for
 * n_sites sites; writes min(n_blocks, cap) rows of 4 int32 and the count.
 * Replaces FlowContext::make_block_windows (src/bindings.cpp:2517-2534). */
int gsmc_block_windows(int64_t n_sites, int64_t core_block_sites, int64_t flank_sites, int32_t* blocks,
                       int64_t cap, int64_t* n_blocks);

/* Blockwise forward-backward into host arrays [n_sites][n_pairs]: each
 * padded window decoded from the prior, core rows kept. pair_batch_size > 0
 * caps the pairs per launch (-1: all). max_streams >= 1 is the number of
 * output slots, at least two: groups of blocks go to the slots in turn,
 * each slot decoding on its own stream with its own workspace and device
 * rows (launches of different slots may overlap on the device), while one
 * copy stream drains finished slots to the host through pinned staging.
 * Results do not depend on either.
 * Replaces FlowContext.run_fb_blockwise (src/bindings.cpp:2991-3345) and
 * gamma_smc_flow_cached_fb_block_gpu[_async] (gamma_smc_flow.cu:908-997). */
int gsmc_run_fb_blockwise(gsmc_context* ctx, const int32_t* pair_i, const int32_t* pair_j, int64_t n_pairs,
                          int64_t core_block_sites, int64_t flank_sites, int64_t pair_batch_size,
                          int max_streams, int flags, float* mean, float* lower, float* upper, float* alpha,
                          float* beta);

/* Per-site summary over n_pairs pairs, without per-pair outputs: for each
 * site s, site_mean[s] = (sum over pairs of the posterior mean, in FP64) /
 * n_pairs, rounded to float; site_min[s] / site_max[s] = their extremes
 * (either may be NULL). Host arrays of n_sites floats. core_block_sites
 * <= 0 decodes the full sequence (the reference's only mode). The sum runs
 * in a fixed order (a butterfly over each 64-pair group, then the groups in
 * order), so results do not depend on launch shape or pair chunking.
 * Replaces FlowContext.run_fb_summary (src/bindings.cpp:2695-2788) and
 * gamma_smc_flow_cached_fb_reduce_gpu (src/kernels/gamma_smc_flow.cu:1574-1683). */
int gsmc_run_fb_summary(gsmc_context* ctx, const int32_t* pair_i, const int32_t* pair_j, int64_t n_pairs,
                        int64_t core_block_sites, int64_t flank_sites, float* site_mean, float* site_min,
                        float* site_max);

/* Per-gene statistics of the posterior means of n_pairs pairs, without
 * per-pair outputs. Gene g covers the context's sites [gene_lo[g],
 * gene_hi[g]) (n_g sites); genes may overlap. With x = max(mean, 1) at each
 * site, pair p's gene values are lin = (sum of x) / n_g and lg = (sum of
 * ln x) / n_g (FP64 sums in site order of the FP32 means and their
 * hardware log2, times ln 2). For each gene with n_g >= 2 the call adds
 * n_pairs to count[g], the sums over pairs of lin, lg and lg^2 to lin_sum,
 * log_sum and log_sq_sum, lowers min_lin / min_log to the pairs' minima,
 * and counts each pair's lg in histogram[g][bin], bin = (number of edges
 * <= lg) - 1 clamped to [0, n_bins - 1] (edges: n_bins + 1 increasing
 * values, n_bins <= 1024). All outputs accumulate, so a caller covering
 * the pairs in several calls initialises them once (0, 0, 0, 0, +inf,
 * +inf, 0). Decodes blockwise as gsmc_run_fb_blockwise (core_block_sites
 * <= 0: the full sequence); the backward kernel sums each pair's values per
 * gene piece in place of storing rows, so no per-pair means are written.
 * Pairs go in launches of at most pair_batch (<= 0: as many as 70 % of the
 * device memory holds); results do not depend on it beyond the order of
 * the FP64 sums over pairs.
 * Replaces the per-chunk host accumulation of the reference's per-gene
 * pipeline (analysis/genome_wide/infer_chromosome.py:157-227), which reads
 * each chunk's (sites, pairs) means back from infer_blockwise first. */
int gsmc_run_fb_gene_stats(gsmc_context* ctx, const int32_t* pair_i, const int32_t* pair_j, int64_t n_pairs,
                           int64_t core_block_sites, int64_t flank_sites, int64_t pair_batch, int64_t n_genes,
                           const int64_t* gene_lo,
                           const int64_t* gene_hi, int n_bins, const double* edges, int64_t* count,
                           double* lin_sum, double* log_sum, double* log_sq_sum, double* min_lin, double* min_log,
                           int64_t* histogram);

/* Forward-only filtering of n_pairs pairs over the full sequence: the
 * filter posterior mean at each site given the sites up to it, [n_sites]
 * [n_pairs] host arrays (lower / upper with GSMC_CI). Same recursion as the
 * reference's half-precision forward-only path: FP16 composed transition
 * tables, gap = (int)(delta + 0.5) bp, explicit het emission. Replaces
 * FlowContext.run_fwd (src/bindings.cpp:2794-2849) and
 * gamma_smc_flow_h2_fwd_gpu (src/kernels/gamma_smc_flow.cu:1139-1265). */
int gsmc_run_fwd(gsmc_context* ctx, const int32_t* pair_i, const int32_t* pair_j, int64_t n_pairs, int flags,
                 float* mean, float* lower, float* upper);

/* Forward states of the cached forward-backward walk (full sequence, FP32
 * tables): after each site, the state's log10 mean (scaled time) into
 * mean_log10 and its log10 CV into cv_log10, host arrays [n_sites][n_pairs].
 * Replaces FlowContext.run_fwd_states (src/bindings.cpp:2851-2893, binding
 * :3704) and gamma_smc_flow_cached_forward_states_gpu
 * (src/kernels/gamma_smc_flow.cu:1542-1566). */
int gsmc_run_fwd_states(gsmc_context* ctx, const int32_t* pair_i, const int32_t* pair_j, int64_t n_pairs,
                        float* mean_log10, float* cv_log10);

/* Device-resident form of gsmc_run_fwd: device pair ids and outputs,
 * queued on `stream` without synchronising. Same reference interface. */
int gsmc_run_fwd_device(gsmc_context* ctx, const int32_t* d_pair_i, const int32_t* d_pair_j, int64_t n_pairs,
                        int flags, float* d_mean, float* d_lower, float* d_upper, void* stream);

/* Device-resident form of gsmc_run_fb_blockwise: pair ids and outputs are
 * device pointers, outputs [n_sites][n_pairs]; the work is queued on
 * `stream` (a hipStream_t, NULL = default stream) and the call returns
 * without synchronising. core_block_sites >= n_sites with flank 0 is the
 * full-sequence run_fb. No allocation happens on the stream's critical path
 * beyond the checkpoint buffer, which the context keeps between calls.
 * Same reference interface as gsmc_run_fb_blockwise, minus the D2H. */
int gsmc_run_fb_device(gsmc_context* ctx, const int32_t* d_pair_i, const int32_t* d_pair_j, int64_t n_pairs,
                       int64_t core_block_sites, int64_t flank_sites, int flags, float* d_mean, float* d_lower,
                       float* d_upper, float* d_alpha, float* d_beta, void* stream);

/* gsmc_run_fb_device without any device allocation: the launch metadata,
 * flank states and checkpoints of the call live in a fixed device array of
 * 64 MiB (GSMC_NOALLOC_SCRATCH_BYTES) that each device holds for the life
 * of the process; a small kernel writes the metadata, so the call makes no
 * allocation, no host-to-device copy and no memset, only kernel launches
 * on `stream`. Same arguments and the same results as gsmc_run_fb_device
 * (checkpoints every 8 sites, static grid). A call needing more scratch
 * than the maximum -- about n_blocks * ceil(n_pairs / 64) * (core / 8 + 2)
 * * 512 bytes -- fails with GSMC_EINVAL before anything is queued. Every
 * call on a device shares that device's array: calls whose work overlaps
 * on one device (other streams, other threads, other contexts) race. No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). */
#define GSMC_NOALLOC_SCRATCH_BYTES (64ull << 20)
int gsmc_run_fb_device_noalloc(gsmc_context* ctx, const int32_t* d_pair_i, const int32_t* d_pair_j, int64_t n_pairs,
                               int64_t core_block_sites, int64_t flank_sites, int flags, float* d_mean,
                               float* d_lower, float* d_upper, float* d_alpha, float* d_beta, void* stream);

/* Launch workspaces for batched ragged jobs and the serving loop: a batch
 * owns the metadata, checkpoint and queue buffers of one launch in flight.
 * gsmc_batch_launch queues every job of the list as one kernel launch on
 * `stream` (work items = 64-pair tiles x block windows of all jobs) and
 * returns; a later launch on the same batch first waits for this one.
 * gsmc_batch_query reports completion without blocking, gsmc_batch_wait
 * blocks. No reference counterpart (the reference decodes one request per
 * launch sequence). */
int gsmc_batch_create(gsmc_context* ctx, gsmc_batch** out);
void gsmc_batch_destroy(gsmc_batch* batch);
int gsmc_batch_launch(gsmc_batch* batch, const gsmc_job* jobs, int64_t n_jobs, int flags, void* stream);
int gsmc_batch_wait(gsmc_batch* batch);
int gsmc_batch_query(gsmc_batch* batch, int* done);
/* Capture a batched launch of `jobs` into an executable HIP graph with its
 * own launch buffers: the plan and its metadata upload happen once, here,
 * and each gsmc_graph_launch replays only the launch's kernels (queue reset,
 * decode, summary epilogue; no copies or memsets) on `stream`. The jobs'
 * pair ids and outputs are read and written at replay time through the
 * pointers given now. Replays of one graph must not overlap (they share its
 * buffers). No reference counterpart (the reference launches every block
 * from the host). */
int gsmc_batch_capture(gsmc_context* ctx, const gsmc_job* jobs, int64_t n_jobs, int flags, gsmc_graph** out);
int gsmc_graph_launch(gsmc_graph* graph, void* stream);
void gsmc_graph_destroy(gsmc_graph* graph);

/* Grow a batch's buffers now (checkpoint bytes, work items and jobs of the
 * largest launch the caller will make), so launches within these bounds
 * never allocate or synchronise the device. A reserved batch never grows its
 * checkpoint buffer: a launch needing more runs in phases (static schedule)
 * or with fewer resident waves (persistent stored-state kernel). */
int gsmc_batch_reserve(gsmc_batch* batch, uint64_t checkpoint_bytes, int64_t max_items, int64_t max_jobs);

/* Bytes the batch's checkpoint / stored-state buffer holds now. After
 * gsmc_batch_reserve, launches keep within the reservation (more phases or
 * fewer resident waves), so this stays at the reserved size. No reference
 * counterpart. */
int gsmc_batch_buffer_bytes(gsmc_batch* batch, uint64_t* checkpoint_bytes);

/* Device bytes of checkpoints one job of n_pairs pairs over n_sites sites
 * needs under the context's tuning (for admission control). */
int gsmc_checkpoint_bytes(gsmc_context* ctx, int64_t n_sites, int64_t core_block_sites, int64_t flank_sites,
                          int64_t n_pairs, uint64_t* bytes);

/* Device memory, stream and event plumbing for callers without a GPU
 * framework (bench, tests). No reference counterpart. */
int gsmc_malloc(void** ptr, uint64_t bytes);
int gsmc_free(void* ptr);
int gsmc_memcpy_htod(void* dst, const void* src, uint64_t bytes);
int gsmc_memcpy_dtoh(void* dst, const void* src, uint64_t bytes);
/* Copies queued on `stream`; the host side must be pinned (gsmc_host_alloc)
 * for the copy to overlap device work. */
int gsmc_memcpy_htod_async(void* dst, const void* src, uint64_t bytes, void* stream);
int gsmc_memcpy_dtoh_async(void* dst, const void* src, uint64_t bytes, void* stream);
/* Pinned host memory. */
int gsmc_host_alloc(void** ptr, uint64_t bytes);
int gsmc_host_free(void* ptr);
int gsmc_memset(void* dst, int value, uint64_t bytes);
int gsmc_synchronize(void);
int gsmc_stream_create(void** stream);
int gsmc_stream_destroy(void* stream);
int gsmc_stream_synchronize(void* stream);
int gsmc_event_create(void** event);
int gsmc_event_destroy(void* event);
int gsmc_event_record(void* event, void* stream);
int gsmc_event_query(void* event, int* done); /* done = 1 once the event completed, without blocking */
int gsmc_event_elapsed_ms(void* start, void* stop, float* ms);
/* Gather-only probe of the texture path: `waves` 64-lane waves each run two
 * independent chains of `steps` table lookups on the context's forward
 * tables (one layer per step shared by the wave, het slab in ~15% of lanes,
 * nothing else in the loop); *ms is the timed launch's duration, so
 * 2 * 64 * waves * steps / ms is the lookup rate the device sustains. out
 * is device memory for 64 * waves floats. No reference counterpart
 * (measurement). */
int gsmc_lookup_probe(gsmc_context* ctx, int64_t waves, int32_t steps, void* out, void* stream, float* ms);

/* Streaming device-to-device copy kernel (bytes a multiple of 16), queued
 * on `stream`; a kernel of known traffic for counter calibration. */
int gsmc_device_copy(void* dst, const void* src, uint64_t bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* GAMMA_SMC_AMD_H */
