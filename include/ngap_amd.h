/* ngap_amd — MI355X (gfx950) non-blocking automata processing, C ABI.
 *
 * Plain pointers and sizes only.  Every function returns an int status
 * (NGAP_OK = 0) and leaves a message for ngap_last_error() on failure; no
 * C++ exception crosses the ABI (a host allocation or thread failure inside a
 * call returns NGAP_ERR_HOST).
 * One engine may be used by one host thread at a time.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   ngap_frontend_*   load_nfa_from_file + ngap::preprocessing +
 *                     nfa_utils::group_nfas_by_num + Graph::ReadNFA
 *                     (code/src/ngap/main.cu:56-78, code/src/ngap/ngap.cu:63-90,
 *                      code/src/commons/nfa_utils.cpp:382-424,
 *                      code/src/commons/graph.cu:169-206)
 *   ngap_load_groups  initGroupCsrWithPrec + GroupMatchset/GroupNodeAttrs/
 *                     GroupAAS::init (code/src/ngap/ngap.cu:3086-3111,
 *                     code/include/commons/group_graph.h:35-109)
 *   ngap_set_input    concat_input_streams_to_array2 + copy_to_device
 *                     (code/src/gpunfautils/abstract_gpunfa.cu:55-81,
 *                      code/src/ngap/ngap.cu:1482-1483)
 *   ngap_run          NonBlockingBuffer::init_nfagroups + the kernel launch in
 *                     ngap::launch_non_blocking_*_groups (code/src/ngap/
 *                     ngap_buffer.cu:475-748, code/src/ngap/ngap.cu:284-1990;
 *                     kernels code/src/ngap/kernel.h:9-55)
 *   ngap_get_reports  D2H + pack + sort by (offset, id) + unique
 *                     (code/src/ngap/ngap.cu:1899-1933, compareResult :19-30)
 */
#ifndef NGAP_AMD_H_
#define NGAP_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define NGAP_OK 0
#define NGAP_ERR_INVALID 1     /* bad argument or malformed input */
#define NGAP_ERR_DEVICE 2      /* HIP runtime error / no device */
#define NGAP_ERR_OOM 3         /* device allocation failed */
#define NGAP_ERR_OVERFLOW 4    /* worklist or report capacity exceeded */
#define NGAP_ERR_STATE 5       /* call out of order (no groups / no input) */
#define NGAP_ERR_HOST 6        /* host memory or threads exhausted */

/* --algorithm names (code/src/ngap/main.cu:114-130) */
#define NGAP_ALGO_BLOCKING 0        /* blockinggroups */
#define NGAP_ALGO_NAP 1             /* NAPgroups */
#define NGAP_ALGO_NONBLOCKING 2     /* nonblockinggroups (O1) */
#define NGAP_ALGO_R1 3              /* nonblockingr1groups (O4, one level) */
#define NGAP_ALGO_R2 4              /* nonblockingr2groups (O4, two levels) */
#define NGAP_ALGO_PC 5              /* nonblockingpcgroups (O1+O2) */
#define NGAP_ALGO_ALL 6             /* nonblockingallgroups (O1+O2+O3) */

/* scheduling of (group, chunk) work */
#define NGAP_SCHED_PLAIN 0     /* one workgroup per (group, chunk) unit */
#define NGAP_SCHED_BATCHED 1   /* persistent waves over a device work queue */

typedef struct ngap_engine ngap_engine;

/* One batched work item: roots at positions [begin, end) of chunk `chunk`
 * (offsets relative to the chunk start) for group `group`, or for every group
 * when group is NGAP_ALL_GROUPS (precompute depth >= 2); its items run to the
 * end of the chunk.  The work items of one (group, chunk) unit must tile
 * [0, chunk_len) exactly; a chunk with an all-group work item takes only
 * all-group work items. */
#define NGAP_ALL_GROUPS 0xFFFFFFFFu
typedef struct {
  uint32_t group, chunk, begin, end;
} ngap_task;

/* One group as Graph/Csr hold it (code/include/commons/graph.h:101-248). */
typedef struct {
  int32_t num_states;
  const int32_t *row_offsets;   /* [num_states+1], edges sorted by (src,dst) */
  const int32_t *col_indices;   /* [row_offsets[num_states]] */
  const uint32_t *symbol_sets;  /* [num_states*8], bit b in word b/32 */
  const uint8_t *node_attrs;    /* bit0 all-input start, bit1 reporting */
  int32_t num_always_active;
  const int32_t *always_active; /* all-input states */
  int32_t num_start;
  const int32_t *start_states;  /* start-of-data states (not all-input) */
} ngap_group_desc;

/* Per-run options.  Names follow the reference flags
 * (code/src/ngap/ngap_option.h:12-68, code/include/commons/common_func.h:50-115). */
typedef struct {
  int32_t algorithm;            /* NGAP_ALGO_* */
  int32_t schedule;             /* NGAP_SCHED_* */
  int32_t precompute_depth;     /* --precompute-depth, 0..3 */
  int32_t active_threshold;     /* --active-threshold: a lane continues with a
                                   matched successor in-lane only while more
                                   lanes than this are busy (else it queues it) */
  int32_t unique;               /* --unique: de-duplicate (state, offset) items */
  int32_t is_dup;               /* start-of-data states in every chunk */
  int32_t report_off;           /* --report-off: count reports only */
  uint32_t data_buffer_fetch_size; /* --data-buffer-fetch-size: accepted, no
                                      effect (no fetch buffer in this engine) */
  uint32_t add_aas_start;       /* --add-aan-start: accepted, no effect */
  uint32_t add_aas_interval;    /* --add-aas-interval: accepted, no effect (all-
                                   input roots come a wave's width of positions
                                   per step; NAPgroups one position per step) */
  uint64_t result_capacity;     /* --result-capacity (reports) */
  uint32_t window;              /* batched: positions per work item, 0 = auto */
  uint32_t spill_capacity;      /* items per wave beyond the LDS ring, 0 = auto */
  uint32_t persistent_blocks;   /* batched: workgroups, 0 = auto (the device's
                                   occupancy, fewer for a job of fewer than
                                   640 positions per wave) */
  int32_t use_graph;            /* batched: replay a captured hipGraph */
  const ngap_task *tasks;       /* batched: work items in run order (NULL = the
                                   engine's static schedule) */
  uint32_t num_tasks;
  int32_t fallback;             /* re-run overflowing work items with a larger
                                   worklist instead of failing */
  int32_t cross_groups;         /* batched static schedule: 1 (or 2) = work
                                   items span all groups when the precompute
                                   depth is >= 2 (one group too: its roots
                                   then come from the 3-gram slots), 0 = one
                                   group per work item */
  uint32_t ring_capacity;       /* batched: LDS worklist items per wave, a power
                                   of two in [64, 2048], 0 = 256 */
  int32_t task_traces;          /* batched: record the items of every work item
                                   (ngap_get_task_items); a wave then finishes
                                   one work item before it starts the next */
  int32_t root_pass;            /* batched engine runs: 1 = compute every work
                                   item's roots in a pass of their own (many
                                   short waves) before the persistent kernel
                                   expands them, 0 = roots inside the
                                   persistent kernel */
  uint32_t root_capacity;       /* root pass: root items reserved per input
                                   position of a work item, 0 = 4; a work item
                                   with more is re-run by the fallback */
  uint64_t tasks_id;            /* batched, with tasks: the tasks_id a run over
                                   these same work items returned in its stats.
                                   The caller promises they are unchanged since
                                   that run, and the engine skips comparing
                                   them with its copy; 0 = compare */
} ngap_run_opts;

typedef struct {
  double kernel_ms;             /* device time of the matching launch(es) */
  uint64_t reports_raw;         /* reports written (before sort/unique) */
  uint64_t overflow_units;      /* units/work items whose worklist overflowed */
  uint64_t items;               /* (state, offset) items expanded */
  uint32_t num_units;           /* (group, chunk) pairs */
  uint32_t num_tasks;           /* batched work items */
  uint32_t fallback_tasks;      /* work items re-run after an overflow */
  double fallback_ms;           /* device time of those re-runs */
  uint64_t tasks_id;            /* batched: id of the work items this run used
                                   (for opts.tasks_id while they are unchanged;
                                   loading groups or input retires it; ids are
                                   unique across the engines of a process) */
} ngap_run_stats;

const char *ngap_last_error(void);
void ngap_default_opts(ngap_run_opts *opts);

int ngap_create(int32_t device, ngap_engine **out);
int ngap_destroy(ngap_engine *eng);

/* Builds the device automata image (memo tables up to max_precompute_depth)
 * on the device (HIP kernels over the uploaded CSR arrays). */
int ngap_load_groups(ngap_engine *eng, const ngap_group_desc *groups,
                     int32_t num_groups, int32_t max_precompute_depth);

/* The same with the builder chosen: NGAP_BUILD_DEVICE (what ngap_load_groups
 * does) or NGAP_BUILD_HOST (host threads, then one upload per array).  Both
 * give the same arrays byte for byte.  The reference builds its memo tables
 * on the host (code/src/ngap/ngap.cu:2358-2513,
 * code/src/commons/precompute_table.cu:25-395). */
#define NGAP_BUILD_DEVICE 0
#define NGAP_BUILD_HOST 1
int ngap_load_groups_ex(ngap_engine *eng, const ngap_group_desc *groups, int32_t num_groups,
                        int32_t max_precompute_depth, int32_t build);

/* The loaded image: where it was built, its size and the wall time of the
 * load (build and upload). */
typedef struct {
  int32_t built_on_device;
  int32_t memo_depth;
  int32_t num_groups;
  uint64_t num_states;
  uint64_t bytes;
  double build_ms;
} ngap_image_summary;
int ngap_image_info(ngap_engine *eng, ngap_image_summary *info);

/* Copies the device image array `name` (meta, succ4, cls, sod, beg1-3,
 * nrep1-3, list1-3, sel2, sel3, mask1, xr1, xe1, xmask1, xr2, xe2, xsod, xs3,
 * xp3, x2rep; layout in ngap_amd/csrc/layout.h) into out (at most cap bytes);
 * *n_bytes gets its size (out may be NULL to query it). */
int ngap_image_array(ngap_engine *eng, const char *name, void *out, uint64_t cap,
                     uint64_t *n_bytes);

/* Uploads num_seg chunks of chunk_len bytes (the concatenated input) and
 * returns when they are on the device (data may be reused at once).  The
 * bytes go through two 8 MiB pinned blocks of the engine, copied in by up to
 * 8 host threads while the other block's copy to the device runs. */
int ngap_set_input(ngap_engine *eng, const uint8_t *data, uint64_t chunk_len,
                   uint32_t num_seg);

int ngap_run(ngap_engine *eng, const ngap_run_opts *opts, ngap_run_stats *stats);

/* ngap_run without device allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The input
 * (num_seg chunks of chunk_len bytes) and the report buffer (report_cap
 * entries) are the caller's device memory; the run's counters, work-item
 * flags and worklist spill areas live in a fixed 48 MiB device array of the
 * library (one per device, held until the process ends), and its work items
 * are generated on the device (the static schedule's shape: all-group
 * windows at depth >= 2, else one group per work item;
 * opts->window or about two work items per wave).  Runs the batched
 * streaming engine on the engine's stream and waits for it.  Reports are
 * written raw in the device encoding ((offset << 32) | (state | group << 22)),
 * unordered and possibly repeated; sorting, removing repeats and swapping
 * the halves gives ngap_get_reports' output.  stats->reports_raw is their
 * count.  Errors: more than 1M work items (use a larger window), a grid
 * whose worklists do not fit the array, NGAP_ERR_OVERFLOW when a worklist
 * (no fallback here) or report_cap overflows.  Calls that overlap on one
 * device race on the shared array. */
int ngap_run_noalloc(ngap_engine *eng, const ngap_run_opts *opts, const uint8_t *d_input,
                     uint64_t chunk_len, uint32_t num_seg, uint64_t *d_reports,
                     uint64_t report_cap, ngap_run_stats *stats);

/* Autotuner of the batched engine (the reference picks these knobs by hand
 * per application, code/scripts/configs/exec_config_ngap_groups_best): times
 * candidate work-item windows, LDS worklist capacities, persistent grid sizes,
 * worklist spill capacities, the root pass (on or off) and the active
 * threshold (the reference's tuning values 0, 8, 16, 24, 32, ngap.cu:1999) on
 * the loaded input by coordinate descent (at most max_runs timed runs; 23
 * cover every axis) and writes the fastest into *best: a copy of *base,
 * batched, with fallback on, window / ring_capacity / persistent_blocks /
 * spill_capacity / root_pass / active_threshold set.  Report
 * sets do not depend on these knobs; the chunk length is part of the input's
 * meaning (matches do not cross chunks) and is not tuned. */
int ngap_autotune(ngap_engine *eng, const ngap_run_opts *base, int32_t max_runs,
                  ngap_run_opts *best, double *best_ms);

/* The batched engine's persistent grid for a ring capacity (0 = default):
 * workgroups and waves per workgroup, for callers that shape their own work
 * items (about two per wave keep every wave busy).  No counterpart in the
 * reference (its grid is (groups, chunks)). */
int ngap_batched_grid(ngap_engine *eng, uint32_t ring_capacity, uint32_t *blocks,
                      uint32_t *waves_per_block);

/* Sorted, de-duplicated reports of the last run, reference encoding
 * ((state | group << 22) << 32) | offset.  *n_out gets the total count; at
 * most cap entries are written (out may be NULL to query the count). */
int ngap_get_reports(ngap_engine *eng, uint64_t *out, uint64_t cap, uint64_t *n_out);

/* Per-work-item item counts of the last batched run (frontier traces), in
 * the order of the work items; the run must have set task_traces. */
int ngap_get_task_items(ngap_engine *eng, uint64_t *out, uint64_t cap, uint64_t *n_out);
/* The work items of the last batched run (the static schedule's choice when
 * none were given). */
int ngap_get_tasks(ngap_engine *eng, ngap_task *out, uint64_t cap, uint64_t *n_out);

/* ---- serving: continuous batching of scan requests on resident groups ----
 * No counterpart in the reference (one input per process there).  A request
 * (group, bytes) is scanned as one chunk.  Requests queue; whenever one of
 * max_inflight launch slots is free and the HBM budget allows, all queued
 * requests that fit are packed into one launch on that slot's stream. */
typedef struct ngap_server ngap_server;

/* opts: algorithm / precompute depth / unique as for ngap_run (report_off is
 * rejected: requests return their reports).  budget_bytes 0 = 200 GB.  A
 * request that alone needs more than the budget fails (state 3). */
int ngap_server_create(ngap_engine *eng, uint64_t budget_bytes, int32_t max_inflight,
                       const ngap_run_opts *opts, ngap_server **out);
int ngap_server_destroy(ngap_server *srv);
int ngap_server_submit(ngap_server *srv, int32_t group, const uint8_t *data, uint64_t len,
                       uint64_t *req_id);
/* Retires finished launches and starts new ones; block != 0 waits for at
 * least one launch to finish when none has. */
int ngap_server_poll(ngap_server *srv, int32_t block);
/* state: 0 queued, 1 in flight, 2 done, 3 failed.  When done, the request's
 * sorted unique reports (offsets relative to the request) are copied to out;
 * a request is forgotten once all of them were delivered.  latency_ms is
 * submit-to-completion wall time. */
int ngap_server_take(ngap_server *srv, uint64_t req_id, int32_t *state, uint64_t *out,
                     uint64_t cap, uint64_t *n_out, double *latency_ms);
int ngap_server_pending(ngap_server *srv, uint64_t *queued, uint64_t *inflight);

/* Memory and admission figures of a server.  held_bytes counts the device
 * buffers of every launch slot (input, last-byte map, reports, work items,
 * worklist spill), in flight or kept for reuse; admission keeps it within
 * budget_bytes (idle slots give their buffers back when a launch needs the
 * room).  A launch whose worklists or report buffer overflowed is re-run with
 * larger ones (redone), which may go past the budget for that re-run.  The
 * resident automata image is not part of the budget. */
typedef struct {
  uint64_t budget_bytes;
  uint64_t queued, inflight_requests;
  uint64_t inflight_bytes, held_bytes;
  uint64_t peak_inflight_bytes, peak_held_bytes;
  uint64_t launches;      /* launches started */
  uint64_t rejected;      /* requests failed: larger than the budget alone */
  uint64_t redone;        /* launches re-run after an overflow */
} ngap_server_info;
int ngap_server_stats(ngap_server *srv, ngap_server_info *info);

/* ---- learned schedule: frontier-size predictor (no reference counterpart) --
 * A 264 -> 64 -> 64 -> 1 ReLU MLP in FP32 predicting log(1 + items) of a
 * (group, 4 KiB window) work item from the window's normalised 256-bin byte
 * histogram and 8 statistics of the group.  Parameters: W1[264][64], b1[64],
 * W2[64][64], b2[64], W3[64], b3. */
typedef struct ngap_mlp ngap_mlp;
int ngap_mlp_create(int32_t device, uint64_t seed, ngap_mlp **out);
int ngap_mlp_destroy(ngap_mlp *m);
int32_t ngap_mlp_num_params(void);
int ngap_mlp_get_params(ngap_mlp *m, float *out);
int ngap_mlp_set_params(ngap_mlp *m, const float *in);
/* y[n] for dense features x[n][264]. */
int ngap_mlp_predict(ngap_mlp *m, const float *x, uint64_t n, float *y);
/* y[n] for pairs[n] = (group, window): features hist[window][256] ++ stats[group][8]. */
int ngap_mlp_predict_pairs(ngap_mlp *m, const float *hist, uint32_t num_windows,
                           const float *stats, uint32_t num_groups, const uint32_t *pairs,
                           uint64_t n, float *y);
/* Gradient of the mean squared error over (x, t) into grad[num_params]. */
int ngap_mlp_gradient(ngap_mlp *m, const float *x, const float *t, uint64_t n, float *grad,
                      float *loss);
/* Adam, mini-batches of `batch`, `epochs` passes; final epoch mean loss. */
int ngap_mlp_train(ngap_mlp *m, const float *x, const float *t, uint64_t n, int32_t epochs,
                   uint32_t batch, float lr, uint64_t seed, float *final_loss);

/* ---- front-end: ANML -> groups in the reference state-id namespace ---- */
typedef struct ngap_frontend ngap_frontend;

/* group_num: --group-num (lowered to the component count when larger),
 * max_nfa_size: --max-nfa-size (-1 off), quit_degree / remove_degree as the
 * reference flags. */
int ngap_frontend_load(const char *anml_path, int32_t group_num, int32_t max_nfa_size,
                       int32_t quit_degree, int32_t remove_degree, ngap_frontend **out);
int ngap_frontend_load_text(const char *anml_text, int32_t group_num,
                            int32_t max_nfa_size, ngap_frontend **out);
int ngap_frontend_free(ngap_frontend *fe);
int32_t ngap_frontend_num_groups(const ngap_frontend *fe);
/* Fills desc with pointers owned by fe (valid until ngap_frontend_free). */
int ngap_frontend_group(const ngap_frontend *fe, int32_t g, ngap_group_desc *desc);
/* Name of state s of group g (ANML id, with _dup_ suffixes after splits). */
const char *ngap_frontend_state_name(const ngap_frontend *fe, int32_t g, int32_t s);
/* Text dump of all groups (same format as the reference-side dumper); NULL
 * when the text could not be built (ngap_last_error() says why). */
const char *ngap_frontend_dump(ngap_frontend *fe);
/* Parses one VASim symbol-set string into 8 words. */
int ngap_parse_symbol_set(const char *s, uint32_t words_out[8]);

#ifdef __cplusplus
}
#endif

#endif /* NGAP_AMD_H_ */
