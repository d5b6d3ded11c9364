/*
 * grass.h -- C ABI of the MI355X GraSS sampler (libgrass_amd.so).
 *
 * The reference has no C ABI: its boundary is the .graph/.offset files, the rw/ks/mrw apps
 * (`-g <cgc_path> [-n <run_times>]`) and compile-time App templates.  Each entry point below
 * names the reference interface it replaces.  Conventions:
 *   - every function returns a grass_status (0 = GRASS_OK) and never throws;
 *   - pointers named *_dev are device (HBM) pointers owned by the caller unless stated;
 *   - `stream` is a hipStream_t (NULL = the graph's own stream); a call only enqueues work on
 *     it unless it says it synchronises;
 *   - a graph handle may be used from several host threads; calls on one stream are ordered.
 */
#ifndef GRASS_H
#define GRASS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GRASS_OK = 0,
    GRASS_EINVAL = 1,     /* bad argument (null handle, size overflow, vertex id >= n, ...) */
    GRASS_EIO = 2,        /* a .graph/.offset file is missing or unreadable */
    GRASS_EFORMAT = 3,    /* malformed compressed graph */
    GRASS_ENOMEM = 4,     /* host or device allocation failed */
    GRASS_EDEVICE = 5,    /* HIP runtime error */
    GRASS_ECAPACITY = 6,  /* an output buffer or the admission budget is too small */
    GRASS_ECANCELLED = 7, /* request cancelled */
    GRASS_ETIMEOUT = 8,   /* wait timed out; the request is still running */
    GRASS_ENOTREADY = 9,  /* poll: the request has not finished */
} grass_status;

typedef enum { GRASS_FMT_LE = 0, GRASS_FMT_GT = 1 } grass_format;

typedef struct grass_graph grass_graph;

typedef struct {
    uint64_t n;            /* vertices */
    uint64_t m;            /* adjacency entries (sum of degrees) */
    uint64_t max_degree;
    uint64_t graph_bytes;  /* size of <p>.graph */
    uint64_t offset_bytes; /* size of <p>.offset */
    uint64_t total_bits;   /* sum of record lengths */
    uint32_t ubl;          /* width of one record-length field in .offset */
    int32_t format;        /* grass_format */
    int32_t device;
    uint64_t device_bytes; /* HBM held by the resident graph */
    uint32_t gt_checkpoint; /* GT: in-block interval of the checkpoints (absolute ids) that the
                               device payload stream built at load stores in place of gaps (the
                               environment variable GRASS_GT_CHECKPOINT, a power of two in
                               [2, 1024], overrides the default 4); 0 for LE */
} grass_graph_info;

/* Text of a status code. */
const char *grass_strerror(int status);

/* Loads <prefix>.graph and <prefix>.offset onto `device` (HIP ordinal): pinned-host staged
 * copies, device byte swap, record-length extraction + scan, full header validation.
 * Replaces loadGraph (LE/src/grass.cu:1015-1095), transferToGPUs (LE/src/grass.cu:1097-1113)
 * and copyPartitionToGPU with decode_offset + cub scan (LE/src/utils.hpp:328-353,
 * LE/src/bcg.cuh:284-329).  The reference returns nullptr on a missing file (grass.cu:1023);
 * this returns GRASS_EIO.  Synchronises. */
int grass_open(const char *prefix, int format, int device, grass_graph **out);

/* Same as grass_open from in-memory file images (host memory). */
int grass_open_memory(const void *graph, uint64_t graph_bytes, const void *offset,
                      uint64_t offset_bytes, int format, int device, grass_graph **out);

/* Same as grass_open_memory from file images already in device memory of `device` (copied on the
 * device, no host staging).  No reference counterpart: the reference reads files only. */
int grass_open_device_images(const void *graph_dev, uint64_t graph_bytes, const void *offset_dev,
                             uint64_t offset_bytes, int format, int device, grass_graph **out);

/* Encodes a CSR in device memory with the LE or GT compressor (as grass_compress_device in
 * grass_ingest.h) and loads the images it made as grass_open would load their files, without
 * either leaving the GPU.  No reference counterpart; for a graph built on the GPU. */
int grass_compress_device_open(const uint64_t *rowptr_dev, const uint32_t *col_dev, uint64_t n,
                               int format, int device, void *stream, grass_graph **out);

/* Frees every device and host resource of the graph.  NULL is accepted. */
int grass_close(grass_graph *g);

int grass_info(const grass_graph *g, grass_graph_info *info);

/* Device views of the resident graph: the payload words the samplers decode (LE: the .graph
 * image as big-endian words; GT: the payload stream built at load, which replaces the image),
 * and the degrees (n).  *starts_dev is set to NULL: record starts are used at load only (the
 * header cache holds each record's payload position).  Valid until grass_close. */
int grass_device_arrays(const grass_graph *g, const uint32_t **words_dev,
                        const uint64_t **starts_dev, const uint32_t **degree_dev);

/* Device memory and copies for callers without a HIP runtime of their own (FFI bindings, the
 * rw/ks/mrw command-line tools).  The reference allocates everything itself inside nextdoor
 * (allocNextDoorDataOnGPU, LE/src/grass.cu:1129-1270) and never returns samples.  The copies
 * are synchronous.  grass_device_free(g, NULL) is a no-op. */
int grass_device_alloc(grass_graph *g, uint64_t bytes, void **ptr_dev);
int grass_device_free(grass_graph *g, void *ptr_dev);
int grass_copy_to_host(grass_graph *g, void *host, const void *src_dev, uint64_t bytes);
int grass_copy_to_device(grass_graph *g, void *dst_dev, const void *host, uint64_t bytes);

/* Waits for `stream` and returns the first error a kernel of this graph recorded since the
 * last call (a decoded neighbour id >= n, a start vertex >= n), then clears it. */
int grass_sync(grass_graph *g, void *stream);

/* Full decode to CSR: rowptr_dev[n+1] (uint64), col_dev[m] (uint32); col_dev == NULL writes
 * the row pointers only.  Replaces running
 * BCGVertex::get_vertex (LE/src/bcg.cuh:131-261, GT/src/bcg.cuh:110-168) over every k.  The
 * first call with col_dev builds per-graph tables it keeps until grass_close: each decode wave's
 * first vertex (4 bytes per wave) and, for GT with a checkpoint interval of at most 4, the runs
 * before each vertex (8 (n + 1) bytes, with one host synchronisation). */
int grass_decode_csr(grass_graph *g, uint64_t *rowptr_dev, uint32_t *col_dev, void *stream);

/* Uniform random walks, one launch: out_dev is row-major [walkers][length+1] uint32; a walk
 * that reaches a vertex without neighbours is padded with n (INVALID_VERTEX, grass.cu:1182).
 * starts_dev == NULL draws start w from Philox(seed, {instance, w, 0xFFFFFFFF, 0}); step s of
 * walker w draws from Philox(seed, {instance, w, s, 0}).  Replaces DeepWalkApp::next
 * (LE/src/apps/randomwalk/randomWalks.cu:62-79) driven by doTransitParallelSampling
 * (LE/src/grass.cu:1299-1900). */
int grass_random_walk(grass_graph *g, const uint32_t *starts_dev, uint64_t walkers,
                      uint32_t length, uint64_t seed, uint32_t instance, uint32_t *out_dev,
                      void *stream);

/* One random-walk request of a batch. */
typedef struct {
    uint64_t walkers;
    uint32_t length;
    uint32_t instance;
    const uint32_t *starts_dev; /* NULL: Philox-drawn starts */
    uint32_t *out_dev;          /* [walkers][length+1] */
} grass_rw_request;

/* Many random-walk requests of different sizes in one persistent launch with a device work
 * queue; each request's output is identical to grass_random_walk on it alone. */
int grass_random_walk_batch(grass_graph *g, const grass_rw_request *reqs, uint32_t count,
                            uint64_t seed, void *stream);

/* grass_random_walk_batch that allocates nothing: no hipMalloc, hipFree or host staging buffer on
 * any call.  No counterpart in the reference; it is for callers that cannot allocate (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation).  The batch's work
 * queue, a copy of `reqs` and its descriptors live in one fixed device array of the library per
 * GPU (about 590 KB, held until the process ends), so count <= GRASS_RW_NOALLOC_MAX_REQUESTS
 * (GRASS_ECAPACITY beyond).  Requests run in their given order (grass_random_walk_batch puts the
 * longest walks first); every output is identical to grass_random_walk_batch's.  Enqueues on
 * `stream` and returns without waiting: keep `reqs` unchanged until the stream has passed the
 * call.  Every call on one GPU shares that array, so calls that overlap on one GPU race: order
 * them on one stream.  Errors a kernel records are reported by grass_sync. */
#define GRASS_RW_NOALLOC_MAX_REQUESTS 8192
int grass_random_walk_batch_noalloc(grass_graph *g, const grass_rw_request *reqs, uint32_t count,
                                    uint64_t seed, void *stream);

/* Measurement companions (no reference counterpart; the reference only prints timers,
 * LE/src/grass.cu:1865-1866,1984-1987): run the same walks without writing them and return
 * their algorithmic HBM bytes -- per walker 4 (start written, +4 if read from starts_dev), per
 * visited vertex 8 (record start) + ceil(header bits/8), per step ceil(payload bits read/8) +
 * 4 (id written).  out_dev of the requests is ignored.  Synchronise. */
int grass_random_walk_bytes(grass_graph *g, const uint32_t *starts_dev, uint64_t walkers,
                            uint32_t length, uint64_t seed, uint32_t instance, void *stream,
                            uint64_t *bytes);
int grass_random_walk_batch_bytes(grass_graph *g, const grass_rw_request *reqs, uint32_t count,
                                  uint64_t seed, void *stream, uint64_t *bytes);

/* ------------------------------------------------------------------ k-hop sampling */

#define GRASS_KHOP_MAX_HOPS 8

/* One k-hop request.  Frontier 0 is the seeds with duplicates dropped (first appearance kept);
 * hop h draws, for every frontier vertex f in order, min(fanout[h], deg f) distinct neighbour
 * positions with Floyd's algorithm (draw k from Philox(seed, {instance, f, h, k})) and emits
 * (f, neighbour) in ascending position; frontier h+1 is hop h's destinations with duplicates
 * dropped (first appearance kept).  fanout[h] <= 64. */
typedef struct {
    const uint32_t *seeds_dev;
    uint64_t num_seeds;
    uint32_t instance;
    uint32_t hops;
    uint32_t fanout[GRASS_KHOP_MAX_HOPS];
} grass_khop_request;

/* Output of a batch, hop-major: the edges of (hop h, request r) are
 * [offsets_dev[h*(count+1)+r], offsets_dev[h*(count+1)+r+1]) of src_dev/dst_dev.
 * offsets_dev holds max_hops*(count+1) entries (max_hops = the largest `hops` of the batch). */
typedef struct {
    uint32_t *src_dev;
    uint32_t *dst_dev;
    uint64_t capacity;     /* entries of src_dev and dst_dev */
    uint64_t *offsets_dev;
    uint64_t total;        /* out: edges of the batch (may exceed capacity: GRASS_ECAPACITY) */
    uint32_t max_hops;     /* out */
} grass_khop_output;

/* Upper bound on the edges a batch can produce (for sizing `capacity`). */
int grass_khop_bound(const grass_graph *g, const grass_khop_request *reqs, uint32_t count,
                     uint64_t *max_edges);

/* Runs every request of the batch in the same launches (a fixed number per hop whatever the
 * batch size; counts stay on the device); count <= 131072.  Each request's edges are identical to
 * running it alone.  Replaces KHopApp (LE/src/apps/khop/khop.cu:1-69: fanout [10,25] with replacement)
 * driven by doTransitParallelSampling (LE/src/grass.cu:1299-1900).  Synchronises to report
 * `total`. */
int grass_khop_batch(grass_graph *g, const grass_khop_request *reqs, uint32_t count, uint64_t seed,
                     grass_khop_output *out, void *stream);

/* ------------------------------------------------------------------ multi-dimensional RW */

#define GRASS_MDRW_MAX_FRONTIER 8192
#define GRASS_MDRW_MAX_TARGET 16384

/* One frontier-sampling instance (Ribeiro & Towsley, IMC 2010).  Entry j of the frontier starts
 * at floor((r1:r0) * n / 2^64) of Philox(seed, {instance, j, 0xFFFFFFFF, 0}) (the walk-start draw);
 * then max(0, target - frontier) steps run, step t drawing Philox(seed, {instance, t,
 * 0xFFFFFFFE, 0}) = (r0, r1, r2, r3): x = floor((r1:r0) * W / 2^64) with W the frontier's
 * degree sum picks the first entry whose running degree sum exceeds x (degree-proportional
 * choice); k = floor((r3:r2) * d / 2^64) picks its neighbour k, which replaces the entry.  An
 * instance whose frontier has no edge left stops early.  1 <= frontier <= GRASS_MDRW_MAX_FRONTIER. */
typedef struct {
    uint32_t instance;
    uint32_t frontier;
} grass_mdrw_request;

/* Output of a batch (device memory owned by the result).  Instance i's part of each array is
 * [x_off_dev[i], x_off_dev[i+1]):
 *   steps   — the traversed edges (step_src, step_dst) in step order;
 *   verts   — the sampled vertex set: initial frontier and every step's destination, ascending;
 *   induced — every adjacency entry (x, y) with x and y both in verts, ordered by x and then by
 *             y's position in x's list. */
typedef struct {
    uint32_t count;
    const uint64_t *step_off_dev;
    const uint32_t *step_src_dev, *step_dst_dev;
    const uint64_t *vert_off_dev;
    const uint32_t *verts_dev;
    const uint64_t *ind_off_dev;
    const uint32_t *ind_src_dev, *ind_dst_dev;
    uint64_t steps, verts, induced;  /* totals */
} grass_mdrw_output;

typedef struct grass_mdrw_result grass_mdrw_result;

/* Runs every instance of the batch in one persistent launch (one wave per instance, taken from a
 * device queue), then sorts each sample and extracts its induced edges.  Replaces MultiRWApp
 * (LE/src/apps/multiRW/multiRW.cu:22-45 next, :74-82 stepTransits, :85-96 initializeSample;
 * GT/src/apps/multiRW/multiRW.cu likewise: 80 roots, 100 steps, uniform root choice, no output)
 * driven by doTransitParallelSampling (LE/src/grass.cu:1299-1900).  target <=
 * GRASS_MDRW_MAX_TARGET.  Synchronous on `stream`; *out is freed with grass_mdrw_free. */
int grass_mdrw(grass_graph *g, const grass_mdrw_request *reqs, uint32_t count, uint32_t target,
               uint64_t seed, void *stream, grass_mdrw_result **out);
int grass_mdrw_output_get(const grass_mdrw_result *res, grass_mdrw_output *out);
/* Copies the result to host memory: offsets arrays hold count+1 entries, the others the totals
 * of grass_mdrw_output.  NULL pointers are skipped.  Synchronises. */
int grass_mdrw_fetch(const grass_mdrw_result *res, uint64_t *step_off, uint32_t *step_src,
                     uint32_t *step_dst, uint64_t *vert_off, uint32_t *verts, uint64_t *ind_off,
                     uint32_t *ind_src, uint32_t *ind_dst);
void grass_mdrw_free(grass_mdrw_result *res);

/* ------------------------------------------------------------------ tuning
 * No reference counterpart (the reference fixes its launch shapes at compile time,
 * LE/src/grass.cu:49-52,1716-1767, LE/src/utils.hpp:449-461).  Launch-shape choices of a resident
 * graph; no choice ever changes a result. */
typedef struct {
    uint32_t rw_stage;              /* walk columns a wave stages in LDS per flush: 8, 16, 32 */
    uint32_t rw_claim;              /* 64-walker granules a wave takes per queue access: 1, 2, 4 */
    uint32_t khop_lane_scale;       /* k-hop sampling lanes: 0 = each frontier vertex takes as
                                       many lanes as it has draws (packed); 1 = per hop, packed
                                       when the hop's draws fill under 70 % of groups of the
                                       smallest width covering its fanout, else those groups
                                       (default); 2 = groups of twice that width */
    uint32_t decode_edges_per_wave; /* full-decode output slots per wave: 512 ... 8192, power of 2 */
} grass_tuning;

int grass_get_tuning(const grass_graph *g, grass_tuning *out);
/* GRASS_EINVAL (and no change) for a value outside the lists above. */
int grass_set_tuning(grass_graph *g, const grass_tuning *t);

#define GRASS_TUNE_RW 1u
#define GRASS_TUNE_KHOP 2u
#define GRASS_TUNE_DECODE 4u

/* Timings of every candidate the autotuner tried (milliseconds, best of 3 runs). */
typedef struct {
    uint32_t rw_count;
    uint32_t rw_stage[9], rw_claim[9];
    double rw_ms[9];
    uint32_t khop_count;
    uint32_t khop_lane_scale[3];
    double khop_ms[3];
    uint32_t decode_count;
    uint32_t decode_edges_per_wave[5];
    double decode_ms[5];
} grass_tune_report;

/* Times each candidate of the selected knobs (what = a mask of GRASS_TUNE_*, 0 = all) on
 * synthetic requests drawn from the graph with Philox -- 128 random-walk requests of 2048
 * walkers and lengths 20-80; 256 2-hop [25,10] requests of lognormal(6, 1.2) seed counts clipped
 * to [1, 8192]; one full decode (skipped when its CSR does not fit in free HBM) -- keeps the
 * fastest in the graph and returns it.  A
 * candidate other than the setting in force at the call must be 3 % faster than that setting to
 * replace it (a near tie keeps the setting).  Synchronises; uses the graph's own stream.  report
 * may be NULL. */
int grass_autotune(grass_graph *g, uint32_t what, grass_tuning *chosen, grass_tune_report *report);

/* ------------------------------------------------------------------ captured batches
 * No reference counterpart (the reference launches and synchronises every step from the host,
 * LE/src/grass.cu:1299-1900).  A plan records a fixed batch -- random-walk requests and k-hop
 * requests with their device buffers -- into a HIP graph, the walks and the k-hop pipeline on
 * two branches, after one eager run that sizes every workspace.  Pass an explicit stream to
 * grass_plan_launch to order later work on that stream after the replay.  grass_plan_launch
 * replays it as one submission.  Between launches the contents of the input buffers (walk starts,
 * k-hop seeds) may change; their addresses, the request shapes and the seed may not.  Launches of
 * one plan must not overlap (they share its workspaces); each replay gives the result an eager
 * grass_random_walk_batch + grass_khop_batch of the same batch gives. */
typedef struct grass_plan grass_plan;
int grass_plan_create(grass_graph *g, const grass_rw_request *rw, uint32_t rw_count,
                      const grass_khop_request *khop, uint32_t khop_count,
                      const grass_khop_output *khop_out, uint64_t seed, grass_plan **out);
int grass_plan_launch(grass_plan *p, void *stream);
/* Edge total of the last completed launch's k-hop part (read after the stream passed it). */
int grass_plan_khop_total(const grass_plan *p, uint64_t *total);
int grass_plan_destroy(grass_plan *p);

/* A plan of one MDRW batch: grass_mdrw on `reqs`, `target` and `seed`, run once eagerly (which
 * sizes the outputs and the working memory) and captured; each grass_plan_launch replays it into
 * the same outputs, whose device views grass_plan_mdrw_output returns (valid until
 * grass_plan_destroy; read them after the launch stream has passed the replay).  The requests and
 * the seed are fixed at creation, so every replay writes the result grass_mdrw returns for them.
 * Replaces the step loop that drives MultiRWApp (LE/src/grass.cu:1453-1861 over
 * LE/src/apps/multiRW/multiRW.cu:22-97) with one submission. */
int grass_plan_create_mdrw(grass_graph *g, const grass_mdrw_request *reqs, uint32_t count,
                           uint32_t target, uint64_t seed, grass_plan **out);
int grass_plan_mdrw_output(const grass_plan *p, grass_mdrw_output *out);

/* ------------------------------------------------------------------ serving
 * No reference counterpart (the reference runs one compile-time app over every vertex,
 * LE/src/grass.cu:1924-1992).  A session keeps the graph resident and runs a scheduler thread
 * that does continuous batching: requests that arrive while others are on the GPU are admitted
 * as soon as HBM for their inputs and outputs is free in the session's arena, grouped into
 * launch groups (one persistent random-walk launch + one k-hop pipeline) and put on one of
 * `max_inflight` streams, so groups overlap on the device.  A request's result depends only on
 * its contents, its instance and the session seed -- never on what it was batched with.  Beside
 * the arena, each of the `max_inflight` slots holds working memory sized at creation for the
 * largest group `max_group_edges` admits (up to 8M edges; about 0.65 GB of device and 36 MB of
 * pinned host memory per slot at 6M), so groups never reallocate it while serving. */

typedef enum { GRASS_REQ_RW = 0, GRASS_REQ_KHOP = 1 } grass_request_kind;

typedef struct {
    int32_t kind;              /* grass_request_kind */
    uint32_t instance;         /* Philox instance */
    /* random walk */
    uint64_t walkers;
    uint32_t length;
    const uint32_t *starts;    /* host, `walkers` ids; NULL = Philox-drawn starts */
    /* k-hop */
    const uint32_t *seeds;     /* host, `num_seeds` ids */
    uint64_t num_seeds;
    uint32_t hops;
    uint32_t fanout[GRASS_KHOP_MAX_HOPS];
} grass_request;

typedef struct {
    uint64_t memory_budget;    /* HBM arena for request inputs and outputs (0 = 4 GiB) */
    uint32_t max_batch;        /* requests per launch group (0 = 512) */
    uint32_t max_inflight;     /* launch groups on the device at once (0 = 2) */
    uint64_t seed;
    uint64_t max_group_edges;  /* edge bound that closes a group early (0 = 6M) */
} grass_session_config;

typedef struct {
    int32_t kind;
    int32_t status;            /* GRASS_OK, GRASS_ECANCELLED, or the error of its launch */
    const uint32_t *walks_dev; /* random walk: [walkers][length+1] */
    uint64_t walkers;
    uint32_t length;
    uint32_t hops;             /* k-hop: edges of hop h are hop_src_dev[h][0..hop_edges[h]) */
    const uint32_t *hop_src_dev[GRASS_KHOP_MAX_HOPS];
    const uint32_t *hop_dst_dev[GRASS_KHOP_MAX_HOPS];
    uint64_t hop_edges[GRASS_KHOP_MAX_HOPS];
    double submit_us, launch_us, done_us; /* host clock, microseconds since session start */
} grass_result;

typedef struct {
    uint64_t submitted, completed, rejected, cancelled, groups;
    uint64_t sampled_edges;
    uint64_t memory_in_use, memory_peak;
} grass_session_counters;

typedef struct grass_session grass_session;

int grass_session_create(grass_graph *g, const grass_session_config *cfg, grass_session **out);
/* Copies the request (the caller may free its arrays on return).  GRASS_ECAPACITY when the
 * request could never fit the arena. */
int grass_session_submit(grass_session *s, const grass_request *req, uint64_t *ticket);
/* GRASS_OK when done (result ready), GRASS_ENOTREADY, GRASS_ECANCELLED, GRASS_EINVAL (unknown). */
int grass_session_poll(grass_session *s, uint64_t ticket);
/* Waits up to timeout_us (< 0: forever); GRASS_ETIMEOUT if still running. */
int grass_session_wait(grass_session *s, uint64_t ticket, int64_t timeout_us);
/* Device views of a finished request's output; valid until grass_session_release. */
int grass_session_result(grass_session *s, uint64_t ticket, grass_result *out);
/* Copies a finished request's output to host memory: a random walk as its [walkers][length+1]
 * matrix; a k-hop request as all its sources (hop 0 first) followed by all its destinations.
 * *bytes receives the size; GRASS_ECAPACITY (nothing copied) when it exceeds `capacity`. */
int grass_session_fetch(grass_session *s, uint64_t ticket, void *host, uint64_t capacity,
                        uint64_t *bytes);

/* Returns the request's HBM to the arena and forgets the ticket. */
int grass_session_release(grass_session *s, uint64_t ticket);
/* A queued request is dropped; a running one is released when it finishes. */
int grass_session_cancel(grass_session *s, uint64_t ticket);
int grass_session_stats(grass_session *s, grass_session_counters *out);
/* Waits for running groups, then frees everything. */
int grass_session_destroy(grass_session *s);

/* Open-loop trace replay for latency measurement: submits reqs[i] at arrival_us[i]
 * (microseconds after the call), releases each result as it completes, and writes
 * latency_us[i] = completion - arrival (-1 when rejected) and, when queue_us is not NULL,
 * queue_us[i] = launch of its group - arrival.  *wall_us receives the time from the first
 * arrival to the last completion. */
int grass_session_replay(grass_session *s, const grass_request *reqs, const double *arrival_us,
                         uint32_t count, double *latency_us, double *queue_us, double *wall_us);

#ifdef __cplusplus
}
#endif
#endif
