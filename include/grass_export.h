/*
 * grass_export.h -- sample export of libgrass_amd.so: the samplers' outputs in the forms a GNN
 * mini-batch consumes, built and handed over in HBM.
 *
 * The reference never returns samples: nextdoor sizes a per-sample buffer of
 * getFinalSampleSize() = sum over steps of the product of the step sizes entries
 * (LE/src/grass.cu:1115-1126) and the kernels' writes into it are commented out
 * (grass.cu:219-227).  For a random walk that layout is grass_random_walk's output as it stands:
 * row w of the [walkers][length + 1] matrix is the initial sample in column 0 followed by the
 * walk's finalSamples row of `length` entries, padded with n (INVALID_VERTEX, grass.cu:1182).
 * The calls below add what a GNN layer needs beyond that: local vertex numbering per sample, and
 * hand-over of library-owned arrays without a copy.
 */
#ifndef GRASS_EXPORT_H
#define GRASS_EXPORT_H

#include <stdint.h>

#include "grass.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ k-hop blocks */

/* Mini-batch blocks of a k-hop batch.  For request r, its sampled vertices are numbered
 * (local ids 0, 1, ...) by first appearance in the sequence: the request's seeds, its hop-0
 * destinations, its hop-1 destinations, ... -- so the distinct seeds come first, in order.
 * nodes_dev[node_off_dev[r] + i] is the vertex of local id i of request r; lsrc_dev[e] and
 * ldst_dev[e] are the local ids (within the edge's request) of the batch's edge e, indexed as
 * the batch's src_dev / dst_dev. */
typedef struct {
    uint32_t *nodes_dev;
    uint64_t node_capacity;  /* entries of nodes_dev; sum of num_seeds + batch total always suffices */
    uint64_t *node_off_dev;  /* count + 1 entries */
    uint32_t *lsrc_dev, *ldst_dev;  /* batch total entries each */
    uint64_t nodes;          /* out: vertices over all requests (GRASS_ECAPACITY above capacity) */
} grass_khop_blocks;

/* Builds the blocks of a batch that grass_khop_batch completed with `reqs`, `count` and `batch`
 * (its seeds unchanged since).  Synchronises to report `nodes`.  GRASS_EINVAL when the batch's
 * offsets do not match the requests. */
int grass_khop_blocks_build(grass_graph *g, const grass_khop_request *reqs, uint32_t count,
                            const grass_khop_output *batch, grass_khop_blocks *out, void *stream);

/* ------------------------------------------------------------------ MDRW samples */

/* Local ids of an MDRW result's edges: the local id of vertex x in instance i is its rank in
 * the instance's ascending verts.  which = GRASS_MDRW_EDGES_INDUCED writes `induced` entries
 * (for ind_src / ind_dst), GRASS_MDRW_EDGES_STEPS `steps` entries (for step_src / step_dst).
 * Enqueued on `stream` (NULL = the graph's stream); `g` is the graph the result came from. */
typedef enum { GRASS_MDRW_EDGES_INDUCED = 0, GRASS_MDRW_EDGES_STEPS = 1 } grass_mdrw_edges;
int grass_mdrw_local_edges(grass_graph *g, const grass_mdrw_result *res, int which,
                           uint32_t *lsrc_dev, uint32_t *ldst_dev, void *stream);

/* ------------------------------------------------------------------ DLPack hand-over */

/* The DLPack exchange structs (DLDevice, DLDataType, DLTensor, DLManagedTensor of the DLPack
 * 0.8 ABI, https://github.com/dmlc/dlpack) under this library's names: a
 * grass_dl_managed_tensor* is a DLManagedTensor* for any DLPack consumer (a "dltensor"
 * PyCapsule for torch.utils.dlpack.from_dlpack).  Device type 10 is kDLROCM. */
typedef struct {
    int32_t device_type;
    int32_t device_id;
} grass_dl_device;
typedef struct {
    uint8_t code; /* 0 = signed int, 1 = unsigned int */
    uint8_t bits;
    uint16_t lanes;
} grass_dl_dtype;
typedef struct {
    void *data;
    grass_dl_device device;
    int32_t ndim;
    grass_dl_dtype dtype;
    int64_t *shape;
    int64_t *strides; /* NULL: compact row-major */
    uint64_t byte_offset;
} grass_dl_tensor;
typedef struct grass_dl_managed_tensor {
    grass_dl_tensor dl_tensor;
    void *manager_ctx;
    void (*deleter)(struct grass_dl_managed_tensor *self);
} grass_dl_managed_tensor;

typedef enum {
    GRASS_MDRW_STEP_OFF = 0,
    GRASS_MDRW_STEP_SRC = 1,
    GRASS_MDRW_STEP_DST = 2,
    GRASS_MDRW_VERT_OFF = 3,
    GRASS_MDRW_VERTS = 4,
    GRASS_MDRW_IND_OFF = 5,
    GRASS_MDRW_IND_SRC = 6,
    GRASS_MDRW_IND_DST = 7
} grass_mdrw_array;

/* One array of an MDRW result as a 1-D DLPack tensor on the result's device, without a copy:
 * offsets as int64, vertex ids as int32 (ids < 2^31).  The tensor holds a reference on the
 * result, so its memory outlives grass_mdrw_free until the consumer calls the deleter (exactly
 * once, as DLPack requires). */
int grass_mdrw_to_dlpack(grass_mdrw_result *res, int which, grass_dl_managed_tensor **out);

#ifdef __cplusplus
}
#endif
#endif
