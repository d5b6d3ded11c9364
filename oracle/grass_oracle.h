/*
 * grass_oracle.h -- CPU oracle for the GraSS compressed-graph sampling path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's decoder and of
 * the sampling semantics, which tests/, __graft_entry__.smoke() and the cpu_baseline leg of
 * bench.py check results against.  No product code links, loads or calls it.
 *
 * Reference anchors (paths relative to the reference checkout):
 *   .graph word view  LE/src/grass.cu:1036-1050      .offset fields  LE/src/bcg.cuh:284-329
 *   n formula         LE/src/grass.cu:1071-1088      exclusive scan  LE/src/utils.hpp:337-342
 *   gamma / fixed     LE/src/bcg.cuh:59-98           LE header       LE/src/bcg.cuh:131-152
 *   LE get_vertex     LE/src/bcg.cuh:196-261         GT header       GT/src/bcg.cuh:110-136
 *   GT get_vertex     GT/src/bcg.cuh:139-168
 * The sampling semantics (Philox keying, start draws, Floyd subsets, per-hop dedup, MDRW
 * degree-proportional selection) have no reference counterpart (the reference draws from
 * cuRAND and never materialises samples); they are defined here and in the HIP kernels alike.
 *
 * Pinning (what checks each function, see tests/test_oracle.py):
 *   or_open / or_get / or_decode_csr : round trip against the golden fixtures written by the
 *       compiled reference compressors (decoded lists == compressor input), brute force over
 *       every (v, k) on the fixtures, the hand-derived worked example.
 *   or_parse_offsets : fixture record lengths summed == graph bits (< 8 bits padding), the
 *       n-overcount fixture.
 *   or_philox4x32_10 : Random123 known-answer vectors and the rocRAND host implementation.
 *   or_rw / or_khop / or_mdrw : invariants against the golden CSR (each walk step is an edge,
 *       dead ends pad with n; k-hop subsets are distinct, ascending, of size min(fanout, d),
 *       frontiers are first-appearance dedups; MDRW steps are edges and the induced list is
 *       exactly the golden CSR restricted to the sample), chi-square tests of uniform neighbour
 *       choice, uniform subsets and degree-proportional frontier choice, and determinism per
 *       (seed, instance).  Bit-exactness against the reference itself is "parity unpinned":
 *       the reference's outputs are never materialised and its RNG is closed cuRAND.
 */
#ifndef GRASS_ORACLE_H
#define GRASS_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    OR_OK = 0,
    OR_EINVAL = -1,    /* bad argument */
    OR_EFORMAT = -2,   /* malformed .graph/.offset */
    OR_ENOMEM = -3,
};

enum { OR_LE = 0, OR_GT = 1 };

typedef struct or_graph or_graph;

/* Big-endian 32-bit word view of a byte stream, zero-padded (LE/src/grass.cu:1036-1050).
 * `words` must hold nbytes/4 + 4 entries. */
void or_pack_words(const uint8_t *bytes, uint64_t nbytes, uint32_t *words);

/* .offset parse: ubl = byte 0, n = floor(8 (nbytes - 1) / ubl) (LE/src/grass.cu:1071-1088),
 * field u = bits [u*ubl, (u+1)*ubl) after byte 0 (LE/src/bcg.cuh:284-329).  `lens` may be NULL
 * to query n_formula; otherwise it must hold n_formula entries. */
int or_parse_offsets(const uint8_t *off, uint64_t nbytes, uint32_t *ubl, uint64_t *n_formula,
                     uint64_t *lens);

/* Opens an in-memory graph (copies both files).  The vertex count is n_formula with the
 * trailing zero-length fields dropped (records are never empty, so those are padding). */
int or_open(const uint8_t *graph, uint64_t graph_bytes, const uint8_t *offset,
            uint64_t offset_bytes, int fmt, or_graph **out);
void or_close(or_graph *g);
/* Byte model only (decoded values never change): GT neighbour reads are counted as the device
 * performs them, through a checkpoint every `interval` in-block positions (default 4; 0 counts
 * the head plus every gap). */
void or_set_gt_checkpoint(or_graph *g, uint32_t interval);
uint64_t or_num_vertices(const or_graph *g);
uint64_t or_total_bits(const or_graph *g);
uint64_t or_record_start(const or_graph *g, uint64_t v);

/* Header fields: out = {d, f1, f2, first_pos, chunks/blocks C, q, R, payload_bit, header_bits}
 * with f1/f2 = rock_bit/sea_bit (LE) or head_bit/body_bit (GT).  For LE d == 1, f1 holds the
 * neighbour itself.  Returns OR_EFORMAT on an undecodable header. */
int or_header(const or_graph *g, uint64_t v, int64_t out[9]);
int64_t or_degree(const or_graph *g, uint64_t v);
/* k-th neighbour (random access); -1 on a bad (v, k). */
int64_t or_get(const or_graph *g, uint64_t v, uint64_t k);
/* Bits of payload the codec reads for neighbour k (LE: rocks and sea; GT: head + i bodies). */
int64_t or_payload_bits(const or_graph *g, uint64_t v, uint64_t k);

/* Full decode: rowptr[n+1], col[m].  With col == NULL only rowptr is written. */
int or_decode_csr(const or_graph *g, uint64_t *rowptr, uint32_t *col, int nthreads);

/* Philox4x32-10 (Salmon et al., SC'11): out = philox(ctr, key). */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* floor(r64 * bound / 2^64) with r64 = out[1]:out[0] of philox(key=seed, ctr={c0,c1,c2,c3}). */
uint64_t or_uniform(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                    uint64_t bound);

/* Uniform random walks.  out is row-major [W][L+1]; a walk that reaches a vertex with no
 * neighbours is padded with n.  starts == NULL draws start w from philox(seed,
 * {instance, w, 0xFFFFFFFF, 0}).  bytes (may be NULL) receives the algorithmic byte count. */
int or_rw(const or_graph *g, uint64_t seed, uint32_t instance, const uint32_t *starts,
          uint64_t W, uint32_t L, uint32_t *out, int nthreads, uint64_t *bytes);

/* k-hop neighbourhood sampling without replacement, per-hop dedup.  Request r uses
 * seeds[seed_off[r] .. seed_off[r+1]), instance inst[r], hops[r] <= OR_MAX_HOPS and fanouts
 * fanout[r*OR_MAX_HOPS + h].  Results are fetched with or_khop_result_*. */
#define OR_MAX_HOPS 8
typedef struct or_khop_result or_khop_result;
int or_khop(const or_graph *g, uint64_t seed, uint32_t R, const uint64_t *seed_off,
            const uint32_t *seeds, const uint32_t *inst, const uint32_t *hops,
            const uint32_t *fanout, int nthreads, or_khop_result **out);
/* edge_off[R+1]; hop_count[R*OR_MAX_HOPS]; total edges; algorithmic bytes */
void or_khop_result_sizes(const or_khop_result *res, uint64_t *edge_off, uint64_t *hop_count,
                          uint64_t *total, uint64_t *bytes);
void or_khop_result_edges(const or_khop_result *res, uint32_t *src, uint32_t *dst);
void or_khop_result_free(or_khop_result *res);

/* Multi-dimensional random walk (frontier sampling).  Instance i has frontier size
 * fsize[i], runs max(0, target - fsize[i]) steps, selects the moving frontier entry with
 * probability proportional to degree and moves it to a uniform neighbour.  Output per
 * instance: the traversed edges in step order and the induced edge list of the sampled
 * vertex set (sorted by (src, dst)). */
typedef struct or_mdrw_result or_mdrw_result;
int or_mdrw(const or_graph *g, uint64_t seed, uint32_t I, const uint32_t *inst,
            const uint32_t *fsize, uint32_t target, int nthreads, or_mdrw_result **out);
/* step_off[I+1], vert_off[I+1], ind_off[I+1], algorithmic bytes */
void or_mdrw_result_sizes(const or_mdrw_result *res, uint64_t *step_off, uint64_t *vert_off,
                          uint64_t *ind_off, uint64_t *bytes);
void or_mdrw_result_data(const or_mdrw_result *res, uint32_t *step_src, uint32_t *step_dst,
                         uint32_t *verts, uint32_t *ind_src, uint32_t *ind_dst);
void or_mdrw_result_free(or_mdrw_result *res);

#ifdef __cplusplus
}
#endif
#endif
