/*
 * grass_ingest.h -- ingest of libgrass_amd.so: synthetic R-MAT edge lists, the reference's
 * graph_sort preprocessing, and the LE / GT compressors that write the .graph/.offset format.
 * The host calls are parallel (std::thread) over flat arrays; the *_device calls do the same
 * on a GPU over arrays in its memory.  All of them are byte-identical to the reference tools on
 * the same input.
 */
#ifndef GRASS_INGEST_H
#define GRASS_INGEST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct grass_csr grass_csr;   /* host CSR owned by the library */
typedef struct grass_blob grass_blob; /* host byte buffer owned by the library */

/* R-MAT edge list: m = edge_factor << scale edges, vertex ids < 2^scale.  Edge i descends
 * `scale` levels; level l picks quadrant (0,0)/(0,1)/(1,0)/(1,1) with probability a/b/c/rest
 * using 32-bit word l%4 of Philox(seed, {i_lo, i_hi, l/4, 0x524D4154}).  The reference has no
 * generator; this defines the synthetic inputs.  src/dst hold m entries. */
int grass_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                     uint64_t seed, uint32_t *src, uint32_t *dst, int nthreads);

/* graph_sort (LE/chunk_compress/graph_sort.cpp:22-82): drops self loops, numbers vertices
 * by first appearance in the stream (u before v), for undirected input makes each pair
 * (min,max), sorts and removes duplicates, then adds the mirror of every edge.  The result
 * is the sorted edge list the compressor reads, held as CSR (n = number of ids assigned). */
int grass_graph_sort(const uint32_t *src, const uint32_t *dst, uint64_t m, int directed,
                     int nthreads, grass_csr **out);

/* Compressor::load_graph (LE/chunk_compress/compressor.hpp:93-128): a (u,v)-sorted edge list,
 * adjacent duplicates skipped, n = max id + 1.  GRASS_EINVAL if the list is not sorted. */
int grass_csr_from_sorted_edges(const uint32_t *src, const uint32_t *dst, uint64_t m,
                                grass_csr **out);

int grass_csr_info(const grass_csr *csr, uint64_t *n, uint64_t *m);
int grass_csr_arrays(const grass_csr *csr, const uint64_t **rowptr, const uint32_t **col);
void grass_csr_free(grass_csr *csr);

/* The LE (format 0) or GT (format 1) compressor in "normal" mode
 * (LE/chunk_compress/compressor.hpp:288-403,130-285; GT/chunk_compress/compressor.hpp:
 * 245-349,120-243): writes the .graph and .offset images.  As in the reference, a GT block head
 * holds |first - v| - 1, so a GT block of row v that starts with v itself (a self loop, which
 * grass_graph_sort drops) encodes to bytes that do not decode back to that row. */
int grass_compress(const grass_csr *csr, int format, int nthreads, grass_blob **graph,
                   grass_blob **offset);

/* The same compressors run on a GPU: the CSR (rowptr n + 1 entries, col rowptr[n] entries) is in
 * device memory of `device`, each row sorted and free of duplicates as grass_graph_sort writes it.
 * The images are byte-identical to grass_compress's and come back as host blobs; `stream` (a
 * hipStream_t, NULL for the default stream) orders the work after the producer of the CSR.  The
 * reference encodes on the CPU only (same compressor.hpp lines as grass_compress); this entry
 * point has no counterpart there and is for callers whose graph is already on the GPU. */
int grass_compress_device(const uint64_t *rowptr_dev, const uint32_t *col_dev, uint64_t n,
                          int format, int device, void *stream, grass_blob **graph,
                          grass_blob **offset);

/* grass_rmat_edges on a GPU: the same m = edge_factor << scale edges, written to src_dev and
 * dst_dev (device memory of `device`, m entries each) in order on `stream`.  The reference has no
 * generator; this is the device form of the synthetic inputs above. */
int grass_rmat_edges_device(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                            uint64_t seed, int device, void *stream, uint32_t *src_dev,
                            uint32_t *dst_dev);

/* grass_graph_sort (graph_sort.cpp:22-82) on a GPU, over an edge list in device memory: the same
 * CSR, written to rowptr_dev (n + 1 entries) and col_dev (*m_out entries).  *n and *m_out receive
 * the sizes; GRASS_ECAPACITY (sizes set, CSR not written) when rowptr_capacity < n + 1 or
 * col_capacity < *m_out.  n <= max id + 1 and *m_out <= m (directed) or 2m (undirected).  Ids
 * must be below 2^31 and m below 2^34 (GRASS_EINVAL otherwise; the host call takes any id below
 * 2^32 - 1).  Device working memory is about 16 bytes per directed edge.  Synchronises on `stream`. */
int grass_graph_sort_device(const uint32_t *src_dev, const uint32_t *dst_dev, uint64_t m,
                            int directed, int device, void *stream, uint64_t *rowptr_dev,
                            uint64_t rowptr_capacity, uint32_t *col_dev, uint64_t col_capacity,
                            uint64_t *n, uint64_t *m_out);

/* The two device calls above for callers that hold their data on the host (the graph_sort and
 * compressor tools' --device mode): the CSR or the edge list is copied to `device`, processed
 * there, and the result comes back as the host calls return it -- the same bytes.  The device
 * calls' limits apply (ids below 2^31, n below 2^31). */
int grass_compress_on_device(const grass_csr *csr, int format, int device, grass_blob **graph,
                             grass_blob **offset);
int grass_graph_sort_on_device(const uint32_t *src, const uint32_t *dst, uint64_t m, int directed,
                               int device, grass_csr **out);

int grass_blob_data(const grass_blob *blob, const uint8_t **data, uint64_t *size);
void grass_blob_free(grass_blob *blob);

/* Writes <prefix>.graph and <prefix>.offset (Compressor::write_cgr,
 * LE/chunk_compress/compressor.hpp:130-187). */
int grass_write_files(const char *prefix, const grass_blob *graph, const grass_blob *offset);

/* Text edge lists, as the reference tools read and write them.  Reading skips `skip_lines`
 * lines, then takes whitespace-separated unsigned integers in pairs (u, v) until the first token
 * that is not a number below 2^32 - 1 (graph_sort.cpp:41-46 `fin >> u >> v`, compressor.hpp:
 * 93-128 fscanf("%ld %ld")); an unpaired last number is dropped.  Parsing is split over
 * nthreads at whitespace.  Writing prints the CSR's entries as "u v" lines in order
 * (graph_sort.cpp:74-77). */
typedef struct grass_edges grass_edges;
int grass_read_edge_list(const char *path, uint64_t skip_lines, int nthreads, grass_edges **out);
int grass_edges_data(const grass_edges *e, const uint32_t **src, const uint32_t **dst, uint64_t *m);
void grass_edges_free(grass_edges *e);
int grass_write_edge_list(const char *path, const grass_csr *csr, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
