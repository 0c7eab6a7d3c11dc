/*
 * mi355graph -- host-side graph ingest, file formats and seeded synthetic
 * workloads for the MI355X SSSP engine (libmi355graph.so, plain C ABI).
 *
 * A host graph is an opaque handle owning a CSR in one orientation:
 *   int64 rowptr[rows+1], int32 col[nnz], 32-bit weights[nnz]
 * whose weight words are either IEEE fp32 or uint32 (MI355G_W_F32/_U32).
 * Pointers returned by mi355g_view stay valid until mi355g_free.
 *
 * Every generator is a pure function of its arguments: it draws from a
 * counter-based SplitMix64 stream keyed by (seed, stream id, index), so the
 * same call yields the same bytes on any machine (the routing-grid lognormal
 * weights go through libm exp/log; see mi355g_gen_routing_grid).
 *
 * All calls return 0 on success or a negative MI355G_E* code; the message of
 * the last failure on the calling thread is available from mi355g_last_error.
 */
#ifndef MI355_GRAPH_H
#define MI355_GRAPH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MI355G_W_F32 = 0, MI355G_W_U32 = 1 };
enum { MI355G_OUT = 0, MI355G_IN = 1 };
enum {
  MI355G_OK = 0,
  MI355G_EINVAL = -1,
  MI355G_ENOMEM = -2,
  MI355G_EIO = -3,
  MI355G_EFORMAT = -4
};

typedef struct mi355g_graph mi355g_graph;

typedef struct {
  int64_t rows;
  int64_t cols;
  int64_t nnz;
  int32_t weight_type;  /* MI355G_W_F32 or MI355G_W_U32 */
  int32_t orientation;  /* MI355G_OUT or MI355G_IN */
  const int64_t* rowptr;
  const int32_t* col;
  const uint32_t* w;    /* weight words (reinterpret as float for W_F32) */
  /* RIPSCSR1 bookkeeping (declared / loaded edge counts of the source file) */
  uint64_t declared_edges;
  uint64_t loaded_edges;
} mi355g_view_t;

const char* mi355g_last_error(void);
int mi355g_view(const mi355g_graph* g, mi355g_view_t* out);
void mi355g_free(mi355g_graph* g);

/* Build from caller arrays (copied). */
int mi355g_from_csr(int64_t rows, int64_t cols, int64_t nnz, const int64_t* rowptr,
                    const int32_t* col, const uint32_t* w, int32_t weight_type,
                    int32_t orientation, mi355g_graph** out);
/* Build from an edge list (u->v, w); rows sorted by column, duplicates kept
 * (dedup_min = 0) or reduced by min (dedup_min = 1). */
int mi355g_from_edges(int64_t n, int64_t m, const int32_t* src, const int32_t* dst,
                      const uint32_t* w, int32_t weight_type, int32_t orientation,
                      int32_t dedup_min, mi355g_graph** out);
/* Stable transpose: OUT <-> IN. */
int mi355g_transpose(const mi355g_graph* g, mi355g_graph** out);

/* ---- file formats ---------------------------------------------------- */
/* RIPSCSR1 (BF/src/gr_to_csr.cpp:34-57, RT/interchange_to_csr.cpp:1350-1388):
 * incoming orientation, int64 rowptr, int32 colind, f32 values. */
int mi355g_read_ripscsr1(const char* path, mi355g_graph** out);
int mi355g_write_ripscsr1(const mi355g_graph* g, const char* path);
/* RIPSADJ1 (BF/src/gr_to_adjacency.cpp:35-37, 180-213): dense incoming matrix
 * A[v*n+u] = w(u->v), INF = no edge; duplicates reduced by min.  The reader
 * reports n (matrix == NULL) or fills matrix[0..n*n). */
int mi355g_write_ripsadj1(const mi355g_graph* g, const char* path);
int mi355g_read_ripsadj1(const char* path, float* matrix, int64_t capacity, int64_t* n);
/* Dense row-major matrix -> CSR of its finite entries (orientation MI355G_OUT:
 * adj[u*n+v] = w(u->v), as dense_to_csr_outgoing in ONE/sssp_hip.cpp:320-348;
 * MI355G_IN: A[v*n+u] = w(u->v)). */
int mi355g_from_dense(const float* adj, int64_t n, int32_t orientation, int32_t skip_diagonal,
                      int32_t reject_negative, mi355g_graph** out);
/* DIMACS .gr text (BF/src/gr_to_csr.cpp:124-235): 1-based ids, rows sorted by
 * column, duplicate (u,v) reduced by min; result has incoming orientation. */
int mi355g_read_dimacs(const char* path, mi355g_graph** out);
/* Galois binary .gr v1 (ADDSF/csr_graph.cpp:153-309): outgoing orientation,
 * u64 row ends, u32 destinations, optional f32 edge data. */
int mi355g_read_galois_gr(const char* path, mi355g_graph** out);
int mi355g_write_galois_gr(const mi355g_graph* g, const char* path);

/* ---- seeded synthetic workloads (outgoing orientation) ------------------ */
/* rows x cols 4-neighbour grid, both directions, independent uint32 weights
 * uniform in [wmin, wmax] per directed edge. */
int mi355g_gen_grid4(int32_t rows, int32_t cols, uint32_t wmin, uint32_t wmax,
                     uint64_t seed, mi355g_graph** out);
/* R-MAT(a,b,c) with 2^scale vertices and edgefactor*2^scale draws; self loops
 * and duplicate undirected pairs removed, then symmetrised with one shared
 * fp32 weight (u24+1)*2^-24 in (0,1] per undirected edge; vertex ids
 * scrambled by a seeded permutation when permute != 0. */
int mi355g_gen_rmat(int32_t scale, int32_t edgefactor, double a, double b, double c,
                    uint64_t seed, int32_t permute, mi355g_graph** out);
/* side x side routing grid: 4-neighbour undirected edges each deleted with
 * probability p_delete, each of the two diagonals of every unit cell added
 * with probability p_diag, symmetric, uint32 weights max(1, ceil(exp(mu +
 * sigma*Z))) with Z standard normal (Box-Muller). */
int mi355g_gen_routing_grid(int32_t side, double p_delete, double p_diag, double mu,
                            double sigma, uint64_t seed, mi355g_graph** out);
/* Chain 0->1->...->n-1 (weight wchain) plus avg_deg random out-edges per
 * vertex (no self loops) with weights uniform in [wmin, wmax] (f32 kind:
 * fp32 uniform; u32 kind: integers). */
int mi355g_gen_chain_random(int64_t n, int32_t avg_deg, double wmin, double wmax,
                            int32_t weight_type, uint64_t seed, mi355g_graph** out);
/* count seeded sources among vertices with out-degree > 0. */
int mi355g_pick_sources(const mi355g_graph* g, int32_t count, uint64_t seed,
                        int32_t* sources);
/* SHA-256 of the CSR bytes (rowptr, col, weights, little endian) as 64 hex
 * chars + NUL. */
int mi355g_sha256(const mi355g_graph* g, char* hex65);

/* Counter-based generator primitives, exposed so workloads built elsewhere
 * (query boxes, min-plus matrices) draw from the same streams. */
uint64_t mi355g_rng_u64(uint64_t seed, uint64_t stream, uint64_t index);
/* fp32 in (0,1]: (top 24 bits + 1) * 2^-24, exact in fp32. */
float mi355g_rng_unit_f32(uint64_t seed, uint64_t stream, uint64_t index);
/* Fill buf[0..count) with unit_f32(seed, stream, first+i). */
void mi355g_fill_unit_f32(uint64_t seed, uint64_t stream, uint64_t first,
                          int64_t count, float* buf);

/* ---- RIPSIFM1 routing metadata sidecar ---------------------------------
 * The FPGA-interchange converter writes a RIPSCSR1 graph plus this sidecar
 * (Routing/interchange_to_csr.cpp:1390-1587, write_metadata; magic
 * "RIPSIFM1", format version 2, incoming orientation): a string table, the
 * device node id of every CSR node, one (tile string, pip index) record per
 * CSR edge in colind order, the pip table (wire0, wire1, forward), sink
 * site-pin nodes, route requests (net, logical net, source and sink site-pin
 * nodes), logical cell / net / port-instance summaries, blocked and
 * sink-stop node lists and the raw netlist payloads.  All records are
 * little-endian u64; "no index" is UINT64_MAX. */
typedef struct mi355g_ifm mi355g_ifm;

typedef struct {
  uint64_t version, orientation;
  uint64_t string_count, node_count, edge_attr_count, pip_data_count;
  uint64_t site_pin_attr_count, route_request_count;
  uint64_t blocked_node_count, sink_stop_node_count;
  uint64_t logical_cell_count, logical_net_count, logical_port_instance_count;
  uint64_t physical_netlist_byte_count, logical_netlist_byte_count;
  uint64_t device_path_string, physical_path_string, logical_path_string;
  uint64_t logical_design_name_string;
  uint64_t route_source_count, route_sink_count;  /* totals over all requests */
} mi355g_ifm_header_t;

/* Reads and checks the whole sidecar (counts against the file size, node and
 * pip references in range, string references in range or UINT64_MAX). */
int mi355g_read_ripsifm1(const char* path, mi355g_ifm** out);
void mi355g_ifm_free(mi355g_ifm* m);
int mi355g_ifm_header(const mi355g_ifm* m, mi355g_ifm_header_t* out);
/* String `index`: bytes (not NUL-terminated) valid until mi355g_ifm_free. */
int mi355g_ifm_string(const mi355g_ifm* m, uint64_t index, const char** data, uint64_t* len);
/* Fixed-width tables copied into caller arrays sized from the header (NULL
 * skips a table): node_device_ids[node_count]; edge_attrs[2 * edge_attr_count]
 * (tile string, pip index); pip_data[3 * pip_data_count] (wire0, wire1,
 * forward); site_pins[3 * site_pin_attr_count] (node, site, pin);
 * blocked[blocked_node_count]; sink_stop[sink_stop_node_count]. */
int mi355g_ifm_tables(const mi355g_ifm* m, uint64_t* node_device_ids, uint64_t* edge_attrs,
                      uint64_t* pip_data, uint64_t* site_pins, uint64_t* blocked,
                      uint64_t* sink_stop);
/* Route requests: nets[2 * count] (net string, logical net index);
 * src_offsets[count + 1] and sources[3 * route_source_count] (node, site,
 * pin); sink_offsets[count + 1] and sinks[3 * route_sink_count]. */
int mi355g_ifm_requests(const mi355g_ifm* m, uint64_t* nets, uint64_t* src_offsets,
                        uint64_t* sources, uint64_t* sink_offsets, uint64_t* sinks);
/* Logical summaries: cells[3 * cell_count] (declaration name, net begin, net
 * count); nets[4 * net_count] (name, cell, port-instance begin, count);
 * ports[7 * port_instance_count] (port name, instance name, port index,
 * instance index, bus index, has bus index, is external port). */
int mi355g_ifm_logical(const mi355g_ifm* m, uint64_t* cells, uint64_t* nets, uint64_t* ports);
/* The raw decompressed .phys / .netlist payloads (valid until mi355g_ifm_free). */
int mi355g_ifm_payloads(const mi355g_ifm* m, const uint8_t** physical, const uint8_t** logical);

#ifdef __cplusplus
}
#endif
#endif
