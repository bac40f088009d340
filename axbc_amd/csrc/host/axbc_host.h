/*
 * Host-side graph tooling of the AxBC MI355X engine.
 *
 * Plain C (C11 + OpenMP).  It builds the input graph the sampled-Brandes hot
 * path runs on, with the same results as AxBC's host code at -p 1x1:
 *   - Graph500 Kronecker generator           (generator/*.c, called from
 *                                              src/axbc2d.c:518-527)
 *   - text edge-list reader                   (src/axbc2d.c:356-445)
 *   - mirror / norm_graph / degree reduction  (src/axbc2d.c:333-351,
 *                                              695-841)
 *   - CSC build                               (src/axbc2d.c:883-907)
 *   - 2D lattice generator for the high-diameter workload (input tooling
 *     with no reference counterpart)
 *
 * Edge lists are packed 64-bit keys (u << 32 | v): vertex ids must fit in
 * 32 bits, which the reference requires too (verify_32bit_fit,
 * src/axbc2d.c:844-863), and a packed key sorts in the reference's (u, v)
 * order (cmpedge_1d, src/axbc2d.c:529-537; cmpedge at 1x1, 676-687).
 */
#ifndef AXBC_HOST_H
#define AXBC_HOST_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AXBC_EDGE(u, v) (((uint64_t)(u) << 32) | (uint64_t)(uint32_t)(v))
#define AXBC_SRC(e) ((uint32_t)((e) >> 32))
#define AXBC_DST(e) ((uint32_t)(e))

/* ---- Kronecker generator (kronecker.c) -------------------------------- */

/* Graph500 Kronecker edge list for 2^scale vertices and `nedges` edges with
 * initiator {a,b,c,d} and two 64-bit user seeds: a malloc'ed array of
 * 2*nedges int64 endpoints, (min,max)-oriented, self-loops and duplicates
 * kept.  The multiset of edges equals the reference generator's make_graph()
 * output (generator/make_graph.c:33-101, 271-353); the order does not (AxBC
 * sorts the list before use, src/axbc2d.c:824). */
uint64_t *axbc_kronecker_edges(int scale, int64_t nedges, uint64_t seed1, uint64_t seed2,
                               const double initiator[4], int nthreads);

/* Low-level pieces, exported for the tests. */
typedef struct { uint32_t z[5]; } axbc_mrg;
void axbc_mrg_seed_from_user(uint64_t seed1, uint64_t seed2, axbc_mrg *st);
void axbc_mrg_skip(axbc_mrg *st, uint64_t e_high, uint64_t e_mid, uint64_t e_low);
uint32_t axbc_mrg_next_uint(axbc_mrg *st);
double axbc_mrg_next_double(axbc_mrg *st);
uint64_t axbc_mrg_random_up_to(axbc_mrg *st, uint64_t n);
int64_t axbc_btrd_binomial(int64_t n, double p, axbc_mrg *st);
void axbc_rand_permutation(const axbc_mrg *st, int64_t n, int64_t *perm, int nthreads);

/* ---- Graph ingest (graph.c) ------------------------------------------ */

/* Pack int64 (u,v) pairs into keys (in place is not possible: new array). */
uint64_t *axbc_pack_pairs(const uint64_t *pairs, uint64_t ned);

/* mirror(): when `undirected` is non-zero append (v,u) for every (u,v) with
 * u != v.  Returns the (realloc'ed) array and updates *ned. */
uint64_t *axbc_mirror(uint64_t *ed, uint64_t *ned, int undirected);

/* Sort keys ascending (parallel LSD radix sort). */
void axbc_sort_edges(uint64_t *ed, uint64_t ned, int nthreads);

/* norm_graph(): sort, then drop self-loops and duplicate edges, keeping the
 * first sorted edge unconditionally as the reference does.  Returns the new
 * edge count. */
uint64_t axbc_norm_graph(uint64_t *ed, uint64_t ned, int nthreads);

/* degree_reduction(): one pass of 1-degree vertex removal on the sorted edge
 * list.  Rewrites `ed` with the kept edges (returned count) and adds to
 * reach[v] (length N, zeroed by the caller) the recorded pendants of v.
 * *rem_ed gets the reference's rem_ed (matched removed edges). */
uint64_t axbc_degree_reduction(uint64_t *ed, uint64_t ned, uint64_t N, uint32_t *reach,
                               uint32_t *rem_ed, int nthreads);

/* build_csc(): col[N+1], row[ned]; column = second endpoint, rows in sorted
 * order.  Returns 0, or -1 when the edge count does not fit 32 bits. */
int axbc_build_csc(const uint64_t *ed, uint64_t ned, uint64_t N, uint32_t **col, uint32_t **row);

/* Text edge list ("u v" per line, lines containing '#' are comments).
 * Returns malloc'ed packed keys (mirrored when undirected) or NULL with an
 * error message in errbuf. */
uint64_t *axbc_read_edge_file(const char *path, uint64_t N, int undirected, uint64_t *ned,
                              char *errbuf, size_t errlen);

/* 2D lattice side x side with 4-neighbour edges; each undirected edge is
 * deleted with probability p_del from a seeded splitmix64 stream; only edges
 * of the largest connected component are kept and vertex ids stay
 * r*side+c.  Returns packed undirected keys (u < v, one copy per edge). */
uint64_t *axbc_lattice_edges(uint32_t side, double p_del, uint64_t seed, uint64_t *ned);

/* An assembled 1x1 graph (src/axbc2d.c:2544-2625). */
typedef struct {
  uint64_t N;
  uint64_t ned;       /* directed adjacency entries in the CSC */
  uint64_t generated; /* edges produced by the generator/reader (pre-mirror) */
  uint32_t rem_ed;    /* rem_ed of the degree reduction */
  uint32_t *col;      /* N+1 */
  uint32_t *row;      /* ned */
  uint32_t *reach;    /* N (zeros unless H==1) */
} axbc_graph;

/* The pipeline after ingest: [H!=0: norm] -> [H==1: degree reduction] ->
 * norm -> CSC.  Takes ownership of `ed` (packed keys, already mirrored). */
int axbc_graph_from_edges(uint64_t *ed, uint64_t ned, uint64_t N, int heuristic, int nthreads,
                          axbc_graph *g);
/* gen_graph (make_graph + mirror) followed by axbc_graph_from_edges. */
int axbc_graph_kronecker(int scale, int edgef, uint64_t seed1, uint64_t seed2, int undirected,
                         int heuristic, int nthreads, axbc_graph *g);
void axbc_graph_free(axbc_graph *g);

#ifdef __cplusplus
}
#endif

#endif /* AXBC_HOST_H */
