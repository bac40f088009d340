/*
 * Checker driver for the reference Graph500 generator.
 *
 * Linked (by oracle/Makefile, target _ref/gen_ref) against the reference's
 * own generator sources, SEQ variant, built with the flags of
 * generator/Makefile.seq.  It writes the raw edge array of
 * make_graph(scale, nedges, seed1, seed2, {.57,.19,.19,.05}) to a file so
 * the tests can compare it, as a multiset, with the engine's generator.
 * Test infrastructure only: nothing in the engine links it.
 *
 * usage: gen_ref SCALE NEDGES SEED1 SEED2 OUTFILE
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
/* The SEQ prototype of make_graph (generator/make_graph.c:33); the shipped
 * header only parses with GRAPH_GENERATOR_MPI defined. */
void make_graph(int log_numverts, int64_t desired_nedges, uint64_t userseed1, uint64_t userseed2,
                const double initiator[4], int64_t *nedges_ptr, int64_t **result_ptr);

int main(int argc, char **argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s SCALE NEDGES SEED1 SEED2 OUTFILE\n", argv[0]);
    return 2;
  }
  int scale = atoi(argv[1]);
  int64_t nedges_req = strtoll(argv[2], NULL, 10);
  uint64_t s1 = strtoull(argv[3], NULL, 10), s2 = strtoull(argv[4], NULL, 10);
  double ini[4] = {.57, .19, .19, .05};
  int64_t nedges = 0, *edges = NULL;
  make_graph(scale, nedges_req, s1, s2, ini, &nedges, &edges);
  FILE *f = fopen(argv[5], "wb");
  if (!f) return 1;
  fwrite(edges, sizeof(int64_t), 2 * (size_t)nedges, f);
  fclose(f);
  printf("%lld\n", (long long)nedges);
  free(edges);
  return 0;
}
