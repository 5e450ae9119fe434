/* Partitions an hMetis file through the C library interface, in C, the way
 * the reference's examples/partition_hypergraph_lib_*.cc do, and prints the
 * objective, the block weights and the partition file.
 *
 * usage: capi_partition <file.hgr> <k> <epsilon> <seed> <partition-file> */
#include <stdio.h>
#include <stdlib.h>

#include "libkahypar.h"

int main(int argc, char** argv) {
  if (argc != 6) {
    fprintf(stderr, "usage: %s <file.hgr> <k> <epsilon> <seed> <partition-file>\n", argv[0]);
    return 2;
  }
  const unsigned int k = (unsigned int)atoi(argv[2]);
  const double eps = atof(argv[3]);
  const int seed = atoi(argv[4]);
  mt_kahypar_context_t* ctx = mt_kahypar_context_new();
  mt_kahypar_initialize_thread_pool(8, true);

  mt_kahypar_hypernode_id_t nv = 0;
  mt_kahypar_hyperedge_id_t ne = 0;
  size_t* idx = NULL;
  mt_kahypar_hyperedge_id_t* pins = NULL;
  mt_kahypar_hyperedge_weight_t* ew = NULL;
  mt_kahypar_hypernode_weight_t* vw = NULL;
  mt_kahypar_read_hypergraph_from_file(argv[1], &nv, &ne, &idx, &pins, &ew, &vw);

  mt_kahypar_hyperedge_weight_t objective = 0;
  mt_kahypar_partition_id_t* part = (mt_kahypar_partition_id_t*)malloc(sizeof(mt_kahypar_partition_id_t) * nv);
  mt_kahypar_partition(nv, ne, eps, k, seed, vw, ew, idx, pins, &objective, ctx, part, false);

  FILE* out = fopen(argv[5], "w");
  if (!out) return 4;
  long long* bw = (long long*)calloc(k, sizeof(long long));
  for (mt_kahypar_hypernode_id_t u = 0; u < nv; ++u) {
    fprintf(out, "%u\n", part[u]);
    bw[part[u]] += vw[u];
  }
  fclose(out);
  printf("objective=%d", objective);
  for (unsigned int b = 0; b < k; ++b) printf(" w%u=%lld", b, bw[b]);
  printf("\n");
  free(bw);
  free(part);
  mt_kahypar_context_free(ctx);
  /* the arrays of mt_kahypar_read_hypergraph_from_file come from new[]; a C
   * program cannot release them and leaves them to process exit */
  return 0;
}
