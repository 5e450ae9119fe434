/* Drop-in for Mt-KaHyPar's C library interface (include/libkahypar.h:47-86 of
 * the reference): the same types, names, argument meaning and ownership
 * rules, with partition() running the MI355X pipeline (Louvain, multilevel
 * coarsening, initial partitioning of the coarsest hypergraph, km1 label
 * propagation at every level).
 *
 * Differences a caller can observe:
 *  - mt_kahypar_initialize_thread_pool only records the host thread count;
 *    the device work does not use a thread pool;
 *  - configure_context_from_file reads the preset keys this path has
 *    (maxnet-ignore, p-*, c-*, i-runs, r-lp-*); FM keys are accepted and the
 *    FM refiner is not run;
 *  - partition() is serialised by a process-wide lock (the reference is not
 *    re-entrant either);
 *  - errors print a message and exit(-1), as the reference's ERROR macro
 *    does (macros.h:108).  mt_kahypar_amd_set_error_mode switches that to a
 *    return code readable with mkp_last_error(). */
#ifndef LIBKAHYPAR_H
#define LIBKAHYPAR_H

#include <stddef.h>
#ifndef __cplusplus
#include <stdbool.h>
#endif

#ifdef __cplusplus
extern "C" {
#endif

#ifndef KAHYPAR_API
#define KAHYPAR_API __attribute__((visibility("default")))
#endif

struct mt_kahypar_context_s;
typedef struct mt_kahypar_context_s mt_kahypar_context_t;

typedef unsigned long int mt_kahypar_hypernode_id_t;
typedef unsigned long int mt_kahypar_hyperedge_id_t;
typedef int mt_kahypar_hypernode_weight_t;
typedef int mt_kahypar_hyperedge_weight_t;
typedef unsigned int mt_kahypar_partition_id_t;

/* libkahypar.h:58 */
KAHYPAR_API mt_kahypar_context_t* mt_kahypar_context_new(void);
/* libkahypar.h:59 */
KAHYPAR_API void mt_kahypar_context_free(mt_kahypar_context_t* kahypar_context);
/* libkahypar.h:60-61 */
KAHYPAR_API void mt_kahypar_configure_context_from_file(mt_kahypar_context_t* kahypar_context,
                                                        const char* ini_file_name);
/* libkahypar.h:63-64 */
KAHYPAR_API void mt_kahypar_initialize_thread_pool(const size_t num_threads, const bool interleaved_allocations);
/* libkahypar.h:66-72; arrays are allocated with new[] and released by the
 * caller with delete[]. */
KAHYPAR_API void mt_kahypar_read_hypergraph_from_file(const char* file_name,
                                                      mt_kahypar_hypernode_id_t* num_vertices,
                                                      mt_kahypar_hyperedge_id_t* num_hyperedges,
                                                      size_t** hyperedge_indices,
                                                      mt_kahypar_hyperedge_id_t** hyperedges,
                                                      mt_kahypar_hyperedge_weight_t** hyperedge_weights,
                                                      mt_kahypar_hypernode_weight_t** vertex_weights);
/* libkahypar.h:74-86; vertex_weights / hyperedge_weights may be NULL (unit
 * weights); objective receives km1. */
KAHYPAR_API void mt_kahypar_partition(const mt_kahypar_hypernode_id_t num_vertices,
                                      const mt_kahypar_hyperedge_id_t num_hyperedges,
                                      const double epsilon,
                                      const mt_kahypar_partition_id_t num_blocks,
                                      const int seed,
                                      const mt_kahypar_hypernode_weight_t* vertex_weights,
                                      const mt_kahypar_hyperedge_weight_t* hyperedge_weights,
                                      const size_t* hyperedge_indices,
                                      const mt_kahypar_hyperedge_id_t* hyperedges,
                                      mt_kahypar_hyperedge_weight_t* objective,
                                      mt_kahypar_context_t* kahypar_context,
                                      mt_kahypar_partition_id_t* partition,
                                      const bool verbose
#ifdef __cplusplus
                                      = false
#endif
);

/* Extensions (no counterpart in the reference). */
/* 0 (default): errors print and exit(-1) like the reference; 1: errors
 * return, leaving the message in mkp_last_error() and *objective = -1. */
KAHYPAR_API void mt_kahypar_amd_set_error_mode(int return_on_error);
/* GPU ordinal used by later partition() calls of this context (default 0). */
KAHYPAR_API void mt_kahypar_amd_set_device(mt_kahypar_context_t* kahypar_context, int device);
/* Copies the context's preset values into *params (an mkp_partition_params
 * of mtkahypar_amd.h); returns 0, or -1 for a NULL argument. */
KAHYPAR_API int mt_kahypar_amd_get_params(const mt_kahypar_context_t* kahypar_context, void* params);

#ifdef __cplusplus
}
#endif

#endif /* LIBKAHYPAR_H */
