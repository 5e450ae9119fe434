/*
 * mtkahypar_amd.h -- C ABI of the MI355X (gfx950) implementation of
 * Mt-KaHyPar's data-parallel core: Louvain local moving on the star
 * expansion, community-restricted heavy-edge clustering and contraction, and
 * km1 label propagation over the partitioned hypergraph.
 *
 * Plain pointers and sizes only.  Every function returns 0 on success and a
 * nonzero status otherwise (1 = invalid argument, 2 = HIP runtime error,
 * 3 = other); mkp_last_error() returns the message of the last failure of the
 * calling thread.  Nothing here calls exit().  Objects created on one context
 * are used from one host thread at a time; distinct contexts are independent.
 *
 * "host" pointers are ordinary host memory; the library copies them to the
 * device.  Device state stays resident between calls.
 */
#ifndef MTKAHYPAR_AMD_H
#define MTKAHYPAR_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MKP_API __attribute__((visibility("default")))

typedef struct mkp_context mkp_context;
typedef struct mkp_hypergraph mkp_hypergraph;
typedef struct mkp_partition mkp_partition;
typedef struct mkp_host_hypergraph mkp_host_hypergraph;

/* ------------------------------------------------------------- errors -- */
MKP_API const char* mkp_last_error(void);

/* ------------------------------------------------------------ context -- */
/* One device plus one stream.  Replaces the TBB arena set up by
 * mt_kahypar_initialize_thread_pool (lib/libkahypar.cc:62-83). */
MKP_API int mkp_context_create(int device, mkp_context** out);
MKP_API int mkp_context_destroy(mkp_context* ctx);
MKP_API int mkp_context_synchronize(mkp_context* ctx);

/* Kernel-family timing with HIP events on the context stream (the
 * counterpart of the reference's utils::Timer phases, utils/timer.h:35-282).
 * Families: 0 pin_count_init, 1 lp_gain, 2 lp_apply, 3 louvain_move,
 * 4 louvain_contract, 5 rating, 6 contraction.  Bytes are the algorithmic
 * (compulsory) HBM bytes of the launches. */
MKP_API int mkp_context_set_profiling(mkp_context* ctx, int enable);
MKP_API int mkp_context_reset_profile(mkp_context* ctx);
MKP_API int mkp_context_get_profile(mkp_context* ctx, int family, double* device_ms, double* algorithmic_bytes,
                                    int64_t* launches);
MKP_API const char* mkp_profile_family_name(int family);
/* Restricts timing to the launches of one kernel (name as reported by
 * mkp_context_profile_kernel; NULL times everything), so that measuring one
 * kernel inside a timed region adds one event pair per launch of it only. */
MKP_API int mkp_context_set_profile_filter(mkp_context* ctx, const char* kernel_name);
/* Choices that never change results: "louvain_group_lanes" (lanes per node
 * of the 17..64-arc Louvain class: 8, 16, 32 or 64; 0, the default, lets the
 * autotuner time the four on the first sub-rounds of a level and keep the
 * fastest per graph kind, class size and mean degree),
 * "louvain_fused_classes" (the Louvain degree classes up to 1024 arcs of a
 * sub-round in one launch: 1, the default, with the 17..64-arc class when it
 * has at most 2^21 arcs; 2 never with it; 3 always with it; 0 one launch per
 * class), "rating_block_passes" and "sleep_wait" (0, the default: host waits
 * on the context's stream spin; 1: they poll for 20 us, then sleep between
 * polls, leaving the core to other threads). */
MKP_API int mkp_context_set_tuning(mkp_context* ctx, const char* key, int64_t value);
/* Per-kernel records: launches timed one at a time (an event pair around a
 * single kernel), keyed by kernel name and template variant, with their
 * summed device time and algorithmic bytes.  *name stays valid until the
 * next reset. */
MKP_API int mkp_context_profile_kernel_count(mkp_context* ctx, int32_t* count);
MKP_API int mkp_context_profile_kernel(mkp_context* ctx, int32_t index, const char** name, double* device_ms,
                                       double* algorithmic_bytes, int64_t* launches);
/* Copies `bytes` bytes device to device `reps` times with 16-byte vector
 * loads and stores and returns the mean time of one copy: a measured copy
 * ceiling, and the calibration kernel for HBM traffic counters. */
MKP_API int mkp_stream_copy(mkp_context* ctx, uint64_t bytes, int32_t reps, double* ms_per_copy);

/* --------------------------------------------------------- hypergraph -- */
typedef struct {
  uint32_t num_vertices;
  uint32_t num_nets;
  uint64_t num_pins;
  uint32_t max_net_size;
  int64_t total_weight;
} mkp_hypergraph_info;

/* Device-resident static CSR hypergraph; incident nets are built on the
 * device in ascending net order.  Replaces StaticHypergraphFactory::construct
 * (datastructures/static_hypergraph_factory.cpp:32-166).  net_offsets has
 * num_nets+1 entries; NULL weights mean unit weights. */
MKP_API int mkp_hypergraph_create(mkp_context* ctx, uint32_t num_vertices, uint32_t num_nets,
                                  const uint64_t* net_offsets, const uint32_t* pins,
                                  const int32_t* vertex_weights, const int32_t* net_weights,
                                  mkp_hypergraph** out);
MKP_API int mkp_hypergraph_destroy(mkp_hypergraph* hg);
MKP_API int mkp_hypergraph_get_info(const mkp_hypergraph* hg, mkp_hypergraph_info* info);
/* StaticHypergraph::setCommunityIDs (datastructures/static_hypergraph.h:790-793) */
MKP_API int mkp_hypergraph_set_communities(mkp_hypergraph* hg, const int32_t* communities);
/* Copies the CSR back (any pointer may be NULL). */
MKP_API int mkp_hypergraph_copy_to_host(const mkp_hypergraph* hg, uint64_t* net_offsets, uint32_t* pins,
                                        uint64_t* vertex_offsets, uint32_t* incident_nets,
                                        int32_t* vertex_weights, int32_t* net_weights, int32_t* communities);

/* ---------------------------------------------------- partition state -- */
/* PartitionedHypergraph (datastructures/partitioned_hypergraph.h:95-126):
 * block ids, block weights, packed pin counts per (net, block)
 * (datastructures/pin_count_in_part.h:45-225) and connectivity bitsets
 * (datastructures/connectivity_set.h:53-221). */
MKP_API int mkp_partition_create(mkp_hypergraph* hg, int32_t k, mkp_partition** out);
MKP_API int mkp_partition_destroy(mkp_partition* p);
/* setOnlyNodePart for every vertex followed by initializePartition
 * (partitioned_hypergraph.h:522-526, 693-698, 1067-1110). */
MKP_API int mkp_partition_set_parts(mkp_partition* p, const int32_t* parts);
/* initializePartition on the block ids already resident on the device. */
MKP_API int mkp_partition_initialize(mkp_partition* p);
/* Device-to-device copy of the block ids of src (same hypergraph) followed by
 * initializePartition. */
MKP_API int mkp_partition_copy_parts(mkp_partition* dst, const mkp_partition* src);
MKP_API int mkp_partition_get_parts(const mkp_partition* p, int32_t* parts);
MKP_API int mkp_partition_get_part_weights(const mkp_partition* p, int64_t* weights);
/* Unpacked pin counts, num_nets*k values (pinCountInPart, partitioned_hypergraph.h:640-645). */
MKP_API int mkp_partition_get_pin_counts(const mkp_partition* p, uint32_t* counts);
/* Connectivity bitsets, num_nets*ceil(k/64) words (connectivity_set.h:80-84,108-116). */
MKP_API int mkp_partition_get_connectivity(const mkp_partition* p, uint64_t* words);
/* isBorderNode for every vertex (partitioned_hypergraph.h:605-620). */
MKP_API int mkp_partition_get_border(const mkp_partition* p, uint8_t* border);
/* metrics::km1 (partition/metrics.cpp:49-65). */
MKP_API int mkp_partition_km1(mkp_partition* p, int64_t* km1);
/* mkp_partition_get_border without any device allocation.  It has no
 * counterpart in the reference: it is for callers that cannot allocate
 * (code under stream capture, tight serving loops, harnesses that forbid
 * allocation).  The per-vertex flags are staged in a fixed device array of
 * 16 MiB per GPU (held until the process ends), so partitions of more than
 * 16,777,216 vertices are refused with status 1.  Calls that overlap on one
 * GPU share that array and race; the call waits for its stream. */
MKP_API int mkp_partition_get_border_noalloc(const mkp_partition* p, uint8_t* border);

/* ---------------------------------------------------------------- LP --- */
typedef struct {
  int32_t max_rounds;              /* r-lp-maximum-iterations (fast preset: 5) */
  int32_t sub_rounds;              /* buckets per round of the synchronous form */
  uint32_t activation_threshold;   /* r-lp-he-size-activation-threshold (100) */
  int32_t rebalancing;             /* r-lp-rebalancing (1) */
  uint64_t seed;
  const int64_t* max_part_weights;     /* k values, host */
  const int64_t* perfect_part_weights; /* k values, host */
} mkp_lp_params;
typedef struct {
  int64_t delta;    /* km1 change */
  int64_t moves;    /* kept moves */
  int64_t reverts;  /* moves undone by the real-delta check */
  int64_t visits;   /* incidences swept by the gain kernels */
  int64_t rounds;
  double device_ms; /* stream time of the refinement */
} mkp_lp_stats;
/* LabelPropagationRefiner<Km1Policy>::refineImpl
 * (refinement/label_propagation/label_propagation_refiner.cpp:33-164), in
 * synchronous sub-rounds. */
MKP_API int mkp_lp_refine(mkp_partition* p, const mkp_lp_params* params, mkp_lp_stats* stats);
/* The same with refinement_nodes (IRefiner::refine, i_refiner.h:49-54): the
 * given vertices (duplicates dropped) form the first active set; an empty
 * list means all vertices, as in initializeActiveNodes
 * (label_propagation_refiner.cpp:194-233). */
MKP_API int mkp_lp_refine_nodes(mkp_partition* p, const mkp_lp_params* params, const uint32_t* nodes,
                                uint32_t num_nodes, mkp_lp_stats* stats);

/* ---------------------------------------------------------------- FM --- */
typedef struct {
  const int64_t* max_part_weights; /* k entries (context.cpp:238-247) */
  int32_t rounds;                  /* r-fm-multitry-rounds (fast preset: 10) */
  int32_t seeds;                   /* r-fm-seed-nodes (25), at most 64 */
  uint32_t ignore_net_size;        /* nets this large are not expanded (1000) */
  uint64_t seed;
  int32_t sub_rounds;              /* groups of searches per round, each applied before the next runs
                                      (0: 8) */
} mkp_fm_params;
typedef struct {
  int64_t improvement; /* km1 decrease */
  int64_t moves;       /* applied moves */
  int64_t rounds;
  int64_t searches;
  double device_ms;    /* wall time of the refinement */
} mkp_fm_stats;
/* MultiTryKWayFM<GainCacheStrategy>::refineImpl
 * (refinement/fm/multitry_kway_fm.cpp:30-150) with LocalizedKWayFM searches
 * (localized_kway_fm_core.cpp:26-190) and GlobalRollback
 * (global_rollback.cpp:146-283), in the synchronous form: the searches of a
 * round run in sub_rounds groups; a group's searches run on the state the
 * earlier groups left, their best prefixes are merged in search order and the
 * best balanced prefix of the merged sequence, with recalculated gains, is
 * applied.  k <= 128. */
MKP_API int mkp_fm_refine(mkp_partition* p, const mkp_fm_params* params, mkp_fm_stats* stats);

/* ----------------------------------------------------------- Louvain --- */
typedef struct {
  int32_t edge_weight;       /* -1 hybrid (density rule of partitioner.cpp:62-72),
                                0 uniform, 1 non_uniform, 2 degree (graph.cpp:46-73) */
  int32_t max_passes;        /* p-max-louvain-pass-iterations (fast preset: 5) */
  double min_move_fraction;  /* p-louvain-min-vertex-move-fraction (0.01) */
  uint64_t sampling;         /* p-vertex-degree-sampling-threshold (200000) */
  uint64_t seed;
  int32_t sub_rounds;        /* buckets per round of the synchronous form */
} mkp_louvain_params;
typedef struct {
  int64_t arcs_swept;  /* arcs of every local-moving round, all levels */
  int64_t rounds;
  int64_t levels;
  int64_t moves;
  double device_ms;
} mkp_louvain_stats;
/* run_parallel_louvain on the star expansion (or the graph itself when every
 * net has two pins), restricted to the hypernodes and stored as the
 * hypergraph's community ids: preprocess() of partitioner.cpp:106-126 with
 * parallel_louvain.cpp:62-66.  communities (host, num_vertices) may be NULL. */
MKP_API int mkp_louvain(mkp_hypergraph* hg, const mkp_louvain_params* params, int32_t* communities,
                        mkp_louvain_stats* stats);
/* Nodes of the graph Louvain runs on (Graph(hypergraph, edge weight),
 * graph.h:66-72): N when every net has two pins (the graph itself), N + E
 * otherwise (the star expansion). */
MKP_API int mkp_louvain_graph_nodes(const mkp_hypergraph* hg, int64_t* num_nodes);
/* run_parallel_louvain(graph, context) (parallel_louvain.h:27) before
 * restrictClusteringToHypernodes (graph.h:123): the clustering of every node
 * of that graph, net e of the star expansion at node N + e as in the
 * reference's Graph.  clustering (host) holds mkp_louvain_graph_nodes()
 * entries; its first N are mkp_louvain's communities, which are stored on the
 * hypergraph as well. */
MKP_API int mkp_louvain_graph(mkp_hypergraph* hg, const mkp_louvain_params* params, int32_t* clustering,
                              mkp_louvain_stats* stats);
/* One evaluation of computeMaxGainCluster (local_moving_modularity.h:87-129)
 * for every node of the level-0 graph (the star expansion with net node
 * N + rank(e), or the graph itself when every net has two pins) on the
 * clustering node_communities (host, one id per node), by the move kernels,
 * and metrics::modularity of it (local_moving_modularity.cpp:28-56).
 * targets (may be NULL): the cluster the node would move to, -1 to stay
 * (including a singleton's move to a larger singleton); gains (may be NULL):
 * max(gain of staying, best candidate's gain) in FP64; modularity may be
 * NULL.  Nothing moves. */
MKP_API int mkp_louvain_evaluate(mkp_hypergraph* hg, const mkp_louvain_params* params,
                                 const int32_t* node_communities, int32_t* targets, double* gains,
                                 double* modularity);

/* -------------------------------------------------------- coarsening --- */
typedef struct mkp_hierarchy mkp_hierarchy;
typedef struct {
  uint32_t contraction_limit;   /* coarsening.contraction_limit (160 * k, context.cpp:256-269) */
  int32_t max_node_weight;      /* context.cpp:271-284 */
  double min_shrink_factor;     /* c-min-shrink-factor (1.01) */
  double max_shrink_factor;     /* c-max-shrink-factor (2.5) */
  uint32_t ignore_net_size;     /* maxnet-ignore (1000) */
  uint64_t sampling;            /* c-vertex-degree-sampling-threshold (200000) */
  uint64_t seed;
  int32_t sub_rounds;           /* buckets per clustering pass of the synchronous form */
} mkp_coarsen_params;
/* One clustering pass (the parallel_for of multilevel_coarsener.h:197-239 with
 * MultilevelVertexPairRater::rate and matchVertices); clusters (host,
 * num_vertices) receives each vertex's representative. */
MKP_API int mkp_cluster_pass(mkp_hypergraph* hg, const mkp_coarsen_params* params, int32_t pass, uint32_t* clusters,
                             uint32_t* nodes_after, int64_t* pin_visits);
/* StaticHypergraph::contract (datastructures/static_hypergraph.cpp:58-513):
 * clusters (host) maps every vertex to its representative; cmap (host, may be
 * NULL) receives the coarse vertex of every vertex. */
MKP_API int mkp_hypergraph_contract(mkp_hypergraph* hg, const uint32_t* clusters, mkp_hypergraph** coarse,
                                    uint32_t* cmap);
/* MultilevelCoarsener::coarsenImpl (multilevel_coarsener.h:143-315): the
 * level hierarchy stays on the device; level 0 is hg itself. */
MKP_API int mkp_coarsen(mkp_hypergraph* hg, const mkp_coarsen_params* params, mkp_hierarchy** out);
MKP_API int mkp_hierarchy_num_levels(const mkp_hierarchy* h, int32_t* levels);
/* Borrowed handle of level i (0 = input, num_levels-1 = coarsest). */
MKP_API int mkp_hierarchy_level(mkp_hierarchy* h, int32_t level, mkp_hypergraph** hg);
MKP_API int mkp_hierarchy_destroy(mkp_hierarchy* h);
/* Projection of doUncoarsen (partition/coarsening/uncoarsening.cpp:111-122):
 * fine (on level i) takes the block of its coarse vertex in coarse (on level
 * i+1), then initializePartition. */
MKP_API int mkp_partition_project(mkp_partition* fine, const mkp_partition* coarse, const mkp_hierarchy* h,
                                  int32_t level);

/* ----------------------------------------------------------- pipeline -- */
typedef struct {
  int32_t k;
  double epsilon;
  uint64_t seed;
  int32_t use_community_detection; /* p-enable-community-detection (1) */
  int32_t lp_sub_rounds;
  int32_t louvain_sub_rounds;
  int32_t coarsening_sub_rounds;
  int32_t ip_runs;                 /* i-runs (20) */
  /* preset values (config/fast_preset.ini); mkp_partition_params_default fills them */
  int32_t louvain_edge_weight;     /* p-louvain-edge-weight-function: -1 hybrid, 0 uniform, 1 non_uniform, 2 degree */
  int32_t louvain_max_passes;      /* p-max-louvain-pass-iterations (5) */
  double louvain_min_move_fraction;/* p-louvain-min-vertex-move-fraction (0.01) */
  uint64_t louvain_sampling;       /* p-vertex-degree-sampling-threshold (200000) */
  double min_shrink_factor;        /* c-min-shrink-factor (1.01) */
  double max_shrink_factor;        /* c-max-shrink-factor (2.5) */
  double max_allowed_weight_multiplier; /* c-s (1) */
  uint32_t contraction_limit_multiplier;/* c-t (160) */
  uint32_t ignore_net_size;        /* maxnet-ignore (1000) */
  uint64_t rating_sampling;        /* c-vertex-degree-sampling-threshold (200000) */
  int32_t lp_enabled;              /* r-lp-type: label_propagation_km1 (1) or do_nothing (0) */
  int32_t lp_max_rounds;           /* r-lp-maximum-iterations (5) */
  uint32_t lp_activation_threshold;/* r-lp-he-size-activation-threshold (100) */
  int32_t lp_rebalancing;          /* r-lp-rebalancing (1) */
  int32_t ip_threads;              /* unused (initial partitioning runs on the device); kept for the layout */
  double large_net_remove_factor;  /* maxnet-remove-factor (0.01): nets larger than max(N * it, 100, 50000) are
                                      set aside during the multilevel phase */
  int32_t fm_rounds;               /* r-fm-multitry-rounds (10 with r-fm-type=fm_gain_cache); default 0: FM
                                      not run (it takes one to two orders of magnitude longer than LP) */
  int32_t fm_seeds;                /* r-fm-seed-nodes (25), at most 64 */
} mkp_partition_params;
/* Initial partitioning of p's hypergraph (the coarsest level) on the device:
 * recursive bisection (recursive_bisection_initial_partitioner.cpp:475-492)
 * in which every bisection is the best balanced of `runs` runs made at once,
 * one workgroup each (greedy growing, BFS or random fill by run index mod 3,
 * then 2-way gain-move passes), in place of the flat pool of
 * flat/pool_initial_partitioner.h:36-75.  Block weights are bounded by
 * (1 + epsilon) * ceil(W / k) (context.cpp:237-250); p is initialised with
 * the result. */
MKP_API int mkp_initial_partition(mkp_partition* p, double epsilon, uint64_t seed, int32_t runs, int64_t* km1);
/* The runs of one bisection of hg as mkp_initial_partition makes them, for
 * inspection and tests: side limits max0 / max1, side-0 target weight
 * target0, run stream key `stream`; sides receives runs x n side ids (host
 * memory), results runs x {cut, balanced, side-0 weight, side-1 weight}.
 * Diagnostic entry point with no reference counterpart (one flat run of the
 * pool, flat/pool_initial_partitioner.h:36-75, is the closest). */
MKP_API int mkp_ip_bisect_runs(mkp_hypergraph* hg, int64_t max0, int64_t max1, int64_t target0, uint64_t stream,
                               int32_t runs, int32_t passes, uint8_t* sides, int64_t* results);
/* Km1Rebalancer (rebalancer.cpp:36-262) on the host: empty blocks get a
 * vertex, then border vertices of overloaded blocks move to the fitting block
 * of best km1 gain (moves that do not worsen km1 first, the rest in
 * increasing gain order while their block is over its bound); max_part_weights
 * holds k bounds.  The partition state is rebuilt on the device. */
MKP_API int mkp_rebalance(mkp_partition* p, const int64_t* max_part_weights, int64_t* km1_delta, int64_t* moves);
/* Fast-preset defaults with the given k, epsilon and seed. */
MKP_API int mkp_partition_params_default(mkp_partition_params* params, int32_t k, double epsilon, uint64_t seed);
typedef struct {
  int64_t km1;
  double imbalance;
  int32_t levels;
  double louvain_ms, coarsening_ms, initial_ms, refinement_ms, total_ms;
  int64_t louvain_arcs, rating_visits, lp_visits;
  int64_t removed_vertices, removed_nets;  /* degree-zero vertices and large nets set aside */
  double fm_ms;                            /* FM part of refinement_ms */
  int64_t fm_improvement;                  /* km1 decrease by FM over all levels */
} mkp_partition_stats;
/* partition() of partitioner.cpp:163-204 with the fast preset (community
 * detection, removal of degree-zero vertices and large nets, multilevel
 * coarsening, initial partitioning of the coarsest hypergraph, LP and then
 * (with fm_rounds > 0) FM at every level (uncoarsening.cpp:94-188), the km1
 * rebalancer when the top
 * level is over its bounds, restore of the removed vertices by bin packing).
 * parts (host, num_vertices) receives the block ids. */
MKP_API int mkp_partition_hypergraph(mkp_hypergraph* hg, const mkp_partition_params* params, int32_t* parts,
                                     mkp_partition_stats* stats);
/* The removals of partition()'s sanitize step (partitioner.cpp:74-86):
 * degree-zero vertices in id order while more than contraction_limit
 * vertices remain (degree_zero_hn_remover.h:46-61) and nets with more than
 * net_threshold pins (large_he_remover.h:47-58), as one compaction on the
 * device.  *reduced receives the reduced hypergraph, owned by the caller
 * (NULL when nothing was removed); vmap (host, num_vertices, may be NULL)
 * the reduced id of every vertex, 0xffffffff for a removed one.  Kept
 * vertices and nets keep their order and a net keeps its pin order;
 * communities are carried along. */
MKP_API int mkp_hypergraph_reduce(mkp_hypergraph* hg, uint32_t contraction_limit, uint32_t net_threshold,
                                  mkp_hypergraph** reduced, uint32_t* vmap, uint32_t* removed_vertices,
                                  uint32_t* removed_nets);
/* max(max(num_vertices * factor, 100), 50000) (partitioner.cpp:37-38,
 * large_he_remover.h:90-92): nets above it are removed. */
MKP_API uint32_t mkp_large_net_threshold(uint32_t num_vertices, double factor);

/* ---------------------------------------------------------- generator -- */
typedef struct {
  uint32_t num_vertices;
  uint32_t num_nets;
  double exponent;          /* net sizes ~ s^-exponent on [min_size, max_size] */
  uint32_t min_size, max_size;
  uint32_t window;          /* locality window width in vertex ids, 0 = none */
  double local_fraction;    /* share of pins drawn inside the window */
  int32_t vertex_weight_min, vertex_weight_max;
  uint64_t seed;
} mkp_powerlaw_params;
typedef struct {
  uint32_t num_rows;        /* rows become vertices */
  uint32_t num_cols;        /* columns become nets */
  double mean_nnz;          /* nnz per row ~ power law with this mean on [1, max_nnz] */
  uint32_t max_nnz;
  uint32_t band;            /* band width in column ids */
  double band_fraction;
  uint64_t seed;
} mkp_rownet_params;
/* Seeded synthetic inputs (SplitMix64 stream). */
MKP_API int mkp_generate_powerlaw(const mkp_powerlaw_params* params, mkp_host_hypergraph** out);
MKP_API int mkp_generate_rownet(const mkp_rownet_params* params, mkp_host_hypergraph** out);
MKP_API int mkp_host_hypergraph_info(const mkp_host_hypergraph* h, mkp_hypergraph_info* info);
MKP_API int mkp_host_hypergraph_copy(const mkp_host_hypergraph* h, uint64_t* net_offsets, uint32_t* pins,
                                     int32_t* vertex_weights, int32_t* net_weights);
/* Host hypergraph from CSR arrays (copied; vw / nw may be NULL for unit
 * weights), e.g. to assemble batch members. */
MKP_API int mkp_host_hypergraph_create(uint32_t n, uint32_t m, const uint64_t* net_offsets, const uint32_t* pins,
                                       const int32_t* vertex_weights, const int32_t* net_weights,
                                       mkp_host_hypergraph** out);
MKP_API int mkp_host_hypergraph_destroy(mkp_host_hypergraph* h);
MKP_API int mkp_hypergraph_create_from_host(mkp_context* ctx, const mkp_host_hypergraph* h,
                                            mkp_hypergraph** out);

/* ---------------------------------------------------------------- io -- */
/* hMetis reader (io/hypergraph_io.cpp:95-177, readHypergraphFile): nets with
 * fewer than two pins are dropped and counted in *removed_single_pin_nets. */
MKP_API int mkp_read_hgr(const char* path, mkp_host_hypergraph** out, uint64_t* removed_single_pin_nets);
MKP_API int mkp_write_hgr(const mkp_host_hypergraph* h, const char* path);
/* writePartitionFile (io/hypergraph_io.cpp:319-333): one block id per line. */
MKP_API int mkp_write_partition_file(const int32_t* parts, uint32_t num_vertices, const char* path);

/* ----------------------------------------------------------- batches --- */
/* The batched form (no counterpart in the reference, which partitions one
 * hypergraph per call): many independent member hypergraphs live in HBM as
 * one disjoint union and every round of every member runs in the same kernel
 * launches.  Members keep local vertex and block ids for all hashing and
 * tie-breaking, their own seeds, block weights and stop state, so each member
 * gets exactly the result it would get alone.  Vertices of member g are
 * numbered consecutively after those of members 0..g-1 in every array. */
typedef struct mkp_batch mkp_batch;
typedef struct mkp_batch_partition mkp_batch_partition;
typedef struct {
  double epsilon;                /* balance of every member (context.cpp:237-250) */
  int32_t max_rounds;            /* r-lp-maximum-iterations (5) */
  int32_t sub_rounds;
  uint32_t activation_threshold; /* r-lp-he-size-activation-threshold (100) */
  int32_t rebalancing;           /* r-lp-rebalancing (1) */
} mkp_batch_lp_params;
MKP_API int mkp_batch_create(mkp_context* ctx, uint32_t count, const mkp_host_hypergraph* const* members,
                             mkp_batch** out);
MKP_API int mkp_batch_destroy(mkp_batch* b);
MKP_API int mkp_batch_get_info(const mkp_batch* b, uint32_t* members, uint32_t* vertices, uint32_t* nets,
                               uint64_t* pins);
/* Borrowed handle of the union hypergraph. */
MKP_API int mkp_batch_hypergraph(mkp_batch* b, mkp_hypergraph** hg);
/* mkp_louvain on every member at once: each member runs all its levels and
 * rounds as it would alone, with its own seed (seeds: one per member, or NULL
 * for params->seed everywhere).  communities (host, all vertices, may be
 * NULL) and the union hypergraph's community ids receive member-local ids;
 * stats (may be NULL) has one entry per member.  The members must resolve to
 * the same edge weighting and either all or none be plain graphs. */
MKP_API int mkp_batch_louvain(mkp_batch* b, const mkp_louvain_params* params, const uint64_t* seeds,
                              int32_t* communities, mkp_louvain_stats* stats);
/* k (one value per member); all vertices start in block 0. */
MKP_API int mkp_batch_partition_create(mkp_batch* b, const int32_t* k, mkp_batch_partition** out);
MKP_API int mkp_batch_partition_destroy(mkp_batch_partition* bp);
/* parts: local block ids of all vertices (host, union order), then
 * initializePartition of every member. */
MKP_API int mkp_batch_partition_set_parts(mkp_batch_partition* bp, const int32_t* parts);
MKP_API int mkp_batch_partition_get_parts(const mkp_batch_partition* bp, int32_t* parts);
/* Device-to-device copy of the block ids of src (same batch and k) followed
 * by initializePartition of every member. */
MKP_API int mkp_batch_partition_copy_parts(mkp_batch_partition* dst, const mkp_batch_partition* src);
/* km1 of every member (host, one value per member). */
MKP_API int mkp_batch_partition_km1(mkp_batch_partition* bp, int64_t* km1);
/* LabelPropagationRefiner on every member at once; seeds and stats have one
 * entry per member. */
MKP_API int mkp_batch_lp_refine(mkp_batch_partition* bp, const mkp_batch_lp_params* params, const uint64_t* seeds,
                                mkp_lp_stats* stats);

/* ------------------------------------------------------------ serving --- */
/* The served form (no counterpart in the reference, which partitions one
 * hypergraph per call): a request loop that takes partition requests of any
 * size as they arrive.  `slots` requests run at once, each on a context of
 * its own (own stream, own device allocation cache), so their kernels overlap
 * on the device; a request is admitted, in arrival order, when a slot is free
 * and its device byte estimate fits the HBM budget beside the estimates of
 * the requests in flight and the allocation caches.  Its hypergraph and
 * partition state stay on the device from admission to completion.  A
 * request's result does not depend on what else is in flight: every request
 * runs mkp_partition_hypergraph with its own parameters and seed, sharing no
 * device state with the others. */
typedef struct mkp_server mkp_server;
typedef struct {
  int32_t device;
  int32_t slots;              /* requests in flight at once (1..64) */
  uint64_t hbm_budget_bytes;  /* admission budget; 0 = device memory less 4 GiB */
  int32_t spin_wait;          /* 0 (default): slot threads sleep while their stream runs; 1: they spin */
  int32_t ip_threads;         /* unused (initial partitioning runs on the device); kept for the layout */
  int32_t host_threads;       /* unused, as ip_threads */
} mkp_server_params;
typedef struct {
  int64_t km1;
  double imbalance;
  int32_t levels;
  int32_t slot;               /* slot that ran the request */
  double queue_ms;            /* submission to admission */
  double service_ms;          /* admission to completion */
  double latency_ms;          /* submission to completion */
  uint64_t estimate_bytes;    /* admission estimate */
  uint64_t peak_bytes;        /* most device bytes its slot held while it ran */
} mkp_request_result;
typedef struct {
  uint64_t submitted, completed, failed;
  int32_t max_in_flight;       /* most requests running at once */
  uint64_t max_reserved_bytes; /* most estimate bytes admitted at once */
  uint64_t budget_bytes;
  uint64_t trims;              /* allocation caches released for admission */
} mkp_server_stats;
/* Device bytes a request needs at most (hypergraph, community detection,
 * coarsening hierarchy, initial partitioning with 20 runs per bisection,
 * partition state and refinement work, rounded to the allocation cache's
 * power-of-two size classes). */
MKP_API int mkp_request_estimate(uint32_t num_vertices, uint32_t num_nets, uint64_t num_pins, uint32_t max_net_size,
                                 int32_t k, uint64_t* bytes);
MKP_API int mkp_server_create(const mkp_server_params* params, mkp_server** out);
/* Finishes the requests already submitted, then stops the slots. */
MKP_API int mkp_server_destroy(mkp_server* s);
/* Queues a request.  h stays owned by the caller and must stay alive until
 * mkp_server_wait has returned for the ticket; params are copied.  A request
 * whose estimate exceeds the whole budget is refused here. */
MKP_API int mkp_server_submit(mkp_server* s, const mkp_host_hypergraph* h, const mkp_partition_params* params,
                              uint64_t* ticket);
/* Blocks until the request is done and releases it; parts (host,
 * num_vertices) and result may be NULL.  A request that failed returns its
 * error here. */
MKP_API int mkp_server_wait(mkp_server* s, uint64_t ticket, int32_t* parts, mkp_request_result* result);
MKP_API int mkp_server_get_stats(mkp_server* s, mkp_server_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
