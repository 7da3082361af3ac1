/*
 * umappp_amd.h — C ABI of the MI355X (gfx950) implementation of umappp's
 * graph-and-layout hot path.
 *
 * The reference (libscran/umappp 3.3.2) is a header-only C++ template library
 * with no ABI of its own; each entry point below replaces one reference
 * function or method, cited as file:line under include/umappp/ of the
 * reference.  The C++ drop-in headers under include/umappp/ of this
 * repository (namespace umappp: Options, initialize(), Status) are written on
 * top of this ABI.
 *
 * Conventions
 *  - Every function returns UMX_OK (0) or a negative umx_status; the message
 *    of the last failure on the calling thread is umx_last_error().
 *  - dtype selects the caller's Float_: UMX_F32 (float) or UMX_F64 (double).
 *    Neighbour indices are int32 (the reference's Index_ = int); row offsets
 *    are int64 (the reference's size_t cumulative_num_edges).
 *  - Pointers are host pointers unless an argument named 'device_ptrs' (or a
 *    function name ending in _device) says they live in GPU memory.
 *  - A handle owns its GPU memory and one HIP stream; handles are not
 *    thread-safe, separate handles are independent.  Calls block until their
 *    results are in the caller's memory, like the reference's calls.
 */
#ifndef UMAPPP_AMD_H
#define UMAPPP_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum umx_dtype { UMX_F32 = 0, UMX_F64 = 1 };

enum umx_status {
    UMX_OK = 0,
    UMX_ERR_ARGUMENT = -1,  /* bad size, option or pointer */
    UMX_ERR_GRAPH = -2,     /* neighbour index out of range, self or duplicate */
    UMX_ERR_DEVICE = -3,    /* a HIP call failed */
    UMX_ERR_STATE = -4,     /* call not valid in the handle's current state */
    UMX_ERR_OVERFLOW = -5,  /* a size or count does not fit its type */
    UMX_ERR_CAPACITY = -6   /* caller's output buffer too small */
};

/* Message for the last non-zero status returned on this thread. */
const char* umx_last_error(void);
/* ABI version (major * 10000 + minor * 100 + patch). */
int umx_version(void);

/* Returns to the driver the device buffers that the library keeps for reuse
 * on the current device (freed graph, layout and kNN buffers, up to 8 GiB per
 * device, reused by later allocations of about the same size).  No
 * counterpart in the reference (host memory there is the C++ allocator's). */
int umx_trim_memory(void);

/* ------------------------------------------------------------------ */
/* Fuzzy similarities                                                  */
/* ------------------------------------------------------------------ */

/*
 * neighbor_similarities — replaces umappp::neighbor_similarities
 * (neighbor_similarities.hpp:46-207) with NeighborSimilaritiesOptions
 * (:38-44).  The kNN lists are passed as CSR: row i holds entries
 * [row_ptr[i], row_ptr[i+1]) of 'val', sorted by increasing distance.  'val'
 * holds distances on entry and fuzzy weights on return (in place, as in the
 * reference).  The per-row solve runs in double precision whatever 'dtype' is.
 * use_newton = 0 selects the bisection-only variant (template argument
 * use_newton_ = false).  rho_out / sigma_out (double[n], may be NULL) receive
 * each row's rho and final sigma; NaN where the reference sets every weight to
 * 1 before computing them.  Host pointers.
 */
int umx_neighbor_similarities(int dtype, int64_t n, const int64_t* row_ptr, void* val,
                              double local_connectivity, double bandwidth, double min_k_dist_scale,
                              int use_newton, double* rho_out, double* sigma_out);

/*
 * combine_neighbor_sets — replaces umappp::combine_neighbor_sets
 * (combine_neighbor_sets.hpp:15-100).  Input is the directed weighted graph
 * (row_ptr, idx, val) with unique, non-self indices per row; output is the
 * symmetrised graph with every row sorted by index, written to out_row_ptr
 * (int64[n+1]), out_idx and out_val (capacity 'cap' entries; 2 * nnz always
 * suffices).  *out_nnz receives the output entry count (also on
 * UMX_ERR_CAPACITY).  Host pointers.
 */
int umx_combine_neighbor_sets(int dtype, int64_t n, const int64_t* row_ptr, const int32_t* idx,
                              const void* val, double mix_ratio, int64_t cap, int64_t* out_row_ptr,
                              int32_t* out_idx, void* out_val, int64_t* out_nnz);

/* ------------------------------------------------------------------ */
/* Graph handle: the kNN graph resident on the GPU                     */
/* ------------------------------------------------------------------ */

typedef struct umx_graph umx_graph;

/* Copies a CSR kNN graph (distances in 'val') to the GPU.  Replaces the
 * NeighborList argument of initialize() (initialize.hpp:69-70).
 * device_ptrs = 1 if row_ptr/idx/val are GPU pointers. */
int umx_graph_create(int dtype, int64_t n, const int64_t* row_ptr, const int32_t* idx, const void* val,
                     int device_ptrs, umx_graph** out);
/* neighbor_similarities on the resident graph (initialize.hpp:71-75). */
int umx_graph_neighbor_similarities(umx_graph* g, double local_connectivity, double bandwidth,
                                    double min_k_dist_scale, int use_newton);
/* combine_neighbor_sets on the resident graph (initialize.hpp:77). */
int umx_graph_combine(umx_graph* g, double mix_ratio);
int umx_graph_info(const umx_graph* g, int64_t* n, int64_t* nnz, int* dtype);
/* Copies the current graph to host memory (row_ptr int64[n+1], idx, val). */
int umx_graph_export(const umx_graph* g, int64_t* row_ptr, int32_t* idx, void* val);
/* rho and sigma of the last neighbor_similarities call (double[n] each). */
int umx_graph_export_rho_sigma(const umx_graph* g, double* rho, double* sigma);
void umx_graph_destroy(umx_graph* g);

/* Exact k-nearest neighbours by brute force on the GPU: the producer of the
 * NeighborList that initialize() takes (the reference gets it from knncolle,
 * initialize.hpp:138-177).  data: n observations of 'dim' values each
 * (point-major, the reference's "data_dim x num_obs" column-major layout),
 * dtype UMX_F32 or UMX_F64.  Writes idx_out (int32[n * k]) and dist_out
 * (Euclidean distances, the data's dtype) row by row in increasing distance,
 * the observation itself excluded, ties to the smaller index; squared
 * distances are summed in double in increasing dimension order.  1 <= k <=
 * 64, k < n.  device_ptrs = 1 if all three pointers are GPU memory. */
int umx_knn_bruteforce(int dtype, int64_t n, int dim, const void* data, int k, int32_t* idx_out, void* dist_out,
                       int device_ptrs);
/* A resident kNN graph straight from data (umx_knn_bruteforce without the
 * round trip through host memory); ready for umx_graph_neighbor_similarities.
 * Replaces initialize()'s data + knncolle::Builder overload
 * (initialize.hpp:165-177) together with the calls that follow. */
int umx_graph_create_knn(int dtype, int64_t n, int dim, const void* data, int k, int device_ptrs, umx_graph** out);

/* Connected components of the (symmetrised) resident graph; the reference's
 * has_multiple_components (spectral_init.hpp:142-168) is *count != 1. */
int umx_graph_components(const umx_graph* g, int64_t* count);
/* spectral_init (spectral_init.hpp:170-207) on the resident symmetrised
 * graph: the eigenvectors of the normalized Laplacian for its num_dim
 * smallest positive eigenvalues (Lanczos in double precision on 2I - L),
 * scaled so that the largest |coordinate| is 'scale', each vector's sign
 * fixed by making its largest-magnitude entry positive, plus N(0, jitter_sd)
 * noise when 'jitter' is set (a counter-based generator keyed by 'seed';
 * the reference's aarand normals are not reproduced).  Writes 'embedding'
 * (n * num_dim, point-major, the graph's dtype; device_ptr = 1 for GPU
 * memory) and sets *success = 1; *success = 0 leaves it untouched, as the
 * reference does when the graph has several components or the eigensolver
 * does not converge.  eigenvalues (may be NULL) receives the num_dim + 1
 * largest eigenvalues of 2I - L found. */
int umx_graph_spectral_init(const umx_graph* g, int num_dim, double scale, int jitter, double jitter_sd,
                            uint64_t seed, void* embedding, int device_ptr, int* success, double* eigenvalues);
/* Diagnostics of the last umx_graph_spectral_init on g (any pointer may be
 * NULL): the Lanczos vectors behind its Ritz pairs, the convergence checks
 * made (the Krylov dimension doubles between them), the Lanczos steps that
 * needed a second re-orthogonalisation pass, the worst relative Ritz
 * residual, and the algorithmic HBM bytes of the kernels it launched (for a
 * roofline).  No counterpart in the reference: spectral_init.hpp does not
 * surface irlba's own diagnostics. */
int umx_graph_spectral_report(const umx_graph* g, int* krylov_dim, int* attempts, int* second_passes,
                              double* worst_residual, double* bytes);

/* Batched graph (no counterpart in the reference, which runs one problem per
 * initialize() call): n_inst independent kNN graphs stored one after the
 * other.  Instance q owns observations [inst_offsets[q], inst_offsets[q+1])
 * (inst_offsets: host int64[n_inst + 1], inst_offsets[0] = 0); row_ptr/val
 * are the concatenated CSR over all observations, and idx holds
 * instance-local neighbour indices, exactly as each instance's own
 * NeighborList would (initialize.hpp:69-70).  The resident graph stores
 * global indices (local + the instance's offset); neighbor_similarities and
 * combine_neighbor_sets then act on every instance at once, and umx_layout_*
 * runs all instances in the same launches.  A neighbour index outside its
 * instance is UMX_ERR_GRAPH. */
int umx_graph_create_batch(int dtype, int32_t n_inst, const int64_t* inst_offsets, const int64_t* row_ptr,
                           const int32_t* idx, const void* val, int device_ptrs, umx_graph** out);
/* Number of instances and (if inst_offsets is not NULL) their offsets
 * (int64[n_inst + 1]).  A plain graph is one instance. */
int umx_graph_instances(const umx_graph* g, int32_t* n_inst, int64_t* inst_offsets);

/* ------------------------------------------------------------------ */
/* Layout handle: EpochData + SGD epochs                               */
/* ------------------------------------------------------------------ */

typedef struct umx_layout umx_layout;

/* similarities_to_epochs (optimize_layout.hpp:47-92) on a resident graph. */
int umx_layout_create(const umx_graph* g, int num_epochs, double negative_sample_rate, umx_layout** out);
/* Same from a symmetric weighted CSR (host or GPU pointers). */
int umx_layout_create_csr(int dtype, int64_t n, const int64_t* row_ptr, const int32_t* idx, const void* val,
                          int device_ptrs, int num_epochs, double negative_sample_rate, umx_layout** out);
/* similarities_to_epochs per instance of a batched graph: num_epochs[q]
 * (host int32[n_inst]) for instance q, whose largest weight and survivor
 * limit are its own (optimize_layout.hpp:49-63); umx_layout_create on a
 * batched graph gives every instance the same num_epochs.  Each instance has
 * its own epoch counter and learning-rate decay; umx_layout_run(L, limit)
 * advances every instance to min(limit, num_epochs[q]), and the handle's
 * total_epochs is the largest num_epochs[q].  Batched layouts run the
 * pull-mode sliced-ELL kernel only (num_dim <= 4). */
int umx_layout_create_batch(const umx_graph* g, const int32_t* num_epochs, double negative_sample_rate,
                            umx_layout** out);
/* Instances of a layout: count, observation offsets (int64[n_inst + 1]) and
 * num_epochs (int32[n_inst]); any output may be NULL. */
int umx_layout_instances(const umx_layout* L, int32_t* n_inst, int64_t* inst_offsets, int32_t* num_epochs);
/* Per-instance optimize_seed (host uint64[n_inst]); by default every
 * instance uses the seed of umx_layout_set_params.  An instance's negative
 * samples depend only on its seed and its own schedule entries, so a batched
 * instance draws the same samples as the same problem run alone. */
int umx_layout_set_instance_seeds(umx_layout* L, const uint64_t* seeds);
/* Sizes and progress: Status::num_observations / num_epochs / epoch
 * (Status.hpp:63-87). */
int umx_layout_info(const umx_layout* L, int64_t* n, int64_t* nnz, int* total_epochs, int* current_epoch);
/* EpochData export (optimize_layout.hpp:27-45): cumulative_num_edges,
 * edge_targets, epochs_per_sample, epoch_of_next_sample,
 * epoch_of_next_negative_sample.  Any pointer may be NULL. */
int umx_layout_export_epochs(const umx_layout* L, int64_t* row_ptr, int32_t* targets, void* epochs_per_sample,
                             void* epoch_of_next_sample, void* epoch_of_next_negative_sample);
/* Optimisation parameters (Status.hpp:103-113 passes these to
 * optimize_layout): embedding dimension, a, b, repulsion_strength (gamma),
 * learning_rate (initial alpha) and the negative-sampling seed. */
int umx_layout_set_params(umx_layout* L, int num_dim, double a, double b, double gamma, double learning_rate,
                          uint64_t seed);
/* Embedding in / out ("column-major num_dim x n", i.e. emb[i*num_dim + d],
 * Status.hpp:93).  device_ptr = 1 for a GPU pointer. */
int umx_layout_set_embedding(umx_layout* L, const void* embedding, int device_ptr);
int umx_layout_get_embedding(const umx_layout* L, void* embedding, int device_ptr);
/* Runs epochs [epoch(), epoch_limit) on the resident embedding with the
 * parallel (Hogwild) kernel — replaces optimize_layout /
 * optimize_layout_parallel as dispatched by Status::run (Status.hpp:101-128). */
int umx_layout_run(umx_layout* L, int epoch_limit);
/* Update scheme of umx_layout_run.  UMX_LAYOUT_PULL (default): every
 * observation's owner lane applies all of its moves — its head steps, the
 * tail steps of the reverse edges (a stored edge and its reverse fire in the
 * same epochs) and its negative samples.  Each epoch reads the embedding as
 * it was at the epoch's start and writes a new one (two buffers), with no
 * atomics, so results are bit-reproducible: the same for any launch shape,
 * and a run stopped and restarted at any epoch equals one run
 * (Options.hpp:203-205).  UMX_LAYOUT_PUSH: each head step also moves the
 * tail with float atomics, as the reference's loop does; not reproducible
 * bit for bit.  Pull mode needs every entry's reverse with the same weight,
 * as combine_neighbor_sets leaves it: a layout made from a graph with
 * one-way entries starts in push mode and refuses pull mode
 * (UMX_ERR_STATE); a batched one is refused (UMX_ERR_GRAPH). */
enum umx_layout_mode { UMX_LAYOUT_PULL = 0, UMX_LAYOUT_PUSH = 1 };
int umx_layout_set_mode(umx_layout* L, int mode);
int umx_layout_get_mode(const umx_layout* L, int* mode);
/* Phases of a pull-mode epoch (1, 2, 4 or 8; default 2).  The observations
 * are dealt to the phases in blocks of 16 consecutive indices (within their
 * instance), block b to phase b mod phases; the phases run one after
 * another, and a phase reads the observations of earlier phases as the
 * epoch left them and all others as they were at its start.  More phases
 * bring the epoch closer to the reference's loop, which moves one
 * observation after another, at the cost of one launch and one wait for the
 * slowest wavefront per phase.  Final UMAP loss against the reference's
 * serial layout: one phase 7-9% above it; two phases 1.4% (1M points) to
 * 3.5% (100k points) above, the 1M-point epoch 508 us; four phases within
 * 1%, 560 us.  Results are reproducible for a given number of phases.
 * get_phases also gives the work items' offsets of the phases (phases + 1
 * values). */
int umx_layout_set_phases(umx_layout* L, int phases);
int umx_layout_get_phases(umx_layout* L, int* phases, int64_t* item_offsets);
/* Launch shape the next umx_layout_run of the pull kernel (num_dim <= 4)
 * uses: LDS tile columns, wavefronts per workgroup, placement groups, and the
 * size of its sliced-ELL schedule copy (64-item chunks, slots).  Builds the
 * work items and the copy if they are not built yet.  Any pointer may be
 * NULL.  No counterpart in the reference. */
int umx_layout_launch_shape(umx_layout* L, int* tile_cols, int* waves, int* groups, int64_t* chunks,
                            int64_t* slots);
/* Load balance of umx_layout_run: one lane runs one work item, normally a
 * whole row (its steps stay sequential, as in the reference).  Rows expected
 * to fire more than fires_per_item edges per epoch (sum of
 * 1 / epochs_per_sample), or longer than edges_per_item, are cut into items
 * of about equal expected firings.  Piece q of a row whose observation is in
 * phase p0 runs in phase (p0 + q) mod phases, from the head as the earlier
 * phases of the epoch left it, and its move is added to the head in edge
 * order after its phase; a row of more pieces than phases sums the pieces
 * that share a phase from one start.  Items are run longest first.
 * Defaults 32.0 and 32: every phase of an epoch waits for its longest item,
 * so long items cost once per phase (a 100k-point epoch: 212 us with the cap
 * at 48 edges, 147 us at 32; the final UMAP cross-entropy stays within 1% of
 * the reference's serial layout at both). */
int umx_layout_set_work_split(umx_layout* L, double fires_per_item, int edges_per_item);
/* LDS tile of the pull kernel: schedule columns staged per wavefront at a
 * time (4, 8, 16 or 32), or 0 (the default) to choose from the number of
 * 64-item chunks: 32 up to 4096 chunks, 16 above (float64: 16 up to 8192,
 * 8 above), the widest that fits the 64 KiB LDS budget with the wavefronts
 * per workgroup. */
int umx_layout_set_tile(umx_layout* L, int tile_cols);
/* Launch shape of the production kernel: 'waves' wavefronts (1, 2 or 4)
 * per workgroup, and 'groups' placement groups: with groups > 1 the
 * instances are dealt to that many groups of about equal expected work and
 * workgroup b only runs items of group b % groups.  Workgroups are dispatched
 * to the XCDs round-robin, so with groups = the number of XCDs each XCD's L2
 * holds the embeddings of its own instances only.  groups = 0 (default)
 * picks the XCD count when the instances can be balanced across it, else 1. */
int umx_layout_set_placement(umx_layout* L, int waves, int groups);
/* Epochs per captured HIP graph: umx_layout_run records 'epochs' launches
 * of the production kernel once (launch q of the block runs epoch c + q,
 * c being a counter in device memory, and the one-wavefront kernel that
 * folds the block's tallies, once per 64 epochs, advances c, so one graph
 * serves every block of epochs) and replays it; 0 launches every epoch
 * directly.  Default 0; timing and tracing always launch directly. */
int umx_layout_set_graph_block(umx_layout* L, int epochs);
/* Autotuner for the production kernel's launch shape: from the current
 * state, runs a few warm-up epochs, then trial_epochs epochs for each
 * candidate (placement groups, waves per workgroup, LDS tile, by coordinate
 * descent), timing each with HIP events, restores the embedding, schedule and
 * tallies it started from, and keeps the fastest.  Results are cached per
 * process by problem shape (dtype, num_dim, n, entries, instances, epochs),
 * so later handles of the same shape take the cached choice at no cost.
 * Outputs (may be NULL) receive the chosen tile, waves and groups. */
int umx_layout_autotune(umx_layout* L, int trial_epochs, int* tile_cols, int* waves, int* groups);
/* The current work items (built on demand), in run order: first edge,
 * length, row (top bit set for a piece of a row cut into several items that
 * starts from the head as an earlier phase of the epoch left it).  *count
 * receives the number of items; up to cap are copied.  With waves > 1 or
 * groups > 1 the list holds padding items of length 0. */
int umx_layout_work_items(umx_layout* L, int64_t* begin, int32_t* length, int32_t* row, int64_t cap,
                          int64_t* count);
/* Deterministic one-thread replay of optimize_layout (optimize_layout.hpp:
 * 116-185) in the reference's edge order and mixed precision.  If 'samples' is
 * non-NULL, negative samples are taken from it in draw order (count entries);
 * otherwise from the counter sampler.  For parity tests. */
int umx_layout_run_serial(umx_layout* L, int epoch_limit, const int32_t* samples, int64_t count);
/* Per-epoch tallies since creation: out[3*e + {0,1,2}] = firing edges, negative
 * draws, self-hits of epoch e; cap = number of int64 slots. */
int umx_layout_tallies(const umx_layout* L, int64_t* out, int64_t cap);
/* Update trace (parity mode).  After umx_layout_set_trace(L, m) with m > 0,
 * every attractive or repulsive step of the next run / run_serial calls is
 * recorded as 3 + 3*num_dim doubles: kind, head, other, the head's and the
 * other point's coordinates before the step, and the head's move.  Kind 0 is
 * an attractive step of edge head->other (in push mode and in the serial
 * replay the tail moves by the negative of the head's move), kind 1 a
 * repulsive step, kind 2 (pull mode) the move that edge other->head's
 * attractive step gives its tail, 'head'.  The first value is
 * kind + 4 * phase, phase being the pull-mode phase of the launch that took
 * the step (0 in push mode and in the serial replay).
 * umx_layout_get_trace copies up to cap_records records (the first m taken in
 * order of recording), reports how many steps were recorded in total, and
 * resets the count.  m = 0 disables tracing. */
int umx_layout_set_trace(umx_layout* L, int64_t max_records);
int umx_layout_get_trace(umx_layout* L, double* out, int64_t cap_records, int64_t* total_records);
/* HIP-event timing of the epoch kernels (on the handle's stream): the
 * summed time and number of kernel launches since timing was switched on.
 * The sliced-ELL kernel is timed launch by launch (an epoch of P phases is P
 * launches; the folds of split rows between them are not counted); the other
 * kernels epoch by epoch. */
int umx_layout_set_timing(umx_layout* L, int on);
int umx_layout_timing(const umx_layout* L, double* kernel_ms, int64_t* launches);
/* Deep copy of a layout: schedule (epochs per sample, next epochs), current
 * epoch, tallies, parameters, settings and embedding, on the same device; the
 * copy and the source then run independently and, from the same state, give
 * the same results bit for bit.  Replaces the copy of umappp::Status, whose
 * members are values (Status.hpp:25-142, members :42-45); the drop-in
 * umappp::Status copy constructor calls it.  Work items, the sliced-ELL copy
 * and captured blocks are built again on the copy's first run. */
int umx_layout_clone(const umx_layout* src, umx_layout** out);
void umx_layout_destroy(umx_layout* L);

/* ------------------------------------------------------------------ */
/* UMAP cross-entropy loss (forward / backward of the autograd loss)    */
/* ------------------------------------------------------------------ */

/* No counterpart in the reference, which only applies per-edge SGD steps
 * (optimize_layout.hpp:145-176); this is the loss those steps descend.  Over
 * an edge list heads[e] -> tails[e] (int32[nnz]) of an embedding
 * (float[n * num_dim], point-major):
 *
 *   L = sum_e c_e [ log(1 + a s^b)
 *                   + gamma * sum_{p < m_e, k != head} (log(1 + a s_k^b) - log a - b log s_k) ]
 *
 * s = max(FLT_EPSILON, squared distance).  c_e = coef[e] (NULL: 1),
 * m_e = num_neg[e] (NULL: m).  Negative p of edge e is the layout's
 * counter-sampler draw for (seed, step, key, p), key = keys[e] (int64; NULL:
 * e), so the samples are recomputed, never stored; a draw equal to the head
 * is skipped.  ab = device double[2] {a, b}.  All pointers are GPU pointers;
 * work is queued on 'stream' (a hipStream_t; NULL = the default stream) and
 * the call returns once it is done.  An edge index outside [0, n) is
 * UMX_ERR_GRAPH.  Arithmetic is in double.  An edge with c_e = 0 contributes
 * nothing and its terms are not evaluated (a caller can pass a whole edge
 * list with a 0/1 mask as coef instead of compacting it).
 *
 * forward: *loss (device double) = L.
 * backward: grad_embedding (float[n * num_dim]) = grad_out * dL/dy and
 * grad_ab (double[2]) = grad_out * (dL/da, dL/db), grad_out a device double
 * (NULL: 1).  Both outputs are overwritten.  clip > 0 clips every term's
 * gradient with respect to its head to [-clip, clip] per coordinate before
 * the sum (per-sample clipping; the term's other point receives minus the
 * clipped value); clip = 4 gives the reference's clipped steps
 * (optimize_layout.hpp:105-110, 145-176).  clip = 0 is the exact gradient. */
int umx_loss_forward(int64_t n, int num_dim, const float* embedding, int64_t nnz, const int32_t* heads,
                     const int32_t* tails, const int64_t* keys, const float* coef, const int32_t* num_neg, int m,
                     const double* ab, double gamma, uint64_t seed, int step, double* loss, void* stream);
int umx_loss_backward(int64_t n, int num_dim, const float* embedding, int64_t nnz, const int32_t* heads,
                      const int32_t* tails, const int64_t* keys, const float* coef, const int32_t* num_neg, int m,
                      const double* ab, double gamma, uint64_t seed, int step, double clip,
                      const double* grad_out, float* grad_embedding, double* grad_ab, void* stream);

/* umx_loss_backward without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving or training loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and errors as umx_loss_backward; its
 * working memory (the double gradient accumulator and the index-error flag)
 * is a fixed array in the library's device memory, held by each GPU the
 * library runs on until the process ends (32 MiB), so n * num_dim may be at
 * most 4194304 (UMX_ERR_CAPACITY beyond).  Every call on one GPU shares that
 * array: calls that overlap on one GPU (other threads or streams) race, and
 * must be serialised by the caller.  Like umx_loss_backward it waits for its
 * stream before returning (to report index errors). */
int umx_loss_backward_noalloc(int64_t n, int num_dim, const float* embedding, int64_t nnz,
                              const int32_t* heads, const int32_t* tails, const int64_t* keys,
                              const float* coef, const int32_t* num_neg, int m, const double* ab,
                              double gamma, uint64_t seed, int step, double clip, const double* grad_out,
                              float* grad_embedding, double* grad_ab, void* stream);

/* ------------------------------------------------------------------ */
/* Host helpers (no GPU work)                                          */
/* ------------------------------------------------------------------ */

/* find_ab (find_ab.hpp:44-168): a and b of 1 / (1 + a x^(2b)). */
int umx_find_ab(double spread, double min_dist, double* a, double* b);
/* choose_num_epochs (initialize.hpp:27-47); num_epochs < 0 means unset. */
int umx_choose_num_epochs(int num_epochs, int64_t size);
/* Symmetric tridiagonal eigen-decomposition used by umx_graph_spectral_init
 * (host code): diagonal d[k] (overwritten with the eigenvalues, unsorted),
 * off-diagonal e[k - 1], eigenvectors into the columns of z (row-major
 * k x k).  UMX_ERR_STATE if the QR iteration does not converge. */
int umx_tridiagonal_eigen(int k, double* d, const double* e, double* z);

#ifdef __cplusplus
}
#endif

#endif /* UMAPPP_AMD_H */
