// Internal declarations shared by the HIP translation units of libumappp_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

namespace umx {

// Error carrying a umx_status code; the C ABI layer turns it into a return
// value plus a thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Thread-local message of the last failed C ABI call (umx_last_error).
void set_last_error(const std::string& msg);

// Runs f, turning exceptions into a umx_status (0 on success).
template <typename F>
int guarded(F&& f) {
    try {
        f();
        return 0;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return -3;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return -3;
    }
}

#define UMX_HIP(call)                                                                                    \
    do {                                                                                                 \
        hipError_t umx_e_ = (call);                                                                      \
        if (umx_e_ != hipSuccess)                                                                        \
            throw ::umx::Error(-3, std::string(#call) + ": " + hipGetErrorString(umx_e_));               \
    } while (0)

constexpr int kWave = 64;

// Device memory from the device's stream-ordered pool (hipMallocAsync), which
// keeps freed memory for reuse instead of returning it to the driver, so the
// many buffers of a graph build cost no driver allocation after the first
// build.  device_alloc returns memory that is ready on every stream;
// device_free may only be called once no queued work uses the memory.
void* device_alloc(size_t bytes);
void device_free(void* p);
// Sets the current device's pool to keep freed memory (idempotent); for
// code that calls hipMallocAsync itself.
void prepare_pool();

// Device buffer owned by a handle.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) device_free(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        p = static_cast<T*>(device_alloc(count * sizeof(T)));
        n = count;
    }
    void swap(DevBuf& o) {
        std::swap(p, o.p);
        std::swap(n, o.n);
    }
};

// ---- graph kernels (graph_kernels.hip) ----

// Lanes per row for the row-group kernels: the smallest power of two >= the
// longest row, between 4 and 64.
int lanes_for_degree(int64_t max_degree);

// Exclusive scan of 'count' int64 values into out[0..count] (out[count] =
// total).  Uses 'tmp' scratch that the function sizes itself.
void exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t count, DevBuf<int64_t>& tmp, hipStream_t s);

// Longest row of a CSR.
int64_t max_row_length(const int64_t* rp, int64_t n, DevBuf<int64_t>& tmp, hipStream_t s);

template <typename T>
void launch_similarities(int64_t n, const int64_t* rp, T* val, double lc, double bw, double mkds, int newton,
                         double* rho, double* sigma, int64_t max_degree, hipStream_t s);

// Symmetrise (combine_neighbor_sets).  Allocates the output CSR.
template <typename T>
void run_combine(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, T mix, int64_t max_degree,
                 DevBuf<int64_t>& out_rp, DevBuf<int32_t>& out_idx, DevBuf<T>& out_val, hipStream_t s);

// Instances of a batched graph or layout: independent problems stored one
// after the other in a single CSR.  Instance q owns observations
// [pt_off[q], pt_off[q + 1]); its rows' entries are contiguous too.  A plain
// problem is the case n_inst = 1 (of_row may then be null).
struct Instances {
    int n_inst = 1;
    const int64_t* pt_off = nullptr;  // device, n_inst + 1 entries
    const int32_t* of_row = nullptr;  // device, instance of each row (null when n_inst == 1)
};

// Turns instance-local neighbour indices into global ones (adds the row's
// instance offset); returns false if a local index lies outside its instance.
bool localize_to_global(int64_t n, const int64_t* rp, int32_t* idx, const Instances& inst, hipStream_t s);

// Instance of every row.
void fill_instance_of_row(int64_t n, const Instances& inst, int32_t* of_row, hipStream_t s);

// similarities_to_epochs.  Allocates the EpochData arrays.  The largest
// weight and the survivor limit are taken per instance; epochs_of_inst holds
// each instance's num_epochs (device; null = num_epochs for all).
template <typename T>
void run_to_epochs(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, int num_epochs,
                   const int32_t* epochs_of_inst, const Instances& inst, T nsr, int64_t max_degree,
                   DevBuf<int64_t>& out_rp, DevBuf<int32_t>& out_tg, DevBuf<T>& eps, DevBuf<T>& nxt,
                   DevBuf<T>& nxt_neg, hipStream_t s);

// ---- spectral initialisation (spectral.hip) ----

struct SpectralParams {
    double scale = 10.0;        // largest |coordinate| (Options::initialize_spectral_scale)
    bool jitter = false;
    double jitter_sd = 1e-4;
    uint64_t jitter_seed = 0;
    double tol = 1e-6;          // Ritz residual / Ritz value for convergence
    int start_dim = 32;         // first Krylov dimension (doubled until converged)
    int max_dim = 512;          // largest Krylov dimension
    uint64_t start_seed = 0x5eedULL;
};

struct SpectralReport {
    int64_t components = 0;
    int krylov_dim = 0;        // Lanczos vectors behind the final Ritz pairs
    int attempts = 0;          // convergence checks (the Krylov dimension doubles between them)
    int second_passes = 0;     // steps that needed a second Gram-Schmidt pass
    double worst_residual = 0;
    double bytes = 0;          // algorithmic HBM bytes of the kernels launched (each array counted once per read)
    std::vector<double> eigenvalues;  // of 2I - L, descending, num_dim + 1 of them
};

// Number of connected components of a symmetric CSR graph.
int64_t count_components(int64_t n, const int64_t* rp, const int32_t* idx, hipStream_t s);

// Spectral coordinates (point-major, n x d) of a symmetric weighted graph;
// false (out untouched) for several components, n <= d + 1, or no
// convergence.
template <typename T>
bool spectral_embedding(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, int max_degree, int d,
                        const SpectralParams& prm, T* out, SpectralReport* rep, hipStream_t s);

// Symmetric tridiagonal eigen-decomposition (host_api.cpp): diagonal d[k],
// off-diagonal e[k - 1]; eigenvalues into d (unsorted), eigenvectors into
// the columns of z (row-major k x k).  False if the iteration does not
// converge.
bool tridiagonal_eigen(int k, std::vector<double>& d, const std::vector<double>& e, std::vector<double>& z);

// ---- exact kNN (knn.hip) ----

// k nearest neighbours (self excluded, ties to the smaller index) of every
// observation of x (n x dim, point-major, device); k <= 64 and k < n.
template <typename T>
void knn_bruteforce(int64_t n, int dim, const T* x, int k, int32_t* idx, T* dist, hipStream_t s);

// ---- layout kernels (layout_kernels.hip) ----

struct EpochArgs {
    int epoch;        // epoch number n
    double alpha;     // learning rate of this epoch, already rounded to T by the host
    uint64_t ekey;    // counter-sampler key of this epoch
};

// Layout update modes: pull (default; each observation's owner applies all of
// its moves, no atomics) and push (tails moved with float atomics).
constexpr int kModePull = 0;
constexpr int kModePush = 1;

// Phases of a pull-mode epoch.  The observations are dealt to 'phases'
// groups (a power of two, 1 to 8) in blocks of 16 consecutive indices within
// their instance, block b to phase b mod phases; phase p's work items run in
// a launch after phase p - 1's, and read the observations of earlier phases
// as this epoch left them (the others as they were at its start), so a point
// sees fresh positions of part of its neighbours, as in the reference's loop,
// which moves one point after another.  Deterministic for a given number of
// phases.  Blocks of 16 points keep each 128-byte line of a two-dimensional
// float embedding in one phase, so a gather reads a line from one of the two
// buffers only and the set of lines an epoch reads is one embedding's worth,
// not two; dealing blocks in turn takes two instructions, which every gather
// of the epoch kernel evaluates, and spreads every run of indices (a cluster
// of points stored together, say) evenly over the phases.
constexpr int kPhaseBlockShift = 4;
__host__ __device__ inline int phase_of(uint64_t local_index, int phases) {
    return static_cast<int>(static_cast<uint32_t>(local_index >> kPhaseBlockShift) & static_cast<uint32_t>(phases - 1));
}
constexpr int kMaxPhases = 8;

// Per-instance constants of the production layout kernel.
struct InstConst {
    int64_t pt_off;    // first observation
    int64_t edge_off;  // first schedule entry (sampler keys use instance-local entry numbers)
    uint32_t n;        // observations
    int32_t epochs;    // num_epochs of this instance
    uint64_t seed;     // optimize_seed of this instance
};

template <typename T>
struct LayoutView {
    int64_t n;
    int64_t nnz;
    const int64_t* rp;
    const int32_t* tg;
    const T* eps;
    T* nxt;
    T* nxt_neg;
    // Embedding.  The push and serial kernels update 'emb' in place.  The
    // pull kernels are double-buffered: epoch e reads every observation from
    // buffer e & 1 (emb for even epochs, emb_alt for odd ones) and writes its
    // result to the other, so an epoch is a pure function of the state at
    // its start, whatever order the wavefronts run in.
    T* emb;
    T* emb_alt;
    // phases of the pull kernels: this launch runs phase 'phase' of
    // 'phases' (its items, chunks and split rows are the view's)
    int phases;
    int phase;
    int dim;
    int mode;
    T a, b, gamma, nsr;
    unsigned long long* tally;  // 3 counters per epoch
    // Per-epoch partial tallies of the parallel kernels: a ring of
    // kTallyRing entries of kTallySlots slots of kTallyStride counters (the
    // first 3 used), folded into tally[] and cleared by finish_layout_epoch
    // after a launch or by fold_layout_epochs after a run of launches.
    unsigned long long* tally_part;
    // Work items of the production kernels: item m covers edges
    // [item_begin[m], item_begin[m] + item_len[m]) of row
    // item_row[m] & 0x7fffffff; the top bit marks a piece of a split row
    // whose row was moved by an earlier phase of the epoch (it starts from
    // the output buffer).  Items are ordered longest first.
    const int64_t* item_begin;
    const int32_t* item_len;
    const int32_t* item_row;
    int64_t n_items;
    // item_aux[m]: kAuxPadding for padding items, kAuxWholeRow for an item
    // that is a whole row, else the piece's slot in split_delta (D values:
    // the piece's move of its head, summed onto the head by a fold after
    // the phase).  Split row r is observation split_row[r] & 0x7fffffff
    // (top bit: as in item_row), whose pieces in the phase hold slots
    // split_first[r] .. + split_count[r] - 1 in edge order.
    const int32_t* item_aux;
    T* split_delta;
    const int32_t* split_row;
    const int64_t* split_first;
    const int32_t* split_count;
    int64_t n_split_rows;
    // Sliced-ELL copy of the schedule for the production pull kernel (see
    // sell_kernels.hip): chunk c of 64 items has chunk_width[c] columns
    // starting at slot chunk_base[c]; slot (k, lane) = base + k * 64 + lane.
    const int32_t* chunk_width;
    const int64_t* chunk_base;
    int64_t n_chunks;
    const int32_t* sell_tg;
    const T* sell_eps;
    T* sell_nx;
    T* sell_nn;
    // Instances (sliced-ELL kernel): item_inst[m] is the instance of item m,
    // inst[q] its constants, alpha_tbl[q * alpha_stride + epoch] its learning
    // rate per epoch.  epoch_ctr[0] is the epoch of the next launch of a
    // captured block.
    const int32_t* item_inst;
    const InstConst* inst;
    uint32_t max_inst_n;  // observations of the largest instance (picks the sampler's arithmetic)
    const T* alpha_tbl;
    int alpha_stride;
    int* epoch_ctr;
    // Optional update trace (parity mode): records of 3 + 3 * dim doubles
    // (kind, head, other, head before, other before, head move).
    double* trace;
    long long trace_cap;
    unsigned long long* trace_len;
};

template <typename T>
void launch_layout_epoch(const LayoutView<T>& v, const EpochArgs& ea, hipStream_t s);

// Partial tallies: a wavefront adds its counts to slot (its index modulo
// kTallySlots); slots are kTallyStride counters (256 bytes) apart, so the
// atomics of an epoch spread over many memory channels instead of queueing
// on one address.
constexpr int kTallySlots = 64;
constexpr int kTallyStride = 32;

// The partial tallies of kTallyRing consecutive epochs are kept apart: epoch
// e adds into ring entry e % kTallyRing (kTallySlots * kTallyStride
// counters), so a run of epochs can be folded by one launch per kTallyRing
// epochs instead of one launch after every epoch.
constexpr int kTallyRing = 64;
constexpr int64_t tally_ring_offset(int epoch) {
    return static_cast<int64_t>(epoch % kTallyRing) * kTallySlots * kTallyStride;
}

// After each parallel epoch launch: adds the partial tallies into the tally
// of the epoch (epoch >= 0, else the epoch in epoch_ctr[0], which is then
// advanced by one) and clears them.  One wavefront; safe in graph capture.
void finish_layout_epoch(unsigned long long* tally_part, unsigned long long* tally, int epoch, int* epoch_ctr,
                         hipStream_t s);

// Folds the partial tallies of epochs [c, c + count), c = epoch_ctr[0] on
// the device, and advances the counter by count (count <= kTallyRing).  One
// launch after a captured block's epoch kernels; safe in graph capture.
void fold_layout_block(unsigned long long* tally_part, unsigned long long* tally, int count, int* epoch_ctr,
                       hipStream_t s);

// Folds the partial tallies of epochs [first, first + count) into their
// tallies and clears them: one launch, count <= kTallyRing.  Launched after
// the epochs' kernels on the same stream.
void fold_layout_epochs(unsigned long long* tally_part, unsigned long long* tally, int first, int count,
                        hipStream_t s);

template <typename T>
void launch_layout_serial(const LayoutView<T>& v, const EpochArgs& ea, const int32_t* samples, int64_t count,
                          unsigned long long* cursor, hipStream_t s);

// Work items of the production layout kernels: one per row, except that rows
// expected to fire more than 'fires_split' edges per epoch (sum of 1 / eps)
// or longer than 'edges_cap' are cut into items of about equal expected
// firings.  Items are sorted longest first.
//
// Placement for the sliced-ELL kernel: a workgroup of 'waves' wavefronts
// takes waves * 64 consecutive items.  With n_groups > 1, each instance
// belongs to group group_of_inst[q] and workgroup b only takes items of
// group b % n_groups (workgroups are dispatched to the XCDs round-robin, so
// a group's instances stay in one XCD's L2); groups are padded with empty
// items (length 0, row 0).  item_inst receives each item's instance.
struct ItemPlacement {
    int waves = 1;
    int n_groups = 1;
    const int32_t* group_of_inst = nullptr;  // device, n_inst entries (n_groups > 1 only)
};

//
// Split rows: item_aux (see LayoutView) and the list of split rows with the
// slots of their pieces, numbered in edge order.
constexpr int32_t kAuxPadding = -2;
constexpr int32_t kAuxWholeRow = -1;
struct SplitPlan {
    DevBuf<int32_t> row;
    DevBuf<int64_t> first;
    DevBuf<int32_t> count;
    int64_t rows = 0;   // split rows
    int64_t slots = 0;  // pieces of split rows
};

//
// Pieces: row_pieces (from count_row_pieces, once per build) holds every
// row's number of items.  Phase 'phase' of 'phases' gets the whole rows of
// its observations and, of a row in phase p0 cut into pieces, piece q when
// (p0 + q) mod phases == phase, so a split row's pieces run one after
// another through the epoch's phases; such a piece's row index has its top
// bit set when a piece of the same row runs in an earlier phase (it starts
// from the head as that phase's fold left it), and so has the row's entry in
// the split list.  The phase's items are padded to whole workgroups
// (waves * 64 items), so that a phase's items are whole chunks and whole
// workgroups.
template <typename T>
void count_row_pieces(int64_t n, const int64_t* rp, const T* eps, double fires_split, int edges_cap,
                      DevBuf<int64_t>& row_pieces, hipStream_t s);
template <typename T>
void build_work_items(int64_t n, const int64_t* rp, const T* eps, double fires_split, int edges_cap,
                      const int64_t* row_pieces, const Instances& inst, const ItemPlacement& place, int phases,
                      int phase, DevBuf<int64_t>& item_begin, DevBuf<int32_t>& item_len, DevBuf<int32_t>& item_row,
                      DevBuf<int32_t>& item_inst, DevBuf<int32_t>& item_aux, SplitPlan& split, int64_t& n_items,
                      hipStream_t s);

// Adds 'by' to every non-negative value of a[0..count) (split-piece slots
// of items built phase by phase).
void offset_nonnegative(int32_t* a, int64_t count, int64_t by, hipStream_t s);
void offset_all(int64_t* a, int64_t count, int64_t by, hipStream_t s);

// Whether every entry i -> t of an EpochData CSR has a reverse entry t -> i
// with the same epochs_per_sample (what pull mode needs); synchronises s.
template <typename T>
bool has_reverse_edges(int64_t n, const int64_t* rp, const int32_t* tg, const T* eps, hipStream_t s);

// After a phase of a double-buffered pull epoch: observation
// split_row[r] & 0x7fffffff of every split row r with pieces in the phase
// gets its state at the start of the phase (the epoch's input buffer, or its
// output buffer when the top bit of split_row[r] is set) plus the moves of
// those pieces, added in edge order, written to the epoch's output buffer.
// 'epoch' as for launch_layout_sell.
template <typename T>
void launch_fold_split(const LayoutView<T>& v, int epoch, hipStream_t s);

// Expected firings per epoch of each instance (sum of 1 / eps over its
// entries), written to fires[n_inst] (device).
template <typename T>
void instance_fires(const int64_t* rp, const T* eps, const Instances& inst, double* fires, hipStream_t s);

// Sliced-ELL schedule state (sell_kernels.hip).
template <typename T>
struct SellState {
    int64_t n_chunks = 0, n_slots = 0, E = 0;
    DevBuf<int32_t> width;
    DevBuf<int64_t> base;
    DevBuf<int32_t> tg;
    DevBuf<T> eps, nx, nn;
    DevBuf<int64_t> slot_of_edge;
};

template <typename T>
void build_sell(int64_t E, const int32_t* tg, const T* eps, const int64_t* item_begin, const int32_t* item_len,
                const int32_t* item_row, int64_t n_items, SellState<T>& st, hipStream_t s);
template <typename T>
void sync_csr_to_sell(const T* nx, const T* nn, SellState<T>& st, hipStream_t s);
template <typename T>
void sync_sell_to_csr(T* nx, T* nn, const SellState<T>& st, hipStream_t s);
// LDS budget of one workgroup of the sliced-ELL kernel, and what a
// configuration needs: waves * tile columns * 64 lanes * (target index +
// negative-sample count), 4 bytes each whatever the value type.
constexpr int kSellLdsMax = 65536;
inline size_t sell_lds_bytes(size_t /*value_bytes*/, int tile_cols, int waves) {
    return static_cast<size_t>(waves) * tile_cols * kWave * (4 + 4);
}

// One epoch of the sliced-ELL kernel.  epoch >= 0: that epoch.  epoch < 0:
// epoch v.epoch_ctr[0] + (-1 - epoch), the counter on the device being
// advanced after the block of such launches by fold_layout_block (or after
// one launch with epoch -1 by finish_layout_epoch), so a captured block of
// launches replays for any epochs.
template <typename T>
void launch_layout_sell(const LayoutView<T>& v, int epoch, int tile_cols, int waves, hipStream_t s);

// Sets epoch_ctr[0] = epoch with a kernel (safe inside graph capture).
void set_epoch_counter(int* epoch_ctr, int epoch, hipStream_t s);

// Counter sampler key of one epoch (host side; the device derives per-edge
// and per-draw keys from it).
uint64_t counter_epoch_key(uint64_t seed, int epoch);

}  // namespace umx
