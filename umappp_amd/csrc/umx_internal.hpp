// Internal declarations shared by the HIP translation units of libumappp_amd.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace umx {

// Error carrying a umx_status code; the C ABI layer turns it into a return
// value plus a thread-local message.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define UMX_HIP(call)                                                                                    \
    do {                                                                                                 \
        hipError_t umx_e_ = (call);                                                                      \
        if (umx_e_ != hipSuccess)                                                                        \
            throw ::umx::Error(-3, std::string(#call) + ": " + hipGetErrorString(umx_e_));               \
    } while (0)

constexpr int kWave = 64;

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
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        release();
        if (count == 0) count = 1;
        UMX_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
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

// similarities_to_epochs.  Allocates the EpochData arrays.
template <typename T>
void run_to_epochs(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, int num_epochs, T nsr,
                   int64_t max_degree, DevBuf<int64_t>& out_rp, DevBuf<int32_t>& out_tg, DevBuf<T>& eps,
                   DevBuf<T>& nxt, DevBuf<T>& nxt_neg, hipStream_t s);

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

template <typename T>
struct LayoutView {
    int64_t n;
    int64_t nnz;
    const int64_t* rp;
    const int32_t* tg;
    const T* eps;
    T* nxt;
    T* nxt_neg;
    T* emb;
    int dim;
    int mode;
    T a, b, gamma, nsr;
    unsigned long long* tally;  // 3 counters per epoch
    // Work items of the production kernels: item m covers edges
    // [item_begin[m], item_begin[m] + item_len[m]) of row
    // item_row[m] & 0x7fffffff; the top bit marks rows split over several
    // items.  Items are ordered heaviest first.
    const int64_t* item_begin;
    const int32_t* item_len;
    const int32_t* item_row;
    int64_t n_items;
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
    // Optional update trace (parity mode): records of 3 + 3 * dim doubles
    // (kind, head, other, head before, other before, head move).
    double* trace;
    long long trace_cap;
    unsigned long long* trace_len;
};

template <typename T>
void launch_layout_epoch(const LayoutView<T>& v, const EpochArgs& ea, hipStream_t s);

template <typename T>
void launch_layout_serial(const LayoutView<T>& v, const EpochArgs& ea, const int32_t* samples, int64_t count,
                          unsigned long long* cursor, hipStream_t s);

// Work items of the production layout kernels: one per row, except that rows
// expected to fire more than 'fires_split' edges per epoch (sum of 1 / eps)
// or longer than 'edges_cap' are cut into items of about equal expected
// firings.  Items are sorted heaviest first.
template <typename T>
void build_work_items(int64_t n, const int64_t* rp, const T* eps, double fires_split, int edges_cap,
                      DevBuf<int64_t>& item_begin, DevBuf<int32_t>& item_len, DevBuf<int32_t>& item_row,
                      int64_t& n_items, hipStream_t s);

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
template <typename T>
void launch_layout_sell(const LayoutView<T>& v, const EpochArgs& ea, int tile_cols, hipStream_t s);

// Counter sampler key of one epoch (host side; the device derives per-edge
// and per-draw keys from it).
uint64_t counter_epoch_key(uint64_t seed, int epoch);

}  // namespace umx
