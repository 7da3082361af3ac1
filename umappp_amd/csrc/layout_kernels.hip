// Layout (SGD epoch) kernels for gfx950 — optimize_layout.hpp:116-185.
//
// k_layout_rows: one lane per head observation (rows of the head-sorted CSR).
//   The head's coordinates stay in registers while the lane walks its edges in
//   the reference's order, so each attractive / repulsive step sees the
//   head's previous steps exactly as the serial code does.  Tail moves are
//   applied with float atomics (no update is lost); the head's accumulated
//   move is added back atomically at the end of the row, so moves that other
//   lanes make to this observation while it is a head are kept as well.
//   Reads of other observations are Hogwild: they may be up to one epoch stale.
// k_layout_serial: a single lane replays the reference's loop in its exact
//   order and mixed precision (double denominators); the parity mode.
//
// Negative samples come from a counter-based sampler keyed by
// (seed, epoch, edge, draw), so no sampler state is stored or shared.
// Schedule updates (epoch_of_next_sample, epoch_of_next_negative_sample,
// num_neg) are IEEE single operations, compiled with -ffp-contract=off, and
// match the reference bit for bit.
#include "umx_internal.hpp"
#include "layout_device.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <algorithm>
#include <cmath>
#include <vector>

namespace umx {

namespace {

using namespace umx::dev;

template <typename T, int D, bool TRACE>
__global__ void __launch_bounds__(kBlock) k_layout_rows(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    if (m < v.n_items && v.item_len[m] > 0) {  // padding items have length 0
        const int64_t i = static_cast<int64_t>(static_cast<uint32_t>(v.item_row[m]) & 0x7fffffffu);
        const int64_t s = v.item_begin[m], e = s + v.item_len[m];
        const Vec<T, D> h0 = load_point<T, D>(v.emb, i);
        Vec<T, D> h = h0;
        const T two_ab = -2 * v.a * v.b;
        const T two_gb = 2 * v.gamma * v.b;
        const uint32_t n32 = static_cast<uint32_t>(v.n);
        for (int64_t j = s; j < e; ++j) {
            const T nx = v.nxt[j];
            if (nx > ep) continue;
            ++fires;
            const int32_t t = v.tg[j];
            const Vec<T, D> r = load_point<T, D>(v.emb, t);
            T d2 = 0;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const T q = h.c[d] - r.c[d];
                d2 += q * q;
            }
            d2 = d2 > eps_of<T>() ? d2 : eps_of<T>();
            const T pd2b = fast_pow(d2, v.b);
            const T gc = prod_div(two_ab * pd2b, d2 * (v.a * pd2b + 1));
            double* rec = TRACE ? trace_slot(v, 0, i, t) : nullptr;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                const T g = alpha * clip4(gc * (h.c[d] - r.c[d]));
                if (TRACE && rec) {
                    rec[3 + d] = h.c[d];
                    rec[3 + D + d] = r.c[d];
                    rec[3 + 2 * D + d] = g;
                }
                h.c[d] += g;
                add_to(v.emb + static_cast<int64_t>(t) * D + d, -g);
            }

            const T ej_eps = v.eps[j];
            const T epns = ej_eps / v.nsr;
            const T nn = v.nxt_neg[j];
            const int nneg = static_cast<int>((ep - nn) / epns);
            const uint64_t ej = edge_key(ec.ekey, static_cast<uint64_t>(j));
            for (int p = 0; p < nneg; ++p) {
                const int32_t k = draw_index(ej, static_cast<uint32_t>(p), n32);
                ++draws;
                if (k == i) {
                    ++self_hits;
                    continue;
                }
                const Vec<T, D> q = load_point<T, D>(v.emb, k);
                T e2 = 0;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const T z = h.c[d] - q.c[d];
                    e2 += z * z;
                }
                e2 = e2 > eps_of<T>() ? e2 : eps_of<T>();
                const T rc = prod_div(two_gb, (static_cast<T>(0.001) + e2) * (v.a * fast_pow(e2, v.b) + 1));
                double* rec = TRACE ? trace_slot(v, 1, i, k) : nullptr;
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    const T g = alpha * clip4(rc * (h.c[d] - q.c[d]));
                    if (TRACE && rec) {
                        rec[3 + d] = h.c[d];
                        rec[3 + D + d] = q.c[d];
                        rec[3 + 2 * D + d] = g;
                    }
                    h.c[d] += g;
                }
            }
            v.nxt[j] = nx + ej_eps;
            v.nxt_neg[j] = nn + static_cast<T>(nneg) * epns;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) add_to(v.emb + i * D + d, h.c[d] - h0.c[d]);
    }
    tally_add(v.tally_part, m / kWave, fires, draws, self_hits);
}

// Any embedding dimension: the head is updated in place with atomics after
// every step (no register copy), the rest as above.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_layout_rows_generic(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    if (m < v.n_items && v.item_len[m] > 0) {  // padding items have length 0
        const int D = v.dim;
        const int64_t i = static_cast<int64_t>(static_cast<uint32_t>(v.item_row[m]) & 0x7fffffffu);
        const int64_t s = v.item_begin[m], e = s + v.item_len[m];
        T* h = v.emb + i * D;
        const uint32_t n32 = static_cast<uint32_t>(v.n);
        for (int64_t j = s; j < e; ++j) {
            const T nx = v.nxt[j];
            if (nx > ep) continue;
            ++fires;
            T* r = v.emb + static_cast<int64_t>(v.tg[j]) * D;
            T d2 = 0;
            for (int d = 0; d < D; ++d) {
                const T q = h[d] - r[d];
                d2 += q * q;
            }
            d2 = d2 > eps_of<T>() ? d2 : eps_of<T>();
            const T pd2b = fast_pow(d2, v.b);
            const T gc = (-2 * v.a * v.b * pd2b) / (d2 * (v.a * pd2b + 1));
            for (int d = 0; d < D; ++d) {
                const T g = alpha * clip4(gc * (h[d] - r[d]));
                add_to(h + d, g);
                add_to(r + d, -g);
            }
            const T epns = v.eps[j] / v.nsr;
            const T nn = v.nxt_neg[j];
            const int nneg = static_cast<int>((ep - nn) / epns);
            const uint64_t ej = edge_key(ec.ekey, static_cast<uint64_t>(j));
            for (int p = 0; p < nneg; ++p) {
                const int32_t k = draw_index(ej, static_cast<uint32_t>(p), n32);
                ++draws;
                if (k == i) {
                    ++self_hits;
                    continue;
                }
                const T* q = v.emb + static_cast<int64_t>(k) * D;
                T e2 = 0;
                for (int d = 0; d < D; ++d) {
                    const T z = h[d] - q[d];
                    e2 += z * z;
                }
                e2 = e2 > eps_of<T>() ? e2 : eps_of<T>();
                const T rc = 2 * v.gamma * v.b / ((static_cast<T>(0.001) + e2) * (v.a * fast_pow(e2, v.b) + 1));
                for (int d = 0; d < D; ++d) add_to(h + d, alpha * clip4(rc * (h[d] - q[d])));
            }
            v.nxt[j] = nx + v.eps[j];
            v.nxt_neg[j] = nn + static_cast<T>(nneg) * epns;
        }
    }
    tally_add(v.tally_part, m / kWave, fires, draws, self_hits);
}

// Serial replay in the reference's order and precision (optimize_layout.hpp:
// 131-181): numerators in T, denominators in double (the 1.0 and 0.001
// literals promote them), results rounded back to T.
template <typename T>
__global__ void k_layout_serial(LayoutView<T> v, EpochConst ec, T ep, T alpha, const int32_t* samples, int64_t count,
                                unsigned long long* cursor, int* overrun) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const int D = v.dim;
    const uint32_t n32 = static_cast<uint32_t>(v.n);
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    unsigned long long cur = samples ? *cursor : 0;
    for (int64_t i = 0; i < v.n; ++i) {
        T* left = v.emb + i * D;
        for (int64_t j = v.rp[i]; j < v.rp[i + 1]; ++j) {
            if (v.nxt[j] > ep) continue;
            ++fires;
            T* right = v.emb + static_cast<int64_t>(v.tg[j]) * D;
            T d2 = 0;
            for (int d = 0; d < D; ++d) {
                const T q = left[d] - right[d];
                d2 += q * q;
            }
            d2 = d2 > eps_of<T>() ? d2 : eps_of<T>();
            const T pd2b = pow(d2, v.b);
            const T gc = static_cast<T>((-2 * v.a * v.b * pd2b) /
                                        (static_cast<double>(d2) * (static_cast<double>(v.a * pd2b) + 1.0)));
            double* rec = trace_slot(v, 0, i, v.tg[j]);
            for (int d = 0; d < D; ++d) {
                const T g = alpha * clip4(gc * (left[d] - right[d]));
                if (rec) {
                    rec[3 + d] = left[d];
                    rec[3 + D + d] = right[d];
                    rec[3 + 2 * D + d] = g;
                }
                left[d] += g;
                right[d] -= g;
            }
            const T epns = v.eps[j] / v.nsr;
            const int nneg = static_cast<int>((ep - v.nxt_neg[j]) / epns);
            const uint64_t ej = edge_key(ec.ekey, static_cast<uint64_t>(j));
            for (int p = 0; p < nneg; ++p) {
                int32_t k;
                if (samples) {
                    if (cur < static_cast<unsigned long long>(count)) {
                        k = samples[cur];
                    } else {
                        k = 0;
                        *overrun = 1;
                    }
                    ++cur;
                    if (k < 0 || k >= v.n) {
                        *overrun = 1;
                        k = 0;
                    }
                } else {
                    k = draw_index(ej, static_cast<uint32_t>(p), n32);
                }
                ++draws;
                if (k == i) {
                    ++self_hits;
                    continue;
                }
                const T* q = v.emb + static_cast<int64_t>(k) * D;
                T e2 = 0;
                for (int d = 0; d < D; ++d) {
                    const T z = left[d] - q[d];
                    e2 += z * z;
                }
                e2 = e2 > eps_of<T>() ? e2 : eps_of<T>();
                const T rc = static_cast<T>(static_cast<double>(2 * v.gamma * v.b) /
                                            ((0.001 + static_cast<double>(e2)) *
                                             (static_cast<double>(v.a * pow(e2, v.b)) + 1.0)));
                double* rec = trace_slot(v, 1, i, k);
                for (int d = 0; d < D; ++d) {
                    const T g = alpha * clip4(rc * (left[d] - q[d]));
                    if (rec) {
                        rec[3 + d] = left[d];
                        rec[3 + D + d] = q[d];
                        rec[3 + 2 * D + d] = g;
                    }
                    left[d] += g;
                }
            }
            v.nxt[j] += v.eps[j];
            v.nxt_neg[j] += static_cast<T>(nneg) * epns;
        }
    }
    if (samples) *cursor = cur;
    unsigned long long* t = v.tally + 3 * static_cast<int64_t>(ec.epoch);
    t[0] += fires;
    t[1] += draws;
    t[2] += self_hits;
}

// Pull mode for embeddings of more than four dimensions (the sliced-ELL
// kernel covers one to four): the owner lane of observation i applies every
// move of i.  A stored edge i->t fires in exactly the same epochs as its
// reverse t->i (the fuzzy union makes weights, hence epochs_per_sample,
// bit-symmetric), so for each firing edge the lane applies i's head step g of
// i->t, then the tail step that t->i gives i, then i's negative samples, in
// the reference's order.  The tail step is the negative of t's head move of
// t->i, -alpha*clip(c*(y_t' - y_i')) = alpha*clip(c*(y_i' - y_t')), taken at
// the positions the pair's first event leaves, y_i' = y_i + g and
// y_t' = y_t - g: in the reference the event i->t moves the tail by -g, and
// the reverse event sees that.
//
// Double-buffered, in phases (see phase_of): a point of an earlier phase is
// read from the epoch's output buffer (this epoch's position), any other
// from its input buffer.  i's working copy is its slot in the output buffer
// (a whole row) or its piece's slot in split_delta (a piece of a split row,
// turned into the piece's move at the end and summed onto the head by
// launch_fold_split).  No lane reads what another lane of the launch writes,
// so the epoch is deterministic.  Plain layouts only (local index = index).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_layout_pull_generic(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    const int32_t aux = m < v.n_items ? v.item_aux[m] : kAuxPadding;
    if (aux != kAuxPadding) {
        const int D = v.dim;
        const T* src = (ec.epoch & 1) ? v.emb_alt : v.emb;
        T* dst = (ec.epoch & 1) ? v.emb : v.emb_alt;
        auto at = [&](int64_t j) -> const T* {
            return (phase_of(static_cast<uint64_t>(j), v.phases) < v.phase ? dst : src) + j * D;
        };
        const int32_t tagged = v.item_row[m];
        const int64_t i = static_cast<int64_t>(static_cast<uint32_t>(tagged) & 0x7fffffffu);
        const int64_t s = v.item_begin[m], e = s + v.item_len[m];
        T* h = aux >= 0 ? v.split_delta + static_cast<int64_t>(aux) * D : dst + i * D;
        // a piece whose row was moved by an earlier phase of this epoch
        // starts where that phase's fold left the head
        const T* base = tagged < 0 ? dst : src;
        for (int d = 0; d < D; ++d) h[d] = base[i * D + d];
        const uint32_t n32 = static_cast<uint32_t>(v.n);
        for (int64_t j = s; j < e; ++j) {
            const T nx = v.nxt[j];
            if (nx > ep) continue;
            ++fires;
            const T* r = at(v.tg[j]);
            // head step g (coefficient c1), then the reverse edge's step at
            // (y_i + g, y_t - g) (coefficient c2); g is recomputed per
            // coordinate rather than stored, so any num_dim works
            T d2 = 0;
            for (int d = 0; d < D; ++d) {
                const T q = h[d] - r[d];
                d2 += q * q;
            }
            d2 = d2 > eps_of<T>() ? d2 : eps_of<T>();
            T pd2b = fast_pow(d2, v.b);
            const T c1 = (-2 * v.a * v.b * pd2b) / (d2 * (v.a * pd2b + 1));
            T e2 = 0;
            for (int d = 0; d < D; ++d) {
                const T g = alpha * clip4(c1 * (h[d] - r[d]));
                const T q = (h[d] + g) - (r[d] - g);
                e2 += q * q;
            }
            e2 = e2 > eps_of<T>() ? e2 : eps_of<T>();
            pd2b = fast_pow(e2, v.b);
            const T c2 = (-2 * v.a * v.b * pd2b) / (e2 * (v.a * pd2b + 1));
            for (int d = 0; d < D; ++d) {
                const T g = alpha * clip4(c1 * (h[d] - r[d]));
                const T h1 = h[d] + g, r1 = r[d] - g;
                h[d] = h1 + alpha * clip4(c2 * (h1 - r1));
            }
            const T epns = v.eps[j] / v.nsr;
            const T nn = v.nxt_neg[j];
            const int nneg = static_cast<int>((ep - nn) / epns);
            const uint64_t ej = edge_key(ec.ekey, static_cast<uint64_t>(j));
            for (int p = 0; p < nneg; ++p) {
                const int32_t k = draw_index(ej, static_cast<uint32_t>(p), n32);
                ++draws;
                if (k == i) {
                    ++self_hits;
                    continue;
                }
                const T* q = at(k);
                T f2 = 0;
                for (int d = 0; d < D; ++d) {
                    const T z = h[d] - q[d];
                    f2 += z * z;
                }
                f2 = f2 > eps_of<T>() ? f2 : eps_of<T>();
                const T rc = 2 * v.gamma * v.b / ((static_cast<T>(0.001) + f2) * (v.a * fast_pow(f2, v.b) + 1));
                for (int d = 0; d < D; ++d) h[d] += alpha * clip4(rc * (h[d] - q[d]));
            }
            v.nxt[j] = nx + v.eps[j];
            v.nxt_neg[j] = nn + static_cast<T>(nneg) * epns;
        }
        if (aux >= 0)
            for (int d = 0; d < D; ++d) h[d] -= base[i * D + d];
    }
    tally_add(v.tally_part, m / kWave, fires, draws, self_hits);
}

// The fold of split rows after a phase of a double-buffered pull epoch (one
// lane per split row with pieces in the phase; see launch_fold_split): the
// head's new position is its position at the phase's start (the input
// buffer, or the output buffer when a fold of an earlier phase of the epoch
// already moved it: the top bit of split_row) plus the phase's piece moves
// in edge order.
template <typename T, int D>
__global__ void __launch_bounds__(kBlock) k_fold_split(LayoutView<T> v, int epoch_arg) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= v.n_split_rows) return;
    const int epn = epoch_arg >= 0 ? epoch_arg : v.epoch_ctr[0] + (-1 - epoch_arg);
    const T* src = (epn & 1) ? v.emb_alt : v.emb;
    T* dst = (epn & 1) ? v.emb : v.emb_alt;
    const int32_t tagged = v.split_row[r];
    const int64_t i = static_cast<int64_t>(static_cast<uint32_t>(tagged) & 0x7fffffffu);
    const int64_t first = v.split_first[r];
    const int cnt = v.split_count[r];
    Vec<T, D> h = load_point<T, D>(tagged < 0 ? dst : src, i);
    for (int q = 0; q < cnt; ++q) {
        const Vec<T, D> d = load_point<T, D>(v.split_delta, first + q);
#pragma unroll
        for (int c = 0; c < D; ++c) h.c[c] += d.c[c];
    }
    store_point<T, D>(dst, i, h);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fold_split_generic(LayoutView<T> v, int epoch_arg) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (r >= v.n_split_rows) return;
    const int D = v.dim;
    const int epn = epoch_arg >= 0 ? epoch_arg : v.epoch_ctr[0] + (-1 - epoch_arg);
    const T* src = (epn & 1) ? v.emb_alt : v.emb;
    T* dst = (epn & 1) ? v.emb : v.emb_alt;
    const int32_t tagged = v.split_row[r];
    const int64_t i = static_cast<int64_t>(static_cast<uint32_t>(tagged) & 0x7fffffffu);
    const int64_t first = v.split_first[r];
    const int cnt = v.split_count[r];
    const T* base = tagged < 0 ? dst : src;
    for (int c = 0; c < D; ++c) {
        T h = base[i * D + c];
        for (int q = 0; q < cnt; ++q) h += v.split_delta[(first + q) * D + c];
        dst[i * D + c] = h;
    }
}

// Work items (see build_work_items).  One thread per row: the expected
// firings per epoch f = sum of 1 / eps over the row; a row with f above the
// split threshold (or longer than the edge cap) is cut into about
// c = ceil(f / threshold) items, a cut falling where the running sum of 1/eps
// would pass q * f / c (q = 1..c-1) or where the edge cap is reached.  The
// same walk counts the items (WRITE = false) and then writes them.
//
// Both walks take kSplitBatch entries per trip: their loads (clamped to the
// row's last entry, so unconditional) and reciprocals come first, then the
// running sums in entry order, so the sums round exactly as one entry at a
// time while a hub row of thousands of entries waits for one load round trip
// per batch instead of one per entry.
constexpr int kSplitBatch = 16;

template <typename T>
__device__ inline void split_weights(const T* __restrict__ eps, int64_t j0, int64_t r1, double (&w)[kSplitBatch]) {
    T e[kSplitBatch];
#pragma unroll
    for (int u = 0; u < kSplitBatch; ++u) e[u] = eps[j0 + u < r1 ? j0 + u : r1 - 1];
    // keeps every load ahead of the first conversion (otherwise the
    // scheduler interleaves them and waits for each load in turn)
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < kSplitBatch; ++u) w[u] = 1.0 / static_cast<double>(e[u]);
}

// Where a row's pieces run: piece q of a row whose observation is in phase
// p0 runs in phase (p0 + q) mod phases, so the pieces of a split row run one
// after another through the epoch's phases (only a row of more pieces than
// phases has several in one phase).  A piece starts from the head as the
// epoch's earlier phases left it when a piece of its row ran in an earlier
// phase (tagged in the top bit of its row index), else from the epoch's
// input.
struct Stagger {
    int p0;         // the row's phase
    int phases;     // a power of two
    int phase;      // the phase whose pieces are written
    int64_t total;  // the row's pieces
};

__host__ __device__ inline int64_t pieces_in_phase(int64_t total, int p0, int phases, int phase) {
    const int64_t r = static_cast<int64_t>((phase - p0) & (phases - 1));
    return r < total ? (total - 1 - r) / phases + 1 : 0;
}

// Whether the pieces of a row (phase p0, 'total' pieces) that run in 'phase'
// start from the epoch's output buffer: some piece ran in an earlier phase.
__host__ __device__ inline bool starts_from_output(int64_t total, int p0, int phases, int phase) {
    const int first = p0 + total - 1 >= phases ? 0 : p0;  // the earliest phase with a piece
    return phase > first;
}

template <typename T, bool WRITE>
__device__ inline int64_t split_row(const int64_t* __restrict__ rp, const T* __restrict__ eps, int64_t i,
                                    double fires_split, int edges_cap, const Stagger& st, int64_t m,
                                    int64_t* __restrict__ ibeg, int32_t* __restrict__ ilen,
                                    int32_t* __restrict__ irow) {
    const int64_t r0 = rp[i], r1 = rp[i + 1];
    if (r1 == r0) return 0;
    const bool from_out = WRITE && starts_from_output(st.total, st.p0, st.phases, st.phase);
    const int32_t tag = static_cast<int32_t>(static_cast<uint32_t>(i) | (from_out ? 0x80000000u : 0u));
    // piece q (in edge order) is written when it runs in st.phase
    int64_t q = 0, written = 0;
    auto piece = [&](int64_t start, int64_t end) {
        if (WRITE && ((st.p0 + q) & (st.phases - 1)) == st.phase) {
            ibeg[m + written] = start;
            ilen[m + written] = static_cast<int32_t>(end - start);
            irow[m + written] = tag;
            ++written;
        }
        ++q;
    };
    // every epochs_per_sample is at least 1, so a row of at most
    // fires_split entries expects at most that many firings: one piece
    if (r1 - r0 <= edges_cap && static_cast<double>(r1 - r0) <= fires_split) {
        piece(r0, r1);
        return WRITE ? written : q;
    }
    double f = 0;
    for (int64_t j0 = r0; j0 < r1; j0 += kSplitBatch) {
        double w[kSplitBatch];
        split_weights(eps, j0, r1, w);
#pragma unroll
        for (int u = 0; u < kSplitBatch; ++u)
            if (j0 + u < r1) f += w[u];
    }
    const double c = f > fires_split ? ceil(f / fires_split) : 1.0;
    if (c == 1.0 && r1 - r0 <= edges_cap) {
        // no cut can occur below (q < c never holds; no piece reaches the cap)
        piece(r0, r1);
        return WRITE ? written : q;
    }
    int64_t start = r0;
    double acc = 0, cq = 1;
    for (int64_t j0 = r0; j0 < r1; j0 += kSplitBatch) {
        double wb[kSplitBatch];
        split_weights(eps, j0, r1, wb);
#pragma unroll
        for (int u = 0; u < kSplitBatch; ++u) {
            const int64_t j = j0 + u;
            if (j < r1) {  // (a guard, not a break: the loop unrolls fully)
                const double w = wb[u];
                const bool cut = j > start && ((cq < c && acc + w > cq * f / c) || j - start >= edges_cap);
                if (cut) {
                    piece(start, j);
                    if (acc + w > cq * f / c) cq += 1;
                    start = j;
                }
                acc += w;
            }
        }
    }
    piece(start, r1);
    return WRITE ? written : q;
}

// The phase of row i (its index within its instance decides).
__device__ inline int row_phase(int64_t i, const int64_t* __restrict__ pt_off, const int32_t* __restrict__ of_row,
                                int phases) {
    if (phases <= 1) return 0;
    const int64_t local = i - (of_row ? pt_off[of_row[i]] : pt_off[0]);
    return phase_of(static_cast<uint64_t>(local), phases);
}

// Pieces of every row (phase-independent), once per build.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_row_pieces(int64_t n, const int64_t* __restrict__ rp,
                                                     const T* __restrict__ eps, double fires_split, int edges_cap,
                                                     int64_t* __restrict__ total) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const Stagger none{0, 1, 0, 0};
    total[i] = split_row<T, false>(rp, eps, i, fires_split, edges_cap, none, 0, nullptr, nullptr, nullptr);
}

// Items of each row in phase 'phase'.
__global__ void __launch_bounds__(kBlock) k_count_items(int64_t n, const int64_t* __restrict__ total,
                                                      const int64_t* __restrict__ pt_off,
                                                      const int32_t* __restrict__ of_row, int phases, int phase,
                                                      int64_t* __restrict__ count) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    count[i] = pieces_in_phase(total[i], row_phase(i, pt_off, of_row, phases), phases, phase);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill_items(int64_t n, const int64_t* __restrict__ rp,
                                                     const T* __restrict__ eps, double fires_split, int edges_cap,
                                                     const int64_t* __restrict__ total,
                                                     const int64_t* __restrict__ pt_off,
                                                     const int32_t* __restrict__ of_row, int phases, int phase,
                                                     const int64_t* __restrict__ offs, int64_t* __restrict__ ibeg,
                                                     int32_t* __restrict__ ilen, int32_t* __restrict__ irow) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    if (offs[i + 1] == offs[i]) return;
    const Stagger st{row_phase(i, pt_off, of_row, phases), phases, phase, total[i]};
    split_row<T, true>(rp, eps, i, fires_split, edges_cap, st, offs[i], ibeg, ilen, irow);
}

// Longest items first: a stable counting sort of the items on a 128-level
// length key, so that long items start at once and the 64 items of a
// wavefront have nearly equal lengths (the sliced-ELL chunk of a wavefront
// is as wide as its longest item, so mixed lengths would be padding).  With
// placement groups the key is group * 128 + level.  Pass A counts keys per
// tile of kBlock items; pass B scatters each item to offset[key][tile] + its
// rank among the tile's items with that key, then places it (see
// place_item).
constexpr int kLevels = 128;

// 127 for the longest items down to 0 for length 1 (length 0 never occurs):
// lengths up to 32 each have a level of their own, longer ones eight levels
// per octave (within 1/8 of each other); lengths from 122880 share key 0.
__device__ inline int length_key(int32_t len) {
    int lvl;
    if (len <= 32) {
        lvl = len;
    } else {
        const int o = 31 - __clz(len);  // floor(log2 len) >= 5
        lvl = 33 + (o - 5) * 8 + ((len >> (o - 3)) & 7);
    }
    lvl = lvl > kLevels - 1 ? kLevels - 1 : lvl;
    return kLevels - 1 - lvl;
}

__device__ inline int item_key(const int32_t* __restrict__ ilen, const int32_t* __restrict__ irow, int64_t m,
                               const int32_t* __restrict__ of_row, const int32_t* __restrict__ group_of_inst) {
    int g = 0;
    if (group_of_inst) {
        const int64_t row = static_cast<int64_t>(static_cast<uint32_t>(irow[m]) & 0x7fffffffu);
        g = group_of_inst[of_row[row]];
    }
    return g * kLevels + length_key(ilen[m]);
}

__global__ void __launch_bounds__(kBlock) k_item_key_hist(const int32_t* __restrict__ ilen,
                                                        const int32_t* __restrict__ irow, int64_t n_items,
                                                        const int32_t* __restrict__ of_row,
                                                        const int32_t* __restrict__ group_of_inst, int n_keys,
                                                        int64_t ntiles, int64_t* __restrict__ hist) {
    extern __shared__ int cnt[];
    for (int k = threadIdx.x; k < n_keys; k += kBlock) cnt[k] = 0;
    __syncthreads();
    const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (m < n_items) atomicAdd(&cnt[item_key(ilen, irow, m, of_row, group_of_inst)], 1);
    __syncthreads();
    for (int k = threadIdx.x; k < n_keys; k += kBlock) hist[static_cast<int64_t>(k) * ntiles + blockIdx.x] = cnt[k];
}

// Final index of the item at sorted position 'at' within group g (whose
// items start at sorted position g0): chunks of 64 items, 'waves' chunks per
// workgroup, and workgroup b serving group b % n_groups.
__device__ inline int64_t place_item(int64_t at, int64_t g0, int g, int waves, int n_groups) {
    if (n_groups == 1 && waves == 1) return at;
    const int64_t r = at - g0;
    const int64_t chunk = r / kWave;
    const int64_t blk = (chunk / waves) * n_groups + g;
    return (blk * waves + chunk % waves) * kWave + (r % kWave);
}

__global__ void __launch_bounds__(kBlock) k_item_scatter(const int64_t* __restrict__ ibeg, const int32_t* __restrict__ ilen,
                                                       const int32_t* __restrict__ irow,
                                                       int64_t n_items, const int32_t* __restrict__ of_row,
                                                       const int32_t* __restrict__ group_of_inst, int n_keys,
                                                       int64_t ntiles, const int64_t* __restrict__ offs, int waves,
                                                       int n_groups, const int32_t* __restrict__ iaux,
                                                       int64_t* __restrict__ obeg, int32_t* __restrict__ olen,
                                                       int32_t* __restrict__ orow, int32_t* __restrict__ oinst,
                                                       int32_t* __restrict__ oaux) {
    extern __shared__ int wave_cnt[];  // [kBlock / kWave][n_keys]
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    for (int k = threadIdx.x; k < (kBlock / kWave) * n_keys; k += kBlock) wave_cnt[k] = 0;
    __syncthreads();
    const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const bool live = m < n_items;
    const int key = live ? item_key(ilen, irow, m, of_row, group_of_inst) : -1;
    // rank among earlier lanes of this wave with the same key: one round per
    // distinct key present in the wave
    int rank = 0;
    bool done = !live;
    while (true) {
        const unsigned long long todo = __ballot(!done);
        if (todo == 0) break;
        const int first = __ffsll(static_cast<long long>(todo)) - 1;
        const int k = __shfl(key, first, kWave);
        const unsigned long long bal = __ballot(!done && key == k);
        if (!done && key == k) {
            rank = __popcll(bal & ((1ULL << lane) - 1ULL));
            done = true;
        }
        if (lane == 0) wave_cnt[wid * n_keys + k] = __popcll(bal);
    }
    __syncthreads();
    if (!live) return;
    for (int w = 0; w < wid; ++w) rank += wave_cnt[w * n_keys + key];
    const int64_t at = offs[static_cast<int64_t>(key) * ntiles + blockIdx.x] + rank;
    const int g = key / kLevels;
    const int64_t g0 = offs[static_cast<int64_t>(g * kLevels) * ntiles];
    const int64_t to = place_item(at, g0, g, waves, n_groups);
    obeg[to] = ibeg[m];
    olen[to] = ilen[m];
    orow[to] = irow[m];
    oaux[to] = iaux[m];
    if (oinst) {
        const int64_t row = static_cast<int64_t>(static_cast<uint32_t>(irow[m]) & 0x7fffffffu);
        oinst[to] = of_row ? of_row[row] : 0;
    }
}

// Padding items: length 0, row 0, instance 0, marked kAuxPadding.
__global__ void __launch_bounds__(kBlock) k_clear_items(int64_t count, int64_t* __restrict__ obeg,
                                                      int32_t* __restrict__ olen, int32_t* __restrict__ orow,
                                                      int32_t* __restrict__ oinst, int32_t* __restrict__ oaux) {
    const int64_t m = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (m >= count) return;
    obeg[m] = 0;
    olen[m] = 0;
    orow[m] = 0;
    oinst[m] = 0;
    oaux[m] = kAuxPadding;
}

// Split rows (rows cut into more than one item).  Pass 1, per row: its
// pieces in the phase (0 unless the row is split) and whether it has any.
__global__ void __launch_bounds__(kBlock) k_split_count(int64_t n, const int64_t* __restrict__ offs,
                                                      const int64_t* __restrict__ total,
                                                      int64_t* __restrict__ pieces, int64_t* __restrict__ is_split) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t c = offs[i + 1] - offs[i];
    const bool split = total[i] > 1 && c > 0;
    pieces[i] = split ? c : 0;
    is_split[i] = split ? 1 : 0;
}

// Pass 2, per row (after exclusive scans of pass 1): the aux value of each
// of its items in the phase before the length sort (the pieces of a row are
// consecutive there, in edge order) and, for a split row, its entry in the
// split list (its tagged row index, as its items carry it).
__global__ void __launch_bounds__(kBlock) k_split_fill(int64_t n, const int64_t* __restrict__ offs,
                                                     const int64_t* __restrict__ total,
                                                     const int32_t* __restrict__ irow,
                                                     const int64_t* __restrict__ slot_off,
                                                     const int64_t* __restrict__ row_off, int32_t* __restrict__ aux,
                                                     int32_t* __restrict__ srow, int64_t* __restrict__ sfirst,
                                                     int32_t* __restrict__ scount) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t p0 = offs[i], c = offs[i + 1] - p0;
    if (c == 0) return;
    if (total[i] == 1) {
        aux[p0] = kAuxWholeRow;
    } else {
        const int64_t s0 = slot_off[i];
        for (int64_t q = 0; q < c; ++q) aux[p0 + q] = static_cast<int32_t>(s0 + q);
        const int64_t r = row_off[i];
        srow[r] = irow[p0];
        sfirst[r] = s0;
        scount[r] = static_cast<int32_t>(c);
    }
}

// Sum of 1 / eps over each instance's entries (blockIdx.y = instance).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_instance_fires(const int64_t* __restrict__ rp, const T* __restrict__ eps,
                                                         const int64_t* __restrict__ pt_off,
                                                         double* __restrict__ fires) {
    const int q = blockIdx.y;
    const int64_t lo = rp[pt_off[q]], hi = rp[pt_off[q + 1]];
    double f = 0;
    for (int64_t e = lo + static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < hi;
         e += static_cast<int64_t>(gridDim.x) * kBlock)
        f += 1.0 / static_cast<double>(eps[e]);
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) f += __shfl_xor(f, m, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0 && f != 0) atomicAdd(fires + q, f);
}

}  // namespace

uint64_t counter_epoch_key(uint64_t seed, int epoch) { return mix64(seed + kGold * (static_cast<uint64_t>(epoch) + 1)); }

template <typename T>
void launch_layout_epoch(const LayoutView<T>& view, const EpochArgs& ea, hipStream_t s) {
    // the epoch's partial tallies go to its ring entry
    LayoutView<T> v = view;
    v.tally_part += tally_ring_offset(ea.epoch);
    const EpochConst ec{ea.epoch, ea.ekey};
    const T ep = static_cast<T>(ea.epoch);
    const T alpha = static_cast<T>(ea.alpha);
    const int64_t blocks = (v.n_items + kBlock - 1) / kBlock;
    const unsigned grid = static_cast<unsigned>(blocks < 1 ? 1 : blocks);
    if (v.mode == kModePush) {
        if (v.trace && v.dim <= 4) {
            switch (v.dim) {
                case 1: k_layout_rows<T, 1, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                case 2: k_layout_rows<T, 2, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                case 3: k_layout_rows<T, 3, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                default: k_layout_rows<T, 4, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            }
        } else {
            switch (v.dim) {
                case 1: k_layout_rows<T, 1, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                case 2: k_layout_rows<T, 2, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                case 3: k_layout_rows<T, 3, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                case 4: k_layout_rows<T, 4, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
                default: k_layout_rows_generic<T><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            }
        }
    } else {
        // pull mode reaches this launcher only for more than four dimensions
        // (one to four run the sliced-ELL kernel)
        k_layout_pull_generic<T><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha);
        UMX_HIP(hipGetLastError());
        launch_fold_split<T>(v, ea.epoch, s);
    }
    UMX_HIP(hipGetLastError());
}

template <typename T>
void launch_layout_serial(const LayoutView<T>& v, const EpochArgs& ea, const int32_t* samples, int64_t count,
                          unsigned long long* cursor, hipStream_t s) {
    const EpochConst ec{ea.epoch, ea.ekey};
    int* overrun = reinterpret_cast<int*>(cursor + 1);
    k_layout_serial<T><<<1, 1, 0, s>>>(v, ec, static_cast<T>(ea.epoch), static_cast<T>(ea.alpha), samples, count,
                                       cursor, overrun);
    UMX_HIP(hipGetLastError());
}

template <typename T>
void count_row_pieces(int64_t n, const int64_t* rp, const T* eps, double fires_split, int edges_cap,
                      DevBuf<int64_t>& row_pieces, hipStream_t s) {
    if (!(fires_split > 0) || edges_cap < 1) throw Error(-1, "work split parameters must be positive");
    row_pieces.alloc(static_cast<size_t>(n > 0 ? n : 1));
    const unsigned grid = static_cast<unsigned>((n + kBlock - 1) / kBlock > 0 ? (n + kBlock - 1) / kBlock : 1);
    if (n > 0) k_row_pieces<T><<<grid, kBlock, 0, s>>>(n, rp, eps, fires_split, edges_cap, row_pieces.p);
    UMX_HIP(hipGetLastError());
}

template <typename T>
void build_work_items(int64_t n, const int64_t* rp, const T* eps, double fires_split, int edges_cap,
                      const int64_t* row_pieces, const Instances& inst, const ItemPlacement& place, int phases,
                      int phase, DevBuf<int64_t>& item_begin, DevBuf<int32_t>& item_len, DevBuf<int32_t>& item_row,
                      DevBuf<int32_t>& item_inst, DevBuf<int32_t>& item_aux, SplitPlan& split, int64_t& n_items,
                      hipStream_t s) {
    if (!(fires_split > 0) || edges_cap < 1) throw Error(-1, "work split parameters must be positive");
    if (place.waves < 1 || place.waves > 4 || place.n_groups < 1 || place.n_groups > 32)
        throw Error(-1, "placement: waves must be 1..4 and groups 1..32");
    if (place.n_groups > 1 && (place.group_of_inst == nullptr || (inst.n_inst > 1 && inst.of_row == nullptr)))
        throw Error(-1, "placement groups need the instance of every row");
    DevBuf<int64_t> count, offs;
    count.alloc(static_cast<size_t>(n > 0 ? n : 1));
    offs.alloc(static_cast<size_t>(n + 1));
    const unsigned grid = static_cast<unsigned>((n + kBlock - 1) / kBlock > 0 ? (n + kBlock - 1) / kBlock : 1);
    if (phases < 1 || phases > kMaxPhases || (phases & (phases - 1)) != 0 || phase < 0 || phase >= phases)
        throw Error(-1, "phases must be a power of two in [1, 8] and phase in [0, phases)");
    const int32_t* of_row_p = inst.n_inst > 1 ? inst.of_row : nullptr;
    if (n > 0) k_count_items<<<grid, kBlock, 0, s>>>(n, row_pieces, inst.pt_off, of_row_p, phases, phase, count.p);
    UMX_HIP(hipGetLastError());
    DevBuf<int64_t> unused;
    exclusive_scan_i64(count.p, offs.p, n, unused, s);
    int64_t real = 0;
    UMX_HIP(hipMemcpyAsync(&real, offs.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    if (real > 0x7fffffffLL) throw Error(-5, "too many work items");
    const size_t M = static_cast<size_t>(real > 0 ? real : 1);
    DevBuf<int64_t> ib;
    DevBuf<int32_t> il, ir;
    ib.alloc(M);
    il.alloc(M);
    ir.alloc(M);
    if (n > 0)
        k_fill_items<T><<<grid, kBlock, 0, s>>>(n, rp, eps, fires_split, edges_cap, row_pieces, inst.pt_off,
                                                of_row_p, phases, phase, offs.p, ib.p, il.p, ir.p);
    UMX_HIP(hipGetLastError());

    // split rows: piece slots numbered row by row in edge order
    DevBuf<int32_t> ia;
    ia.alloc(M);
    {
        DevBuf<int64_t> pieces, flags, slot_off, row_off;
        pieces.alloc(static_cast<size_t>(n > 0 ? n : 1));
        flags.alloc(static_cast<size_t>(n > 0 ? n : 1));
        slot_off.alloc(static_cast<size_t>(n + 1));
        row_off.alloc(static_cast<size_t>(n + 1));
        if (n > 0) k_split_count<<<grid, kBlock, 0, s>>>(n, offs.p, row_pieces, pieces.p, flags.p);
        UMX_HIP(hipGetLastError());
        exclusive_scan_i64(pieces.p, slot_off.p, n, unused, s);
        exclusive_scan_i64(flags.p, row_off.p, n, unused, s);
        int64_t tot[2] = {0, 0};
        UMX_HIP(hipMemcpyAsync(&tot[0], slot_off.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipMemcpyAsync(&tot[1], row_off.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        if (tot[0] > 0x7fffffffLL) throw Error(-5, "too many pieces of split rows");
        split.slots = tot[0];
        split.rows = tot[1];
        split.row.alloc(static_cast<size_t>(split.rows));
        split.first.alloc(static_cast<size_t>(split.rows));
        split.count.alloc(static_cast<size_t>(split.rows));
        if (n > 0)
            k_split_fill<<<grid, kBlock, 0, s>>>(n, offs.p, row_pieces, ir.p, slot_off.p, row_off.p, ia.p,
                                                 split.row.p, split.first.p, split.count.p);
        UMX_HIP(hipGetLastError());
        UMX_HIP(hipStreamSynchronize(s));
    }

    const int G = place.n_groups;
    const int n_keys = G * kLevels;
    const int32_t* of_row = inst.n_inst > 1 ? inst.of_row : nullptr;
    const int32_t* goi = G > 1 ? place.group_of_inst : nullptr;
    const int64_t ntiles = (real + kBlock - 1) / kBlock;
    DevBuf<int64_t> hist, hoff;
    hist.alloc(static_cast<size_t>(n_keys * (ntiles > 0 ? ntiles : 1)));
    hoff.alloc(static_cast<size_t>(n_keys * ntiles + 1));
    // Final item count: each group padded to whole workgroups of
    // waves * 64 items, every group given as many workgroups as the largest
    // (one group: the items padded to whole workgroups).
    const int64_t per_blk = static_cast<int64_t>(place.waves) * kWave;
    int64_t padded = (real + per_blk - 1) / per_blk * per_blk;
    if (real > 0) {
        k_item_key_hist<<<static_cast<unsigned>(ntiles), kBlock, sizeof(int) * n_keys, s>>>(
            il.p, ir.p, real, of_row, goi, n_keys, ntiles, hist.p);
        UMX_HIP(hipGetLastError());
        exclusive_scan_i64(hist.p, hoff.p, n_keys * ntiles, unused, s);
        if (G > 1) {
            std::vector<int64_t> starts(static_cast<size_t>(G) + 1);
            for (int g = 0; g < G; ++g)
                UMX_HIP(hipMemcpyAsync(&starts[g], hoff.p + static_cast<int64_t>(g * kLevels) * ntiles,
                                       sizeof(int64_t), hipMemcpyDeviceToHost, s));
            UMX_HIP(hipStreamSynchronize(s));
            starts[G] = real;
            int64_t most = 0;
            for (int g = 0; g < G; ++g) most = std::max(most, (starts[g + 1] - starts[g] + per_blk - 1) / per_blk);
            padded = most * G * per_blk;
        }
    }
    if (padded > 0x7fffffffLL) throw Error(-5, "too many work items");
    n_items = padded;
    const size_t P = static_cast<size_t>(padded > 0 ? padded : 1);
    item_begin.alloc(P);
    item_len.alloc(P);
    item_row.alloc(P);
    item_inst.alloc(P);
    item_aux.alloc(P);
    if (padded > real)
        k_clear_items<<<static_cast<unsigned>((padded + kBlock - 1) / kBlock), kBlock, 0, s>>>(
            padded, item_begin.p, item_len.p, item_row.p, item_inst.p, item_aux.p);
    if (real > 0) {
        const size_t lds = sizeof(int) * static_cast<size_t>((kBlock / kWave) * n_keys);
        k_item_scatter<<<static_cast<unsigned>(ntiles), kBlock, lds, s>>>(
            ib.p, il.p, ir.p, real, of_row, goi, n_keys, ntiles, hoff.p, place.waves, G, ia.p, item_begin.p,
            item_len.p, item_row.p, item_inst.p, item_aux.p);
        UMX_HIP(hipGetLastError());
    }
    UMX_HIP(hipStreamSynchronize(s));
}

template <typename T>
void instance_fires(const int64_t* rp, const T* eps, const Instances& inst, double* fires, hipStream_t s) {
    UMX_HIP(hipMemsetAsync(fires, 0, sizeof(double) * static_cast<size_t>(inst.n_inst), s));
    k_instance_fires<T><<<dim3(64, static_cast<unsigned>(inst.n_inst)), kBlock, 0, s>>>(rp, eps, inst.pt_off, fires);
    UMX_HIP(hipGetLastError());
}

__global__ void __launch_bounds__(kBlock) k_offset_nonneg(int32_t* __restrict__ a, int64_t count, int64_t by) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (q < count && a[q] >= 0) a[q] = static_cast<int32_t>(a[q] + by);
}
__global__ void __launch_bounds__(kBlock) k_offset_all(int64_t* __restrict__ a, int64_t count, int64_t by) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (q < count) a[q] += by;
}

// Pull mode needs every stored edge i -> t to have a reverse t -> i with the
// same epochs_per_sample (then both fire in the same epochs).  One lane per
// row; a linear search of row t (rows need not be sorted); *bad is set when
// an edge has no such reverse.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_check_reverse(int64_t n, const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ tg, const T* __restrict__ eps,
                                                        int* __restrict__ bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
        const int64_t t = tg[e];
        const T w = eps[e];
        bool found = false;
        if (t >= 0 && t < n) {
            for (int64_t f = rp[t]; f < rp[t + 1] && !found; ++f) found = tg[f] == i && eps[f] == w;
        }
        if (!found) {
            atomicOr(bad, 1);
            return;
        }
    }
}

__global__ void k_clear_flag(int* p) { *p = 0; }

template <typename T>
bool has_reverse_edges(int64_t n, const int64_t* rp, const int32_t* tg, const T* eps, hipStream_t s) {
    if (n <= 0) return true;
    DevBuf<int> flag;
    flag.alloc(1);
    k_clear_flag<<<1, 1, 0, s>>>(flag.p);
    k_check_reverse<T><<<static_cast<unsigned>((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(n, rp, tg, eps, flag.p);
    UMX_HIP(hipGetLastError());
    int h = 0;
    UMX_HIP(hipMemcpyAsync(&h, flag.p, sizeof(int), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    return h == 0;
}

void offset_nonnegative(int32_t* a, int64_t count, int64_t by, hipStream_t s) {
    if (count <= 0 || by == 0) return;
    k_offset_nonneg<<<static_cast<unsigned>((count + kBlock - 1) / kBlock), kBlock, 0, s>>>(a, count, by);
    UMX_HIP(hipGetLastError());
}
void offset_all(int64_t* a, int64_t count, int64_t by, hipStream_t s) {
    if (count <= 0 || by == 0) return;
    k_offset_all<<<static_cast<unsigned>((count + kBlock - 1) / kBlock), kBlock, 0, s>>>(a, count, by);
    UMX_HIP(hipGetLastError());
}

template <typename T>
void launch_fold_split(const LayoutView<T>& v, int epoch, hipStream_t s) {
    if (v.n_split_rows <= 0) return;
    const unsigned grid = static_cast<unsigned>((v.n_split_rows + kBlock - 1) / kBlock);
    switch (v.dim) {
        case 1: k_fold_split<T, 1><<<grid, kBlock, 0, s>>>(v, epoch); break;
        case 2: k_fold_split<T, 2><<<grid, kBlock, 0, s>>>(v, epoch); break;
        case 3: k_fold_split<T, 3><<<grid, kBlock, 0, s>>>(v, epoch); break;
        case 4: k_fold_split<T, 4><<<grid, kBlock, 0, s>>>(v, epoch); break;
        default: k_fold_split_generic<T><<<grid, kBlock, 0, s>>>(v, epoch); break;
    }
    UMX_HIP(hipGetLastError());
}

template void count_row_pieces<float>(int64_t, const int64_t*, const float*, double, int, DevBuf<int64_t>&,
                                      hipStream_t);
template void build_work_items<float>(int64_t, const int64_t*, const float*, double, int, const int64_t*,
                                      const Instances&,
                                      const ItemPlacement&, int, int, DevBuf<int64_t>&, DevBuf<int32_t>&,
                                      DevBuf<int32_t>&, DevBuf<int32_t>&, DevBuf<int32_t>&, SplitPlan&, int64_t&,
                                      hipStream_t);
template void count_row_pieces<double>(int64_t, const int64_t*, const double*, double, int, DevBuf<int64_t>&,
                                       hipStream_t);
template void build_work_items<double>(int64_t, const int64_t*, const double*, double, int, const int64_t*,
                                       const Instances&,
                                       const ItemPlacement&, int, int, DevBuf<int64_t>&, DevBuf<int32_t>&,
                                       DevBuf<int32_t>&, DevBuf<int32_t>&, DevBuf<int32_t>&, SplitPlan&, int64_t&,
                                       hipStream_t);
template bool has_reverse_edges<float>(int64_t, const int64_t*, const int32_t*, const float*, hipStream_t);
template bool has_reverse_edges<double>(int64_t, const int64_t*, const int32_t*, const double*, hipStream_t);
template void launch_fold_split<float>(const LayoutView<float>&, int, hipStream_t);
template void launch_fold_split<double>(const LayoutView<double>&, int, hipStream_t);
template void instance_fires<float>(const int64_t*, const float*, const Instances&, double*, hipStream_t);
template void instance_fires<double>(const int64_t*, const double*, const Instances&, double*, hipStream_t);

template void launch_layout_epoch<float>(const LayoutView<float>&, const EpochArgs&, hipStream_t);
template void launch_layout_epoch<double>(const LayoutView<double>&, const EpochArgs&, hipStream_t);
template void launch_layout_serial<float>(const LayoutView<float>&, const EpochArgs&, const int32_t*, int64_t,
                                          unsigned long long*, hipStream_t);
template void launch_layout_serial<double>(const LayoutView<double>&, const EpochArgs&, const int32_t*, int64_t,
                                           unsigned long long*, hipStream_t);

}  // namespace umx
