// Production layout epoch for gfx950 over a sliced-ELL copy of the schedule.
//
// Layout in HBM.  Work items (rows, heavy rows cut into pieces, sorted
// longest first) are grouped 64 at a time into chunks, one chunk per
// wavefront.  A chunk of width W (its longest item) holds 64 * W slots stored
// column-major: slot(k, lane) = chunk_base + k * 64 + lane, slot k of a lane
// being edge k of its item (padding slots never fire).  So a wavefront's load
// of "column k" of the targets, epochs_per_sample, epoch_of_next_sample or
// epoch_of_next_negative_sample is 64 consecutive values: every schedule byte
// is fetched once per epoch, fully coalesced, instead of each lane walking its
// own CSR row and touching a different cache line per load.
//
// Kernel.  A wavefront reads its chunk TW columns at a time (coalesced, all
// loads of a tile in flight together) and advances the schedule of the
// firing slots at once: the fire test, the number of negative samples and
// both updated epochs depend on the schedule alone, so they are computed in
// this pass (for every column, in straight-line code, so that the columns'
// IEEE divisions overlap) and stored straight away, and the stores complete while the
// slots are walked (the vector memory counter retires in order, so stores
// issued just before the next tile's loads would hold those loads back).
// Each lane keeps its slots' targets and sample counts in LDS (it reads back
// only its own entries, so no barriers are needed) and walks its firing
// slots with count-trailing-zeros: on every trip each lane applies one whole
// firing edge (pull mode: head step, reverse-edge tail step, negative
// samples, exactly as in k_layout_pull_generic), fetching the tail and
// kNegFetch negative samples in one batch of unconditional independent loads.
// Schedule loads and stores are non-temporal, so the stream of schedule
// bytes does not evict from the L2 the embedding that the gathers read.  The
// CSR arrays remain the canonical schedule state; the host syncs the two
// copies when it switches between this kernel and the CSR kernels (serial
// replay, push mode, export).
//
// Determinism.  The embedding is double-buffered: epoch e reads from buffer
// e & 1 and writes every point's result to the other buffer.  The epoch
// runs in phases, one launch each (observations dealt to phases in blocks of
// 16 indices, see phase_of): phase p reads the points of earlier phases
// from the output buffer, as this epoch left them, and every other point
// from the input buffer.  Every item writes, moved or not: a whole row its
// head's new position, a piece of a split row its move into its own slot,
// which a fold after the phase adds to the head in edge order.  The pieces
// of a split row run one per phase in turn, starting at the row's own phase
// (piece q in phase (p0 + q) mod phases), each from the head as the earlier
// phases' folds left it (the first from the input buffer), so the row's
// steps follow one another as in one long item; only a row of more pieces
// than phases has several pieces, summed from one start, in a phase.  No
// lane reads what another lane of its launch writes and
// nothing is accumulated with atomics, so the result of an epoch does not
// depend on the order the wavefronts run in, the tile width, the wavefronts
// per workgroup or the placement: the same inputs and number of phases give
// the same embedding bit for bit, and a run that stops at any epoch and
// restarts equals one run.
//
// Batches.  The same kernel runs many independent problems (instances) in one
// launch: each lane reads its item's instance constants (observation offset
// and count, seed, epochs, learning rates), draws negative samples inside its
// instance, and keys the sampler by the instance's seed and its own entry
// numbers, so an instance sees the samples it would see when run alone.  The
// epoch number can live in device memory (launch q of a block runs epoch
// counter + q), advanced by a one-wavefront kernel after the block's
// launches (the one that also folds their partial tallies), so a block of
// launches captured once as a HIP graph replays for any epochs.
#include "layout_device.hpp"
#include "umx_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>

namespace umx {

namespace {

using namespace umx::dev;

// Negative samples fetched per batch: the reference's default rate (5) is
// one batch for most firing edges; a few draw a sixth in a second batch.
constexpr int kNegFetch = 5;
#ifndef UMX_LONG_PRIO
#define UMX_LONG_PRIO 1
#endif

// ---------------------------------------------------------------------------
// Build: chunk widths, chunk offsets, constant slot data and the slot map.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kWave) k_chunk_width(const int32_t* __restrict__ item_len, int64_t n_items,
                                                     int32_t* __restrict__ width, int64_t* __restrict__ slots) {
    const int64_t c = blockIdx.x;
    const int64_t m = c * kWave + threadIdx.x;
    int w = m < n_items ? item_len[m] : 0;
#pragma unroll
    for (int s = kWave / 2; s >= 1; s >>= 1) {
        const int o = __shfl_xor(w, s, kWave);
        w = o > w ? o : w;
    }
    if (threadIdx.x == 0) {
        width[c] = w;
        slots[c] = static_cast<int64_t>(w) * kWave;
    }
}

template <typename T>
__global__ void __launch_bounds__(kWave) k_fill_sell(const int64_t* __restrict__ item_begin,
                                                   const int32_t* __restrict__ item_len,
                                                   const int32_t* __restrict__ item_row, int64_t n_items,
                                                   const int32_t* __restrict__ width, const int64_t* __restrict__ base,
                                                   const int32_t* __restrict__ tg, const T* __restrict__ eps,
                                                   int32_t* __restrict__ stg, T* __restrict__ seps,
                                                   T* __restrict__ snx, T* __restrict__ snn,
                                                   int64_t* __restrict__ slot_of_edge) {
    const int64_t c = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t m = c * kWave + lane;
    const int w = width[c];
    const int64_t b0 = base[c];
    const bool live = m < n_items;
    const int64_t e0 = live ? item_begin[m] : 0;
    const int len = live ? item_len[m] : 0;
    const int32_t self = live ? static_cast<int32_t>(static_cast<uint32_t>(item_row[m]) & 0x7fffffffu) : 0;
    // kFillUnroll columns per trip: the loads of a trip are issued together
    // (padding slots of a nonempty item read its first edge again; empty
    // items, which are padding, and lanes past the last item load nothing)
    // before any store, so a long chunk waits for one round trip per trip
    // instead of one per column
    constexpr int kFillUnroll = 8;
    int k = 0;
    for (; k + kFillUnroll <= w; k += kFillUnroll) {
        int32_t t[kFillUnroll];
        T ep[kFillUnroll];
        if (len > 0) {
#pragma unroll
            for (int u = 0; u < kFillUnroll; ++u) {
                const int64_t e = k + u < len ? e0 + k + u : e0;
                t[u] = tg[e];
                ep[u] = eps[e];
            }
        }
        __builtin_amdgcn_sched_barrier(0);  // every load of the trip before its stores
#pragma unroll
        for (int u = 0; u < kFillUnroll; ++u) {
            const int64_t slot = b0 + static_cast<int64_t>(k + u) * kWave + lane;
            if (k + u < len) {
                stg[slot] = t[u];
                seps[slot] = ep[u];
                slot_of_edge[e0 + k + u] = slot;
            } else {
                stg[slot] = self;
                seps[slot] = static_cast<T>(1);
                snx[slot] = static_cast<T>(INFINITY);  // padding never fires
                snn[slot] = static_cast<T>(0);
            }
        }
    }
    for (; k < w; ++k) {
        const int64_t slot = b0 + static_cast<int64_t>(k) * kWave + lane;
        if (k < len) {
            const int64_t e = e0 + k;
            stg[slot] = tg[e];
            seps[slot] = eps[e];
            slot_of_edge[e] = slot;
        } else {
            stg[slot] = self;
            seps[slot] = static_cast<T>(1);
            snx[slot] = static_cast<T>(INFINITY);  // padding never fires
            snn[slot] = static_cast<T>(0);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_csr_to_sell(int64_t E, const int64_t* __restrict__ slot_of_edge,
                                                   const T* __restrict__ nx, const T* __restrict__ nn,
                                                   T* __restrict__ snx, T* __restrict__ snn) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= E) return;
    const int64_t s = slot_of_edge[e];
    snx[s] = nx[e];
    snn[s] = nn[e];
}

template <typename T>
__global__ void __launch_bounds__(256) k_sell_to_csr(int64_t E, const int64_t* __restrict__ slot_of_edge,
                                                   const T* __restrict__ snx, const T* __restrict__ snn,
                                                   T* __restrict__ nx, T* __restrict__ nn) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (e >= E) return;
    const int64_t s = slot_of_edge[e];
    nx[e] = snx[s];
    nn[e] = snn[s];
}

// ---------------------------------------------------------------------------
// Epoch kernel: one wavefront per chunk.
// ---------------------------------------------------------------------------
template <typename T, int TW>
struct SellTile {
    int32_t nneg[TW][kWave];
    int32_t tg[TW][kWave];
};

// One wavefront per chunk; a workgroup of W wavefronts (W = blockDim.x / 64)
// takes chunks blockIdx.x * W .. + W - 1.  Each lane reads its item's
// instance constants (observation offset and count, sampler seed, epochs,
// learning rate); a plain layout is one instance.  The epoch is epoch_arg,
// or v.epoch_ctr[0] + (-1 - epoch_arg) when epoch_arg < 0 (captured blocks
// of epochs: launch q of a block passes -1 - q, and a one-wavefront kernel
// after the block's launches advances the counter by the block length).
template <typename T, int D, int TW, bool TRACE, bool SMALL>
__global__ void __launch_bounds__(256) k_layout_sell(LayoutView<T> v, int epoch_arg) {
    extern __shared__ __align__(16) unsigned char sell_lds[];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    SellTile<T, TW>& tile = reinterpret_cast<SellTile<T, TW>*>(sell_lds)[wid];
    const int64_t c = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + wid;
    const int epn = epoch_arg >= 0 ? epoch_arg : v.epoch_ctr[0] + (-1 - epoch_arg);
    const T ep = static_cast<T>(epn);
    // double buffer: the epoch reads buffer epn & 1 and writes the other;
    // points of earlier phases are read from the output buffer
    const T* __restrict__ src = (epn & 1) ? v.emb_alt : v.emb;
    T* __restrict__ dst = (epn & 1) ? v.emb : v.emb_alt;

    uint32_t fires = 0, draws = 0, self_hits = 0;  // one chunk's counts fit in 32 bits
    if (c < v.n_chunks) {
        const int64_t m = c * kWave + lane;
        const int32_t aux = m < v.n_items ? v.item_aux[m] : kAuxPadding;
        const bool has_item = aux != kAuxPadding;
        bool live = has_item;  // the item's instance runs this epoch
        // the chunk's width is the same for the whole wavefront: read as a
        // uniform value, the per-column guards below are scalar branches
        const int width = __builtin_amdgcn_readfirstlane(v.chunk_width[c]);
#if UMX_LONG_PRIO
        // Wide chunks (the pieces of heavy rows) are the longest chains of
        // sequential steps of a launch, and wavefronts of short chunks share
        // their SIMD: the wider the chunk, the higher its wavefront's issue
        // priority, so that the launch does not wait for a long chain that ran
        // at a fraction of the SIMD's rate.  (Priority orders issue among the
        // wavefronts of a SIMD; it changes no result.)
        if (width >= 40) __builtin_amdgcn_s_setprio(3);
        else if (width >= 32) __builtin_amdgcn_s_setprio(2);
        else if (width >= 24) __builtin_amdgcn_s_setprio(1);
#endif
        const int64_t base = v.chunk_base[c];

        int64_t i = 0, e0 = 0, edge_off = 0, pt_off = 0;
        int32_t i32 = 0;
        uint32_t n32 = 0;
        uint64_t ekey = 0;
        T alpha = 0;
        Vec<T, D> h0{}, h{};
        if (has_item) {
            const int q = v.item_inst[m];
            const InstConst ic = v.inst[q];
            live = epn < ic.epochs;
            pt_off = ic.pt_off;
            edge_off = ic.edge_off;
            n32 = ic.n;
            ekey = mix64(ic.seed + kGold * (static_cast<uint64_t>(epn) + 1));
            alpha = live ? v.alpha_tbl[static_cast<int64_t>(q) * v.alpha_stride + epn] : static_cast<T>(0);
            const int32_t tagged = v.item_row[m];
            i = static_cast<int64_t>(static_cast<uint32_t>(tagged) & 0x7fffffffu);
            i32 = static_cast<int32_t>(i);
            e0 = v.item_begin[m];
            // a piece of a split row whose row an earlier phase of this
            // epoch moved (top bit) starts where that phase's fold left it
            h0 = load_point<T, D>(tagged < 0 ? dst : src, i);
            h = h0;
        }
        const T two_ab = -2 * v.a * v.b;
        const T two_gb = 2 * v.gamma * v.b;
        // where point j (global index, in this lane's instance) is read from:
        // points of earlier phases from the output buffer (in the first phase,
        // and with one phase, the comparison never holds)
        auto from = [&](int64_t j) -> const T* {
            const uint32_t local = static_cast<uint32_t>(j) - static_cast<uint32_t>(pt_off);
            return phase_of(local, v.phases) < v.phase ? dst : src;
        };
        // The same read with 32-bit byte offsets (SMALL: the host checked that
        // both buffers, which lie one after the other in one allocation, end
        // within 4 GiB of the first): buffer offset plus the point's offset
        // from one base, and the phase as one bit-field extract.
        const char* const emb_base = reinterpret_cast<const char*>(v.emb);
        const uint32_t alt_off = static_cast<uint32_t>(reinterpret_cast<const char*>(v.emb_alt) - emb_base);
        const uint32_t src_off = (epn & 1) ? alt_off : 0u;
        const uint32_t dst_off = alt_off - src_off;
        const uint32_t phase_bits = static_cast<uint32_t>(31 - __clz(v.phases));
        auto gather = [&](int64_t j) -> Vec<T, D> {
            if constexpr (SMALL) {
                const uint32_t local = static_cast<uint32_t>(j) - static_cast<uint32_t>(pt_off);
                const uint32_t ph = __builtin_amdgcn_ubfe(local, static_cast<uint32_t>(kPhaseBlockShift), phase_bits);
                const uint32_t off = (ph < static_cast<uint32_t>(v.phase) ? dst_off : src_off) +
                                     static_cast<uint32_t>(j) * static_cast<uint32_t>(D * sizeof(T));
                return load_point<T, D>(reinterpret_cast<const T*>(emb_base + static_cast<uint64_t>(off)), 0);
            } else {
                return load_point<T, D>(from(j), j);
            }
        };
        // Draws p0 .. p0 + kNegFetch - 1 of an edge and their points, as one
        // batch of unconditional loads (slots past nneg read the head's own
        // point and are marked -1); conditional loads would each be followed
        // by a wait, leaving one or two gathers in flight instead of all.
        auto fetch = [&](uint64_t ej, int p0, int nneg, int32_t* kk, Vec<T, D>* q) {
#pragma unroll
            for (int u = 0; u < kNegFetch; ++u) {
                const uint32_t p = static_cast<uint32_t>(p0 + u);
                const int32_t d =
                    static_cast<int32_t>(pt_off + (SMALL ? draw_index_small(ej, p, n32) : draw_index(ej, p, n32)));
                const bool use = p0 + u < nneg;
                kk[u] = use ? d : -1;
                const int64_t pt = use ? static_cast<int64_t>(d) : i;
                q[u] = gather(pt);
            }
        };

        // The schedule of the tile of columns t0 .. t0 + TW - 1, coalesced,
        // all columns in flight together (columns past the chunk's end are
        // given values that never fire).
        T xs[TW], ys[TW], es[TW];
        int32_t ts[TW];
        auto load_tile = [&](int t0) {
            const int tw = width - t0 < TW ? width - t0 : TW;
#pragma unroll
            for (int k = 0; k < TW; ++k) {
                if (k < tw) {
                    const int64_t slot = base + static_cast<int64_t>(t0 + k) * kWave + lane;
                    xs[k] = ld_stream(v.sell_nx + slot);
                    ys[k] = ld_stream(v.sell_nn + slot);
                    es[k] = ld_stream(v.sell_eps + slot);
                    ts[k] = ld_stream(v.sell_tg + slot);
                } else {
                    xs[k] = ep + 1;
                    ys[k] = ep;
                    es[k] = 1;
                    ts[k] = 0;
                }
            }
        };
        if (width > 0) load_tile(0);
        for (int t0 = 0; t0 < width; t0 += TW) {
            // The fire test, sample count and both updated epochs of every
            // column in straight-line code (a column that does not fire has
            // them computed and dropped), so the columns' division chains
            // interleave instead of running one column at a time.
            unsigned int mask = 0;
            T nxv[TW], nnv[TW];
#pragma unroll
            for (int k = 0; k < TW; ++k) {
                const bool fire = live && !(xs[k] > ep);
                const T epns = es[k] / v.nsr;
                const T q = (ep - ys[k]) / epns;
                const int nneg = static_cast<int>(fire ? q : static_cast<T>(0));
                tile.tg[k][lane] = ts[k];
                tile.nneg[k][lane] = nneg;
                nxv[k] = xs[k] + es[k];
                nnv[k] = ys[k] + static_cast<T>(nneg) * epns;
                mask |= (fire ? 1u : 0u) << k;
            }
            // The firing slots' schedule is stored at once (it depends on the
            // schedule only), so the stores complete while the slots are
            // walked.
#pragma unroll
            for (int k = 0; k < TW; ++k) {
                if (mask & (1u << k)) {
                    const int64_t slot = base + static_cast<int64_t>(t0 + k) * kWave + lane;
                    st_stream(v.sell_nx + slot, nxv[k]);
                    st_stream(v.sell_nn + slot, nnv[k]);
                }
            }
            while (mask) {
                const int k = __builtin_ctz(mask);
                mask &= mask - 1;
                // instance-local entry number of this edge: keys the sampler
                const int64_t j = e0 + t0 + k - edge_off;
                ++fires;
                const int32_t t = tile.tg[k][lane];
                const int nneg = tile.nneg[k][lane];
                draws += static_cast<uint32_t>(nneg > 0 ? nneg : 0);
                const uint64_t ej = edge_key(ekey, static_cast<uint64_t>(j));
                Vec<T, D> r = gather(t);
                int32_t kk[kNegFetch];
                Vec<T, D> q[kNegFetch];
                fetch(ej, 0, nneg, kk, q);
                // every load of the edge issued before the first step waits
                // for the tail (the scheduler would otherwise interleave the
                // later draws' address arithmetic and loads with the steps)
                __builtin_amdgcn_sched_barrier(0);
                // the reverse edge's step is taken where the pair's first
                // event leaves the two points: the head step g of i -> t
                // moves the tail by -g (as in the reference)
#pragma unroll
                for (int half = 0; half < 2; ++half) {
                    const T d2 = sqdist<T, D>(h, r);
                    const T pd2b = fast_pow(d2, v.b);
                    const T gc = prod_div(two_ab * pd2b, d2 * (v.a * pd2b + 1));
                    double* rec = TRACE ? trace_slot(v, half == 0 ? 0 : 2, i, t) : nullptr;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const T g = alpha * clip4(gc * (h.c[d] - r.c[d]));
                        if (TRACE && rec) {
                            rec[3 + d] = h.c[d];
                            rec[3 + D + d] = r.c[d];
                            rec[3 + 2 * D + d] = g;
                        }
                        h.c[d] += g;
                        if (half == 0) r.c[d] -= g;
                    }
                }
                // Negative samples, a batch of kNegFetch at a time.  Every
                // slot of a batch is evaluated and its step kept by a select
                // (slots past the edge's count hold -1, never equal to the
                // row; a draw equal to the row is skipped, as in the
                // reference).  The first batch is applied in straight-line
                // code with no branch around its uses, so its loads, issued
                // with the tail's, stay ahead of the attractive steps, all in
                // flight together (a use inside a branch or a loop lets the
                // compiler move the load there, after the steps before it).
                auto apply_batch = [&]() {
#pragma unroll
                    for (int u = 0; u < kNegFetch; ++u) {
                        self_hits += kk[u] == i32 ? 1u : 0u;
                        const bool apply = kk[u] >= 0 && kk[u] != i32;
                        const T e2 = sqdist<T, D>(h, q[u]);
                        const T rc = prod_div(two_gb, (static_cast<T>(0.001) + e2) * (v.a * fast_pow(e2, v.b) + 1));
                        double* rec = TRACE && apply ? trace_slot(v, 1, i, kk[u]) : nullptr;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            const T g = alpha * clip4(rc * (h.c[d] - q[u].c[d]));
                            if (TRACE && rec) {
                                rec[3 + d] = h.c[d];
                                rec[3 + D + d] = q[u].c[d];
                                rec[3 + 2 * D + d] = g;
                            }
                            h.c[d] = apply ? h.c[d] + g : h.c[d];
                        }
                    }
                };
                apply_batch();
                for (int p0 = kNegFetch; p0 < nneg; p0 += kNegFetch) {
                    fetch(ej, p0, nneg, kk, q);
                    apply_batch();
                }
            }
            if (t0 + TW < width) load_tile(t0 + TW);
        }
        // every item writes, moved or not (the output buffer must hold the
        // whole state): a whole row its new position, a piece of a split row
        // its move, summed onto the head in edge order by the fold
        if (aux >= 0) {
            Vec<T, D> dlt;
#pragma unroll
            for (int d = 0; d < D; ++d) dlt.c[d] = h.c[d] - h0.c[d];
            store_point<T, D>(v.split_delta, aux, dlt);
        } else if (has_item) {
            store_point<T, D>(dst, i, h);
        }
    }
    tally_add(v.tally_part + tally_ring_offset(epn), c, fires, draws, self_hits);
}

__global__ void k_set_epoch(int* ctr, int epoch) { ctr[0] = epoch; }

// Folds the partial tallies of one epoch launch (ring entry of the epoch)
// into the epoch's tally and clears them (one lane per slot); with epoch < 0
// the epoch is epoch_ctr[0], advanced by one afterwards.
__global__ void __launch_bounds__(kTallySlots) k_finish_epoch(unsigned long long* __restrict__ part,
                                                             unsigned long long* __restrict__ tally, int epoch,
                                                             int* __restrict__ ctr) {
    static_assert(kTallySlots == kWave, "one lane per partial-tally slot");
    const int e = epoch >= 0 ? epoch : ctr[0];
    unsigned long long* p = part + tally_ring_offset(e) + static_cast<int64_t>(threadIdx.x) * kTallyStride;
    unsigned long long c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) c[q] = atomicExch(p + q, 0ULL);
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
#pragma unroll
        for (int q = 0; q < 3; ++q) c[q] += __shfl_xor(c[q], m, kWave);
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (c[q]) atomicAdd(tally + 3 * static_cast<int64_t>(e) + q, c[q]);
        if (epoch < 0) ctr[0] = e + 1;
    }
}

// The fold of a captured block: epochs epoch_ctr[0] .. epoch_ctr[0] + count - 1
// in turn (count <= kTallyRing), then the counter advanced by count.  One
// wavefront, launched after the block's epoch kernels.
__global__ void __launch_bounds__(kTallySlots) k_fold_block(unsigned long long* __restrict__ part,
                                                           unsigned long long* __restrict__ tally, int count,
                                                           int* __restrict__ ctr) {
    static_assert(kTallySlots == kWave, "one lane per partial-tally slot");
    const int base = ctr[0];
    for (int b = 0; b < count; ++b) {
        const int e = base + b;
        unsigned long long* p = part + tally_ring_offset(e) + static_cast<int64_t>(threadIdx.x) * kTallyStride;
        unsigned long long c[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            c[q] = p[q];
            p[q] = 0ULL;
        }
#pragma unroll
        for (int m = kWave / 2; m >= 1; m >>= 1) {
#pragma unroll
            for (int q = 0; q < 3; ++q) c[q] += __shfl_xor(c[q], m, kWave);
        }
        if (threadIdx.x == 0) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
                if (c[q]) atomicAdd(tally + 3 * static_cast<int64_t>(e) + q, c[q]);
        }
    }
    // every lane read the counter before the loop; the one wavefront runs in
    // lockstep, so no lane reads it after this store
    if (threadIdx.x == 0) ctr[0] = base + count;
}

// The same fold for epochs first .. first + gridDim.x - 1, one workgroup of
// one wavefront per epoch (the epochs lie in distinct ring entries).
__global__ void __launch_bounds__(kTallySlots) k_fold_epochs(unsigned long long* __restrict__ part,
                                                            unsigned long long* __restrict__ tally, int first) {
    static_assert(kTallySlots == kWave, "one lane per partial-tally slot");
    const int e = first + static_cast<int>(blockIdx.x);
    unsigned long long* p = part + tally_ring_offset(e) + static_cast<int64_t>(threadIdx.x) * kTallyStride;
    unsigned long long c[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        c[q] = p[q];
        p[q] = 0ULL;
    }
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
#pragma unroll
        for (int q = 0; q < 3; ++q) c[q] += __shfl_xor(c[q], m, kWave);
    }
    if (threadIdx.x == 0) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
            if (c[q]) atomicAdd(tally + 3 * static_cast<int64_t>(e) + q, c[q]);
    }
}

template <typename T, int TW, bool TRACE, bool SMALL>
void launch_sell_dim(const LayoutView<T>& v, int epoch, unsigned grid, unsigned block, size_t lds, hipStream_t s) {
    switch (v.dim) {
        case 1: k_layout_sell<T, 1, TW, TRACE, SMALL><<<grid, block, lds, s>>>(v, epoch); break;
        case 2: k_layout_sell<T, 2, TW, TRACE, SMALL><<<grid, block, lds, s>>>(v, epoch); break;
        case 3: k_layout_sell<T, 3, TW, TRACE, SMALL><<<grid, block, lds, s>>>(v, epoch); break;
        default: k_layout_sell<T, 4, TW, TRACE, SMALL><<<grid, block, lds, s>>>(v, epoch); break;
    }
}

template <typename T, int TW>
void launch_sell_tw(const LayoutView<T>& v, int epoch, int waves, hipStream_t s) {
    const unsigned grid = static_cast<unsigned>((v.n_chunks + waves - 1) / waves);
    const unsigned block = static_cast<unsigned>(waves * kWave);
    const size_t lds = sizeof(SellTile<T, TW>) * static_cast<size_t>(waves);
    if (lds > static_cast<size_t>(kSellLdsMax))
        throw Error(-1, "LDS tile times waves per workgroup exceeds 64 KiB; use a smaller tile or fewer waves");
    // 32-bit arithmetic when every instance has at most 2^24 observations
    // and both embedding buffers (one allocation, buffer 1 right after
    // buffer 0) end within 4 GiB of the first: the sampler's draw and the
    // points' addresses are the same either way; the choice is made here,
    // once, so that the kernel carries one of the two forms)
    const uint64_t buf_bytes = static_cast<uint64_t>(v.n) * static_cast<uint64_t>(v.dim) * sizeof(T);
    const bool small = v.max_inst_n <= kSmallDraw && v.emb_alt == v.emb + static_cast<int64_t>(v.n) * v.dim &&
                       2 * buf_bytes <= 0xffffffffull;
    if (v.trace) {
        if (!small)
            throw Error(-1, "the update trace needs instances of at most 2^24 observations and at most 4 GiB of "
                            "embedding buffers");
        launch_sell_dim<T, TW, true, true>(v, epoch, grid, block, lds, s);
    } else if (small) {
        launch_sell_dim<T, TW, false, true>(v, epoch, grid, block, lds, s);
    } else {
        launch_sell_dim<T, TW, false, false>(v, epoch, grid, block, lds, s);
    }
}

}  // namespace

template <typename T>
void build_sell(int64_t E, const int32_t* tg, const T* eps, const int64_t* item_begin, const int32_t* item_len,
                const int32_t* item_row, int64_t n_items, SellState<T>& st, hipStream_t s) {
    st.n_chunks = (n_items + kWave - 1) / kWave;
    const int64_t nc = st.n_chunks;
    st.width.alloc(static_cast<size_t>(nc > 0 ? nc : 1));
    st.base.alloc(static_cast<size_t>(nc + 1));
    DevBuf<int64_t> slots;
    slots.alloc(static_cast<size_t>(nc > 0 ? nc : 1));
    if (nc > 0) {
        k_chunk_width<<<static_cast<unsigned>(nc), kWave, 0, s>>>(item_len, n_items, st.width.p, slots.p);
        UMX_HIP(hipGetLastError());
    }
    DevBuf<int64_t> unused;
    exclusive_scan_i64(slots.p, st.base.p, nc, unused, s);
    UMX_HIP(hipMemcpyAsync(&st.n_slots, st.base.p + nc, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    const size_t ns = static_cast<size_t>(st.n_slots > 0 ? st.n_slots : 1);
    st.tg.alloc(ns);
    st.eps.alloc(ns);
    st.nx.alloc(ns);
    st.nn.alloc(ns);
    st.slot_of_edge.alloc(static_cast<size_t>(E > 0 ? E : 1));
    if (nc > 0) {
        k_fill_sell<T><<<static_cast<unsigned>(nc), kWave, 0, s>>>(item_begin, item_len, item_row, n_items,
                                                                    st.width.p, st.base.p, tg, eps, st.tg.p,
                                                                    st.eps.p, st.nx.p, st.nn.p, st.slot_of_edge.p);
        UMX_HIP(hipGetLastError());
    }
    st.E = E;
    UMX_HIP(hipStreamSynchronize(s));
}

template <typename T>
void sync_csr_to_sell(const T* nx, const T* nn, SellState<T>& st, hipStream_t s) {
    if (st.E == 0) return;
    const unsigned grid = static_cast<unsigned>((st.E + 255) / 256);
    k_csr_to_sell<T><<<grid, 256, 0, s>>>(st.E, st.slot_of_edge.p, nx, nn, st.nx.p, st.nn.p);
    UMX_HIP(hipGetLastError());
}

template <typename T>
void sync_sell_to_csr(T* nx, T* nn, const SellState<T>& st, hipStream_t s) {
    if (st.E == 0) return;
    const unsigned grid = static_cast<unsigned>((st.E + 255) / 256);
    k_sell_to_csr<T><<<grid, 256, 0, s>>>(st.E, st.slot_of_edge.p, st.nx.p, st.nn.p, nx, nn);
    UMX_HIP(hipGetLastError());
}

template <typename T>
void launch_layout_sell(const LayoutView<T>& v, int epoch, int tile_cols, int waves, hipStream_t s) {
    if (waves < 1 || waves > 4) throw Error(-1, "waves per workgroup must be 1..4");
    if (v.n_chunks > 0) {
        switch (tile_cols) {
            case 4: launch_sell_tw<T, 4>(v, epoch, waves, s); break;
            case 8: launch_sell_tw<T, 8>(v, epoch, waves, s); break;
            case 32: launch_sell_tw<T, 32>(v, epoch, waves, s); break;
            default: launch_sell_tw<T, 16>(v, epoch, waves, s); break;
        }
        UMX_HIP(hipGetLastError());
    }
}

void finish_layout_epoch(unsigned long long* tally_part, unsigned long long* tally, int epoch, int* epoch_ctr,
                         hipStream_t s) {
    k_finish_epoch<<<1, kTallySlots, 0, s>>>(tally_part, tally, epoch, epoch_ctr);
    UMX_HIP(hipGetLastError());
}

void fold_layout_epochs(unsigned long long* tally_part, unsigned long long* tally, int first, int count,
                        hipStream_t s) {
    if (count <= 0) return;
    if (count > kTallyRing || first < 0) throw Error(-1, "tally fold covers at most one ring of epochs");
    k_fold_epochs<<<static_cast<unsigned>(count), kTallySlots, 0, s>>>(tally_part, tally, first);
    UMX_HIP(hipGetLastError());
}

void fold_layout_block(unsigned long long* tally_part, unsigned long long* tally, int count, int* epoch_ctr,
                       hipStream_t s) {
    if (count <= 0) return;
    if (count > kTallyRing) throw Error(-1, "tally fold covers at most one ring of epochs");
    k_fold_block<<<1, kTallySlots, 0, s>>>(tally_part, tally, count, epoch_ctr);
    UMX_HIP(hipGetLastError());
}

void set_epoch_counter(int* epoch_ctr, int epoch, hipStream_t s) {
    k_set_epoch<<<1, 1, 0, s>>>(epoch_ctr, epoch);
    UMX_HIP(hipGetLastError());
}

#define UMX_INST(T)                                                                                            \
    template void build_sell<T>(int64_t, const int32_t*, const T*, const int64_t*, const int32_t*,            \
                                const int32_t*, int64_t, SellState<T>&, hipStream_t);                          \
    template void sync_csr_to_sell<T>(const T*, const T*, SellState<T>&, hipStream_t);                         \
    template void sync_sell_to_csr<T>(T*, T*, const SellState<T>&, hipStream_t);                               \
    template void launch_layout_sell<T>(const LayoutView<T>&, int, int, int, hipStream_t);
UMX_INST(float)
UMX_INST(double)
#undef UMX_INST

}  // namespace umx
