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

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

namespace umx {

namespace {

constexpr int kBlock = 256;
constexpr uint64_t kGold = 0x9e3779b97f4a7c15ULL;
constexpr uint64_t kOdd = 0xd1b54a32d192ed03ULL;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ inline uint64_t edge_key(uint64_t ekey, uint64_t j) { return mix64(ekey ^ (j * kOdd)); }
__device__ inline int32_t draw_index(uint64_t ej, uint32_t p, uint32_t n) {
    const uint64_t r = mix64(ej + kGold * (static_cast<uint64_t>(p) + 1));
    return static_cast<int32_t>(__umul64hi(r, static_cast<uint64_t>(n)));
}

template <typename T>
__device__ inline T clip4(T v) {  // std::min(std::max(v, -4), 4), optimize_layout.hpp:105-110
    const T lo = v < static_cast<T>(-4) ? static_cast<T>(-4) : v;
    return static_cast<T>(4) < lo ? static_cast<T>(4) : lo;
}
template <typename T>
__device__ inline T eps_of();
template <>
__device__ inline float eps_of<float>() {
    return FLT_EPSILON;
}
template <>
__device__ inline double eps_of<double>() {
    return DBL_EPSILON;
}

// Fast power for the production kernel: exp2(b * log2(x)) on the hardware
// transcendental units for float (x >= FLT_EPSILON here, so no denormals).
__device__ inline float fast_pow(float x, float y) { return __builtin_amdgcn_exp2f(y * __builtin_amdgcn_logf(x)); }
__device__ inline double fast_pow(double x, double y) { return pow(x, y); }

template <typename T>
__device__ inline void add_to(T* p, T v) {
    unsafeAtomicAdd(p, v);
}

template <typename T, int D>
struct Vec {
    T c[D];
};

template <typename T, int D>
__device__ inline Vec<T, D> load_point(const T* emb, int64_t i) {
    Vec<T, D> v;
#pragma unroll
    for (int d = 0; d < D; ++d) v.c[d] = emb[i * D + d];
    return v;
}
template <>
__device__ inline Vec<float, 2> load_point<float, 2>(const float* emb, int64_t i) {
    const float2 q = *reinterpret_cast<const float2*>(emb + i * 2);
    return Vec<float, 2>{{q.x, q.y}};
}
template <>
__device__ inline Vec<float, 4> load_point<float, 4>(const float* emb, int64_t i) {
    const float4 q = *reinterpret_cast<const float4*>(emb + i * 4);
    return Vec<float, 4>{{q.x, q.y, q.z, q.w}};
}

// Wave-aggregated tally update: one atomic per wavefront per counter.
__device__ inline void tally_add(unsigned long long* t, unsigned long long fires, unsigned long long draws,
                                 unsigned long long self_hits) {
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        fires += __shfl_xor(fires, m, kWave);
        draws += __shfl_xor(draws, m, kWave);
        self_hits += __shfl_xor(self_hits, m, kWave);
    }
    if ((threadIdx.x & (kWave - 1)) == 0) {
        if (fires) atomicAdd(t, fires);
        if (draws) atomicAdd(t + 1, draws);
        if (self_hits) atomicAdd(t + 2, self_hits);
    }
}

// Reserves one trace record; nullptr when tracing is off or the buffer is full.
template <typename T>
__device__ inline double* trace_slot(const LayoutView<T>& v, int kind, int64_t head, int64_t other) {
    if (!v.trace) return nullptr;
    const unsigned long long slot = atomicAdd(v.trace_len, 1ULL);
    if (slot >= static_cast<unsigned long long>(v.trace_cap)) return nullptr;
    double* rec = v.trace + slot * (3 + 3 * static_cast<long long>(v.dim));
    rec[0] = kind;
    rec[1] = static_cast<double>(head);
    rec[2] = static_cast<double>(other);
    return rec;
}

struct EpochConst {
    int epoch;
    uint64_t ekey;
};

template <typename T, int D, bool TRACE>
__global__ void __launch_bounds__(kBlock) k_layout_rows(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    if (i < v.n) {
        const int64_t s = v.rp[i], e = v.rp[i + 1];
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
            const T gc = (two_ab * pd2b) / (d2 * (v.a * pd2b + 1));
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
                const T rc = two_gb / ((static_cast<T>(0.001) + e2) * (v.a * fast_pow(e2, v.b) + 1));
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
    tally_add(v.tally + 3 * static_cast<int64_t>(ec.epoch), fires, draws, self_hits);
}

// Any embedding dimension: the head is updated in place with atomics after
// every step (no register copy), the rest as above.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_layout_rows_generic(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    if (i < v.n) {
        const int D = v.dim;
        const int64_t s = v.rp[i], e = v.rp[i + 1];
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
    tally_add(v.tally + 3 * static_cast<int64_t>(ec.epoch), fires, draws, self_hits);
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

template <typename T, int D>
__device__ inline void store_point(T* emb, int64_t i, const Vec<T, D>& v) {
#pragma unroll
    for (int d = 0; d < D; ++d) emb[i * D + d] = v.c[d];
}
template <>
__device__ inline void store_point<float, 2>(float* emb, int64_t i, const Vec<float, 2>& v) {
    *reinterpret_cast<float2*>(emb + i * 2) = make_float2(v.c[0], v.c[1]);
}
template <>
__device__ inline void store_point<float, 4>(float* emb, int64_t i, const Vec<float, 4>& v) {
    *reinterpret_cast<float4*>(emb + i * 4) = make_float4(v.c[0], v.c[1], v.c[2], v.c[3]);
}

template <typename T, int D>
__device__ inline T sqdist(const Vec<T, D>& x, const Vec<T, D>& y) {
    T s = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const T q = x.c[d] - y.c[d];
        s += q * q;
    }
    return s > eps_of<T>() ? s : eps_of<T>();
}

constexpr int kNegBatch = 4;  // negative samples whose loads are issued together

// Pull-mode epoch: the owner lane of observation i applies every move of i.
// A stored edge i->t fires in exactly the same epochs as its reverse t->i
// (the fuzzy union makes weights, hence epochs_per_sample, bit-symmetric), so
// for each firing edge the lane applies i's head step of i->t and then the
// tail step that t->i gives i (the reference computes it as the negative of
// t's head move, -alpha*clip(c*(y_t - y_i)) = alpha*clip(c*(y_i - y_t))),
// followed by i's negative samples.  Only the owner writes i, once, at the
// end of its row: the epoch is a pure gather with no atomics.
template <typename T, int D, bool TRACE>
__global__ void __launch_bounds__(kBlock) k_layout_pull(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    if (i < v.n) {
        const int64_t s = v.rp[i], e = v.rp[i + 1];
        Vec<T, D> h = load_point<T, D>(v.emb, i);
        const T two_ab = -2 * v.a * v.b;
        const T two_gb = 2 * v.gamma * v.b;
        const uint32_t n32 = static_cast<uint32_t>(v.n);
        bool moved = false;
        for (int64_t j = s; j < e; ++j) {
            const T nx = v.nxt[j];
            if (nx > ep) continue;
            ++fires;
            moved = true;
            const int32_t t = v.tg[j];
            const T ej_eps = v.eps[j];
            const T nn = v.nxt_neg[j];
            const Vec<T, D> r = load_point<T, D>(v.emb, t);
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                const T d2 = sqdist<T, D>(h, r);
                const T pd2b = fast_pow(d2, v.b);
                const T gc = (two_ab * pd2b) / (d2 * (v.a * pd2b + 1));
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
                }
            }

            const T epns = ej_eps / v.nsr;
            const int nneg = static_cast<int>((ep - nn) / epns);
            const uint64_t ej = edge_key(ec.ekey, static_cast<uint64_t>(j));
            for (int p0 = 0; p0 < nneg; p0 += kNegBatch) {
                int32_t kk[kNegBatch];
                Vec<T, D> q[kNegBatch];
#pragma unroll
                for (int u = 0; u < kNegBatch; ++u) {
                    kk[u] = p0 + u < nneg ? draw_index(ej, static_cast<uint32_t>(p0 + u), n32) : -1;
                    q[u] = load_point<T, D>(v.emb, kk[u] >= 0 ? kk[u] : i);
                }
#pragma unroll
                for (int u = 0; u < kNegBatch; ++u) {
                    if (kk[u] < 0) break;
                    ++draws;
                    if (kk[u] == i) {
                        ++self_hits;
                        continue;
                    }
                    const T e2 = sqdist<T, D>(h, q[u]);
                    const T rc = two_gb / ((static_cast<T>(0.001) + e2) * (v.a * fast_pow(e2, v.b) + 1));
                    double* rec = TRACE ? trace_slot(v, 1, i, kk[u]) : nullptr;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const T g = alpha * clip4(rc * (h.c[d] - q[u].c[d]));
                        if (TRACE && rec) {
                            rec[3 + d] = h.c[d];
                            rec[3 + D + d] = q[u].c[d];
                            rec[3 + 2 * D + d] = g;
                        }
                        h.c[d] += g;
                    }
                }
            }
            v.nxt[j] = nx + ej_eps;
            v.nxt_neg[j] = nn + static_cast<T>(nneg) * epns;
        }
        if (moved) store_point<T, D>(v.emb, i, h);
    }
    tally_add(v.tally + 3 * static_cast<int64_t>(ec.epoch), fires, draws, self_hits);
}

// Pull mode for any embedding dimension: the owner updates its point in place.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_layout_pull_generic(LayoutView<T> v, EpochConst ec, T ep, T alpha) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    unsigned long long fires = 0, draws = 0, self_hits = 0;
    if (i < v.n) {
        const int D = v.dim;
        const int64_t s = v.rp[i], e = v.rp[i + 1];
        T* h = v.emb + i * D;
        const uint32_t n32 = static_cast<uint32_t>(v.n);
        for (int64_t j = s; j < e; ++j) {
            const T nx = v.nxt[j];
            if (nx > ep) continue;
            ++fires;
            const T* r = v.emb + static_cast<int64_t>(v.tg[j]) * D;
            for (int half = 0; half < 2; ++half) {
                T d2 = 0;
                for (int d = 0; d < D; ++d) {
                    const T q = h[d] - r[d];
                    d2 += q * q;
                }
                d2 = d2 > eps_of<T>() ? d2 : eps_of<T>();
                const T pd2b = fast_pow(d2, v.b);
                const T gc = (-2 * v.a * v.b * pd2b) / (d2 * (v.a * pd2b + 1));
                for (int d = 0; d < D; ++d) h[d] += alpha * clip4(gc * (h[d] - r[d]));
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
                for (int d = 0; d < D; ++d) h[d] += alpha * clip4(rc * (h[d] - q[d]));
            }
            v.nxt[j] = nx + v.eps[j];
            v.nxt_neg[j] = nn + static_cast<T>(nneg) * epns;
        }
    }
    tally_add(v.tally + 3 * static_cast<int64_t>(ec.epoch), fires, draws, self_hits);
}

}  // namespace

uint64_t counter_epoch_key(uint64_t seed, int epoch) { return mix64(seed + kGold * (static_cast<uint64_t>(epoch) + 1)); }

template <typename T>
void launch_layout_epoch(const LayoutView<T>& v, const EpochArgs& ea, hipStream_t s) {
    const EpochConst ec{ea.epoch, ea.ekey};
    const T ep = static_cast<T>(ea.epoch);
    const T alpha = static_cast<T>(ea.alpha);
    const int64_t blocks = (v.n + kBlock - 1) / kBlock;
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
    } else if (v.trace && v.dim <= 4) {
        switch (v.dim) {
            case 1: k_layout_pull<T, 1, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            case 2: k_layout_pull<T, 2, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            case 3: k_layout_pull<T, 3, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            default: k_layout_pull<T, 4, true><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
        }
    } else {
        switch (v.dim) {
            case 1: k_layout_pull<T, 1, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            case 2: k_layout_pull<T, 2, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            case 3: k_layout_pull<T, 3, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            case 4: k_layout_pull<T, 4, false><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
            default: k_layout_pull_generic<T><<<grid, kBlock, 0, s>>>(v, ec, ep, alpha); break;
        }
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

template void launch_layout_epoch<float>(const LayoutView<float>&, const EpochArgs&, hipStream_t);
template void launch_layout_epoch<double>(const LayoutView<double>&, const EpochArgs&, hipStream_t);
template void launch_layout_serial<float>(const LayoutView<float>&, const EpochArgs&, const int32_t*, int64_t,
                                          unsigned long long*, hipStream_t);
template void launch_layout_serial<double>(const LayoutView<double>&, const EpochArgs&, const int32_t*, int64_t,
                                           unsigned long long*, hipStream_t);

}  // namespace umx
