// UMAP cross-entropy loss and its gradient (forward / backward kernels of the
// autograd loss).  No reference counterpart: the reference only applies
// per-edge SGD steps (optimize_layout.hpp:145-176); this is the loss whose
// gradient those steps follow.  Over an edge list (head, tail) with
// per-edge coefficients c_e and m_e negative samples:
//
//   L = sum_e c_e [ log(1 + a s^b)
//                   + gamma * sum_{p < m_e, k != head} (log(1 + a s_k^b) - log a - b log s_k) ]
//
// s = max(FLT_EPSILON, |y_head - y_other|^2).  Minus the gradient with
// respect to the head is the reference's attractive step direction, and its
// repulsive direction with s in place of 0.001 + s.  Negative p of edge e is
// the layout's counter-sampler draw for (seed, step, key_e, p): the backward
// pass recomputes the samples instead of storing them.
//
// Arithmetic is in double (the embedding is float): the loss is a sum of
// tens of millions of terms and the gradient of a point often nearly cancels
// over its terms.  One thread per edge; the head's gradient is accumulated
// in registers and summed over the wavefront's lanes with the same head
// before one atomic per run, the tail's and the negative samples' go out with
// double atomics into a scratch buffer, which a last kernel scales by the
// upstream gradient and rounds to float.  An edge's negative samples are
// gathered in batches ahead of their arithmetic and atomics.
//
// Optionally (clip > 0) the backward pass clips each term's gradient with
// respect to the head to [-clip, clip] per coordinate before summing
// (per-sample gradient clipping; the other endpoint receives minus the
// clipped value).  With clip = 4 a gradient step of size alpha is the
// reference's epoch of clipped per-edge steps taken from one snapshot.
#include "layout_device.hpp"
#include "umappp_amd.h"
#include "umx_internal.hpp"

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>
#include <type_traits>

namespace umx {
namespace {

using namespace umx::dev;

constexpr int kMaxLossDim = 16;
constexpr int kLossBlock = 256;

struct LossView {
    int64_t n;
    int dim;
    const float* emb;
    int64_t nnz;
    const int32_t* heads;
    const int32_t* tails;
    const int64_t* keys;
    const float* coef;
    const int32_t* nneg;
    int m;
    const double* ab;  // device: a, b
    double gamma;
    uint64_t ekey;     // counter-sampler key of 'step'
    int* bad;          // set when an index lies outside [0, n)
};

template <int D>
struct DVec {
    double c[D];
};

template <int D>
__device__ inline DVec<D> load_d(const float* emb, int64_t i, int dim) {
    DVec<D> v;
#pragma unroll
    for (int d = 0; d < D; ++d) v.c[d] = d < dim ? static_cast<double>(emb[i * dim + d]) : 0.0;
    return v;
}

template <int D>
__device__ inline double sqd(const DVec<D>& x, const DVec<D>& y) {
    double s = 0;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const double q = x.c[d] - y.c[d];
        s += q * q;
    }
    return s;
}

__device__ inline double block_sum(double v, double* lds) {
#pragma unroll
    for (int o = kWave / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) lds[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int q = 0; q < kLossBlock / kWave; ++q) t += lds[q];
    return t;
}

__device__ inline bool edge_ok(const LossView& v, int64_t e, int64_t& i, int64_t& j) {
    i = v.heads[e];
    j = v.tails[e];
    if (i < 0 || i >= v.n || j < 0 || j >= v.n) {
        atomicOr(v.bad, 1);
        return false;
    }
    return true;
}

// Negative samples of one edge fetched per batch: all draws of a batch and
// their points are loaded before any is used, so the gathers overlap instead
// of waiting on one another (and on the previous sample's atomics).
template <int D>
struct Batch {
    static constexpr int kCount = D <= 4 ? 8 : 2;
};

template <int D>
__device__ inline int gather_batch(const LossView& v, uint64_t ej, uint32_t n32, int64_t i, int p0, int m,
                                   int64_t (&kk)[Batch<D>::kCount], DVec<D> (&yk)[Batch<D>::kCount]) {
    int got = 0;
#pragma unroll
    for (int u = 0; u < Batch<D>::kCount; ++u) {
        kk[u] = p0 + u < m ? static_cast<int64_t>(draw_index(ej, static_cast<uint32_t>(p0 + u), n32)) : -1;
        if (kk[u] >= 0 && kk[u] != i) yk[u] = load_d<D>(v.emb, kk[u], v.dim);
        got += kk[u] >= 0;
    }
    return got;
}

// D = compile-time bound on the embedding dimension (1, 2, 3, 4 or 16).
template <int D>
__global__ void __launch_bounds__(kLossBlock) k_loss_forward(LossView v, double* __restrict__ loss) {
    __shared__ double lds[kLossBlock / kWave];
    constexpr int PF = Batch<D>::kCount;
    const double a = v.ab[0], b = v.ab[1];
    const double feps = FLT_EPSILON;
    const double log_a = log(a);
    const uint32_t n32 = static_cast<uint32_t>(v.n);
    double acc = 0;
    for (int64_t e = static_cast<int64_t>(blockIdx.x) * kLossBlock + threadIdx.x; e < v.nnz;
         e += static_cast<int64_t>(gridDim.x) * kLossBlock) {
        int64_t i, j;
        if (!edge_ok(v, e, i, j)) continue;
        const double c = v.coef ? static_cast<double>(v.coef[e]) : 1.0;
        if (c == 0) continue;  // every term of the edge is scaled by c
        const int m = v.nneg ? v.nneg[e] : v.m;
        const uint64_t ej = edge_key(v.ekey, v.keys ? static_cast<uint64_t>(v.keys[e]) : static_cast<uint64_t>(e));
        const DVec<D> yi = load_d<D>(v.emb, i, v.dim);
        const DVec<D> yj = load_d<D>(v.emb, j, v.dim);
        int64_t kk[PF];
        DVec<D> yk[PF];
        gather_batch<D>(v, ej, n32, i, 0, m, kk, yk);
        double s = sqd<D>(yi, yj);
        s = s > feps ? s : feps;
        double t = log1p(a * exp(b * log(s)));
        for (int p0 = 0; p0 < m; p0 += PF) {
            if (p0 > 0) gather_batch<D>(v, ej, n32, i, p0, m, kk, yk);
#pragma unroll
            for (int u = 0; u < PF; ++u) {
                if (kk[u] < 0) break;
                if (kk[u] == i) continue;
                double s2 = sqd<D>(yi, yk[u]);
                s2 = s2 > feps ? s2 : feps;
                const double ls = log(s2);
                t += v.gamma * (log1p(a * exp(b * ls)) - log_a - b * ls);
            }
        }
        acc += c * t;
    }
    const double tot = block_sum(acc, lds);
    if (threadIdx.x == 0 && tot != 0) unsafeAtomicAdd(loss, tot);
}

// Per-term gradient with respect to the head, clipped to [-clip, clip] per
// coordinate when clip > 0 (the reference's per-step clip,
// optimize_layout.hpp:105-110, applied to minus the step).
__device__ inline double clip_term(double g, double clip) {
    if (clip > 0) g = g < -clip ? -clip : (g > clip ? clip : g);
    return g;
}

// Adds a lane's head gradient to g64 with one atomic per run of equal heads
// in the wavefront (edges in CSR order share heads across ~k neighbouring
// lanes, and atomics to one address serialise).  Runs are maximal groups of
// adjacent lanes with one head; a segmented suffix sum leaves each run's
// total in its first lane, which adds it.  Equal heads in separate runs
// simply add separately.
template <int D>
__device__ inline void add_head_segmented(double* __restrict__ g64, int64_t i, int dim, DVec<D> gi) {
    const int lane = threadIdx.x & (kWave - 1);
    const int64_t prev = __shfl_up(i, 1, kWave);
    const bool start = lane == 0 || prev != i;
    const unsigned long long starts = __ballot(start);
    const unsigned long long later = lane + 1 < kWave ? (starts >> (lane + 1)) << (lane + 1) : 0ULL;
    const int run_end = later ? __ffsll(static_cast<long long>(later)) - 1 : kWave;  // first lane of the next run
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        // lane holds the sum over [lane, min(lane + o, run_end)); lane + o's
        // partial covers the next o lanes of the same run
        const bool take = lane + o < run_end;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const double x = __shfl_down(gi.c[d], o, kWave);
            if (take) gi.c[d] += x;
        }
    }
    if (i >= 0 && start) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            if (d >= dim) break;
            unsafeAtomicAdd(g64 + i * dim + d, gi.c[d]);
        }
    }
}

template <int D>
__global__ void __launch_bounds__(kLossBlock) k_loss_backward(LossView v, double clip, double* __restrict__ g64,
                                                            double* __restrict__ gab) {
    __shared__ double lds[kLossBlock / kWave];
    constexpr int PF = Batch<D>::kCount;
    const double a = v.ab[0], b = v.ab[1];
    const double feps = FLT_EPSILON;
    const uint32_t n32 = static_cast<uint32_t>(v.n);
    const int dim = v.dim;
    double da = 0, db = 0;
    // every lane of a wavefront runs the same number of trips and reaches
    // the head reduction (a wavefront collective: ballot and shuffles) at one
    // call site; lanes past the end, with a bad edge or with a zero
    // coefficient take part with head -1 and a zero gradient
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kLossBlock;
    const int64_t first = static_cast<int64_t>(blockIdx.x) * kLossBlock + (threadIdx.x & ~(kWave - 1));
    for (int64_t e0 = first; e0 < v.nnz; e0 += stride) {
        const int64_t e = e0 + (threadIdx.x & (kWave - 1));
        int64_t i = -1, j = -1;
        DVec<D> gi{};
        bool active = e < v.nnz && edge_ok(v, e, i, j);
        const double c = active && v.coef ? static_cast<double>(v.coef[e]) : 1.0;
        if (c == 0) active = false;  // every term of the edge, and its gradient, is scaled by c
        if (active) {
            const int m = v.nneg ? v.nneg[e] : v.m;
            const uint64_t ej =
                edge_key(v.ekey, v.keys ? static_cast<uint64_t>(v.keys[e]) : static_cast<uint64_t>(e));
            const DVec<D> yi = load_d<D>(v.emb, i, dim);
            const DVec<D> yj = load_d<D>(v.emb, j, dim);
            int64_t kk[PF];
            DVec<D> yk[PF];
            gather_batch<D>(v, ej, n32, i, 0, m, kk, yk);
            {
                double s = sqd<D>(yi, yj);
                const bool clamped = !(s > feps);
                if (clamped) s = feps;
                const double ls = log(s), sb = exp(b * ls), den = 1.0 + a * sb;
                da += c * sb / den;
                db += c * a * sb * ls / den;
                if (!clamped) {
                    const double f = 2.0 * c * a * b * sb / (s * den);
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        if (d >= dim) break;
                        const double g = clip_term(f * (yi.c[d] - yj.c[d]), clip);
                        gi.c[d] += g;
                        unsafeAtomicAdd(g64 + j * dim + d, -g);
                    }
                }
            }
            for (int p0 = 0; p0 < m; p0 += PF) {
                if (p0 > 0) gather_batch<D>(v, ej, n32, i, p0, m, kk, yk);
#pragma unroll
                for (int u = 0; u < PF; ++u) {
                    if (kk[u] < 0) break;
                    if (kk[u] == i) continue;
                    double s = sqd<D>(yi, yk[u]);
                    const bool clamped = !(s > feps);
                    if (clamped) s = feps;
                    const double ls = log(s), sb = exp(b * ls), den = 1.0 + a * sb;
                    da += c * v.gamma * (sb / den - 1.0 / a);
                    db += -c * v.gamma * ls / den;
                    if (!clamped) {
                        const double f = -2.0 * c * v.gamma * b / (s * den);
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            if (d >= dim) break;
                            const double g = clip_term(f * (yi.c[d] - yk[u].c[d]), clip);
                            gi.c[d] += g;
                            unsafeAtomicAdd(g64 + kk[u] * dim + d, -g);
                        }
                    }
                }
            }
        }
        add_head_segmented<D>(g64, active ? i : -1, dim, gi);
    }
    const double ta = block_sum(da, lds);
    const double tb = block_sum(db, lds);
    if (threadIdx.x == 0) {
        if (ta != 0) unsafeAtomicAdd(gab, ta);
        if (tb != 0) unsafeAtomicAdd(gab + 1, tb);
    }
}

__global__ void __launch_bounds__(kLossBlock) k_loss_finish(int64_t count, const double* __restrict__ g64,
                                                          const double* __restrict__ grad_out,
                                                          float* __restrict__ grad_emb, double* __restrict__ gab) {
    const double go = grad_out ? *grad_out : 1.0;
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kLossBlock + threadIdx.x;
    if (q < count) grad_emb[q] = static_cast<float>(go * g64[q]);
    if (q == 0) {
        gab[0] *= go;
        gab[1] *= go;
    }
}

__global__ void k_zero_f64(double* p, int64_t count) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kLossBlock + threadIdx.x;
    if (q < count) p[q] = 0.0;
}
__global__ void k_zero_i32(int* p) { *p = 0; }

// Grid of the grid-stride loss kernels: at most 2048 workgroups (8 per CU),
// so that the one scalar atomic per workgroup (loss, d/da, d/db), which
// serialises at a single address, stays a small cost.
unsigned loss_grid(int64_t nnz) {
    const int64_t want = (nnz + kLossBlock - 1) / kLossBlock;
    return static_cast<unsigned>(want < 1 ? 1 : (want > 2048 ? 2048 : want));
}

template <typename F>
void dispatch_dim(int dim, F&& f) {
    switch (dim) {
        case 1: f(std::integral_constant<int, 1>{}); break;
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 3: f(std::integral_constant<int, 3>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        default: f(std::integral_constant<int, kMaxLossDim>{}); break;
    }
}

struct LossArgs {
    int64_t n;
    int dim;
    const float* emb;
    int64_t nnz;
    const int32_t* heads;
    const int32_t* tails;
    const int64_t* keys;
    const float* coef;
    const int32_t* num_neg;
    int m;
    const double* ab;
    double gamma;
    uint64_t seed;
    int step;
};

void check_args(const LossArgs& x) {
    if (x.n < 0 || x.n >= 0x7fffffffLL) throw Error(UMX_ERR_OVERFLOW, "n must fit a 32-bit index");
    if (x.dim < 1 || x.dim > kMaxLossDim) throw Error(UMX_ERR_ARGUMENT, "num_dim must lie in [1, 16]");
    if (x.nnz < 0 || x.m < 0) throw Error(UMX_ERR_ARGUMENT, "negative edge or sample count");
    if (x.ab == nullptr) throw Error(UMX_ERR_ARGUMENT, "ab is NULL");
    if (x.nnz > 0 && (x.emb == nullptr || x.heads == nullptr || x.tails == nullptr))
        throw Error(UMX_ERR_ARGUMENT, "emb/heads/tails is NULL");
    if (x.nnz > 0 && x.n == 0) throw Error(UMX_ERR_GRAPH, "edges over an empty embedding");
}

LossView view_of(const LossArgs& x, int* bad) {
    LossView v;
    v.n = x.n;
    v.dim = x.dim;
    v.emb = x.emb;
    v.nnz = x.nnz;
    v.heads = x.heads;
    v.tails = x.tails;
    v.keys = x.keys;
    v.coef = x.coef;
    v.nneg = x.num_neg;
    v.m = x.m;
    v.ab = x.ab;
    v.gamma = x.gamma;
    v.ekey = counter_epoch_key(x.seed, x.step);
    v.bad = bad;
    return v;
}

// Error flag of a call: a device int, cleared, checked once the stream is done.
struct BadFlag {
    int* p = nullptr;
    hipStream_t s;
    explicit BadFlag(hipStream_t st) : s(st) {
        prepare_pool();
        UMX_HIP(hipMallocAsync(reinterpret_cast<void**>(&p), sizeof(int), s));
        k_zero_i32<<<1, 1, 0, s>>>(p);
        UMX_HIP(hipGetLastError());
    }
    ~BadFlag() {
        if (p) (void)hipFreeAsync(p, s);
    }
    void check() {
        int h = 0;
        UMX_HIP(hipMemcpyAsync(&h, p, sizeof(int), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        if (h) throw Error(UMX_ERR_GRAPH, "an edge index lies outside [0, n)");
    }
};

void loss_forward(const LossArgs& x, double* loss, hipStream_t s) {
    check_args(x);
    if (loss == nullptr) throw Error(UMX_ERR_ARGUMENT, "loss is NULL");
    BadFlag bad(s);
    k_zero_f64<<<1, kLossBlock, 0, s>>>(loss, 1);
    const LossView v = view_of(x, bad.p);
    if (x.nnz > 0)
        dispatch_dim(x.dim, [&](auto Dc) { k_loss_forward<decltype(Dc)::value><<<loss_grid(x.nnz), kLossBlock, 0, s>>>(v, loss); });
    UMX_HIP(hipGetLastError());
    bad.check();
}

void check_backward_args(const LossArgs& x, double clip, float* grad_emb, double* grad_ab) {
    check_args(x);
    if (grad_ab == nullptr || (x.n > 0 && grad_emb == nullptr)) throw Error(UMX_ERR_ARGUMENT, "output is NULL");
    if (!(clip >= 0)) throw Error(UMX_ERR_ARGUMENT, "clip must be >= 0");
}

// The backward kernels: g64 (n * dim doubles) accumulates the embedding
// gradient and is cleared here first; 'bad' must already be cleared.
void backward_launch(const LossArgs& x, double clip, const double* grad_out, float* grad_emb, double* grad_ab,
                     double* g64, int* bad, hipStream_t s) {
    const int64_t count = x.n * x.dim;
    const unsigned zg = static_cast<unsigned>((count + kLossBlock - 1) / kLossBlock);
    if (count > 0) k_zero_f64<<<zg, kLossBlock, 0, s>>>(g64, count);
    k_zero_f64<<<1, kLossBlock, 0, s>>>(grad_ab, 2);
    const LossView v = view_of(x, bad);
    if (x.nnz > 0)
        dispatch_dim(x.dim, [&](auto Dc) {
            k_loss_backward<decltype(Dc)::value><<<loss_grid(x.nnz), kLossBlock, 0, s>>>(v, clip, g64, grad_ab);
        });
    k_loss_finish<<<count > 0 ? zg : 1, kLossBlock, 0, s>>>(count, g64, grad_out, grad_emb, grad_ab);
}

void loss_backward(const LossArgs& x, double clip, const double* grad_out, float* grad_emb, double* grad_ab,
                   hipStream_t s) {
    check_backward_args(x, clip, grad_emb, grad_ab);
    BadFlag bad(s);
    const int64_t count = x.n * x.dim;
    double* g64 = nullptr;
    UMX_HIP(hipMallocAsync(reinterpret_cast<void**>(&g64), sizeof(double) * static_cast<size_t>(count > 0 ? count : 1), s));
    backward_launch(x, clip, grad_out, grad_emb, grad_ab, g64, bad.p, s);
    const hipError_t e = hipGetLastError();
    (void)hipFreeAsync(g64, s);
    UMX_HIP(e);
    bad.check();
}

// ---- the backward pass without allocation ----
//
// Device memory of the no-allocation entry point: the index-error flag (the
// first 256 bytes) and the double gradient accumulator, for at most
// kNoAllocValues = n * num_dim values.  Each device holds its own copy at its
// own address until the process ends; every call on a device shares it.
constexpr int64_t kNoAllocValues = int64_t(1) << 22;  // 32 MiB of doubles
constexpr int64_t kNoAllocFlagWords = 32;             // 256 bytes
static __device__ __align__(256) double scratch[kNoAllocFlagWords + kNoAllocValues];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of 'scratch' on the current device, looked up once per
// device (host code must never pass the array's own name to a kernel: it is
// a host-side stand-in, not a device address).
double* scratch_address() {
    int dev = 0;
    UMX_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) throw Error(UMX_ERR_ARGUMENT, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || p == nullptr)
            throw Error(UMX_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<double*>(p);
    }
    return g_scratch_ptr[dev];
}

void loss_backward_noalloc(const LossArgs& x, double clip, const double* grad_out, float* grad_emb, double* grad_ab,
                           hipStream_t s) {
    check_backward_args(x, clip, grad_emb, grad_ab);
    if (x.n * x.dim > kNoAllocValues)
        throw Error(UMX_ERR_CAPACITY, "n * num_dim exceeds the no-allocation limit of 4194304 gradient values");
    double* base = scratch_address();
    int* bad = reinterpret_cast<int*>(base);
    k_zero_i32<<<1, 1, 0, s>>>(bad);
    backward_launch(x, clip, grad_out, grad_emb, grad_ab, base + kNoAllocFlagWords, bad, s);
    UMX_HIP(hipGetLastError());
    int h = 0;
    UMX_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    if (h) throw Error(UMX_ERR_GRAPH, "an edge index lies outside [0, n)");
}

}  // namespace
}  // namespace umx

extern "C" {

int umx_loss_forward(int64_t n, int num_dim, const float* embedding, int64_t nnz, const int32_t* heads,
                     const int32_t* tails, const int64_t* keys, const float* coef, const int32_t* num_neg, int m,
                     const double* ab, double gamma, uint64_t seed, int step, double* loss, void* stream) {
    return umx::guarded([&] {
        const umx::LossArgs x{n, num_dim, embedding, nnz, heads, tails, keys, coef, num_neg, m, ab, gamma, seed, step};
        umx::loss_forward(x, loss, static_cast<hipStream_t>(stream));
    });
}

int umx_loss_backward_noalloc(int64_t n, int num_dim, const float* embedding, int64_t nnz, const int32_t* heads,
                              const int32_t* tails, const int64_t* keys, const float* coef, const int32_t* num_neg,
                              int m, const double* ab, double gamma, uint64_t seed, int step, double clip,
                              const double* grad_out, float* grad_embedding, double* grad_ab, void* stream) {
    return umx::guarded([&] {
        const umx::LossArgs x{n, num_dim, embedding, nnz, heads, tails, keys, coef, num_neg, m, ab, gamma, seed, step};
        umx::loss_backward_noalloc(x, clip, grad_out, grad_embedding, grad_ab, static_cast<hipStream_t>(stream));
    });
}

int umx_loss_backward(int64_t n, int num_dim, const float* embedding, int64_t nnz, const int32_t* heads,
                      const int32_t* tails, const int64_t* keys, const float* coef, const int32_t* num_neg, int m,
                      const double* ab, double gamma, uint64_t seed, int step, double clip, const double* grad_out,
                      float* grad_embedding, double* grad_ab, void* stream) {
    return umx::guarded([&] {
        const umx::LossArgs x{n, num_dim, embedding, nnz, heads, tails, keys, coef, num_neg, m, ab, gamma, seed, step};
        umx::loss_backward(x, clip, grad_out, grad_embedding, grad_ab, static_cast<hipStream_t>(stream));
    });
}

}  // extern "C"
