// Spectral initialisation on the GPU (spectral_init.hpp:24-207 of the
// reference): the eigenvectors of the normalized graph Laplacian
// L = I - D^-1/2 W D^-1/2 for its smallest positive eigenvalues give the
// starting coordinates.  The reference shifts to 2I - L, asks irlba for the
// num_dim + 1 largest singular vectors and drops the first (spectral_init.hpp:
// 80-126); this file finds the same eigenvectors of M = 2I - L = I + A,
// A = D^-1/2 W D^-1/2, by Lanczos iteration in double precision:
//
//  * connected components first (spectral_init.hpp:142-168): a graph with
//    several components fails, as in the reference;
//  * Lanczos with full re-orthogonalisation (the three-term recurrence, one
//    classical Gram-Schmidt pass against all previous vectors, and a second
//    pass where the first cancelled most of the direction) builds V (k
//    vectors of length n, resident in HBM) and
//    the tridiagonal T; T's eigen-decomposition (implicit symmetric QR, host)
//    gives Ritz pairs; the Krylov dimension doubles until the num_dim + 1
//    largest Ritz pairs have residuals below tol * theta, else the
//    initialisation fails (the reference fails when irlba does not converge);
//  * the Ritz vectors of eigenvalues 2 .. num_dim + 1 are scaled so that the
//    largest magnitude is 'scale' (spectral_init.hpp:128-137); each vector's
//    sign is fixed by making its largest-magnitude entry positive (the
//    reference's sign is whatever irlba returns);
//  * optional jitter adds N(0, jitter_sd) noise (spectral_init.hpp:190-204)
//    from a counter-based generator (the reference's aarand normals are not
//    reproduced).
#include "layout_device.hpp"
#include "umx_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace umx {
namespace {

using namespace umx::dev;

constexpr int kB = 256;

inline unsigned blocks_for(int64_t n) { return static_cast<unsigned>(n > 0 ? (n + kB - 1) / kB : 1); }

// ---- connected components: hooking to the smaller root + pointer jumping ----
__global__ void k_cc_init(int64_t n, int32_t* parent) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) parent[i] = static_cast<int32_t>(i);
}

__device__ inline int32_t cc_find(const int32_t* parent, int32_t x) {
    int32_t p = parent[x];
    while (p != x) {
        x = p;
        p = parent[x];
    }
    return x;
}

// One round of hooking, G lanes per row: the row's root is found once, each
// lane hooks the larger of it and an edge's root under the smaller.  Parents
// only decrease (no cycles) and rounds repeat until no edge joins two roots,
// so a root that another lane re-hooks meanwhile costs a round, not
// correctness.
__global__ void k_cc_hook(int64_t n, int G, const int64_t* __restrict__ rp, const int32_t* __restrict__ idx,
                          int32_t* parent, int* changed) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    const int64_t i = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (i >= n) return;
    const int32_t a = cc_find(parent, static_cast<int32_t>(i));
    for (int64_t e = rp[i] + lane; e < rp[i + 1]; e += G) {
        const int32_t b = cc_find(parent, idx[e]);
        if (a != b) {
            const int32_t hi = a > b ? a : b, lo = a > b ? b : a;
            atomicMin(parent + hi, lo);
            *changed = 1;
        }
    }
}

__global__ void k_check_indices(int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ idx,
                                int* __restrict__ bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e)
        if (idx[e] < 0 || idx[e] >= n) atomicOr(bad, 1);
}

__global__ void k_cc_compress(int64_t n, int32_t* parent) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) parent[i] = cc_find(parent, static_cast<int32_t>(i));
}

__global__ void k_cc_count(int64_t n, const int32_t* parent, unsigned long long* roots) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    const bool root = i < n && parent[i] == i;
    const unsigned long long b = __ballot(root);
    if ((threadIdx.x & (kWave - 1)) == 0 && b) atomicAdd(roots, static_cast<unsigned long long>(__popcll(b)));
}

// ---- normalized adjacency ----
template <typename T>
__global__ void k_inv_sqrt_degree(int64_t n, const int64_t* __restrict__ rp, const T* __restrict__ val,
                                  double* __restrict__ isd) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    double s = 0;
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e) s += static_cast<double>(val[e]);
    isd[i] = 1.0 / sqrt(s);  // spectral_init.hpp:45-49 (degree > 0 in a connected graph with n > 1)
}

// a = D^-1/2 W D^-1/2 entry by entry, G lanes per row (coalesced).
template <typename T>
__global__ void k_norm_values(int64_t n, int G, const int64_t* __restrict__ rp, const int32_t* __restrict__ idx,
                              const T* __restrict__ val, const double* __restrict__ isd, double* __restrict__ a) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    const int64_t i = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (i >= n) return;
    const double si = isd[i];
    for (int64_t e = rp[i] + lane; e < rp[i + 1]; e += G) a[e] = static_cast<double>(val[e]) * isd[idx[e]] * si;
}

// y = x + A x (M = 2I - L), one row per group of 'G' lanes.
__global__ void k_spmv(int64_t n, int G, const int64_t* __restrict__ rp, const int32_t* __restrict__ idx,
                       const double* __restrict__ a, const double* __restrict__ x, double* __restrict__ y) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    const int64_t i = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    double s = 0;
    if (i < n)
        for (int64_t e = rp[i] + lane; e < rp[i + 1]; e += G) s += a[e] * x[idx[e]];
    for (int o = G / 2; o >= 1; o >>= 1) s += __shfl_xor(s, o, kWave);
    if (i < n && lane == 0) y[i] = x[i] + s;
}

// out[c] += <V[c], w> for c < m (V column-major: vector c at V + c * n; out
// zeroed by the caller).  A workgroup takes a slice of the rows for kDotCols
// columns at once (blockIdx.y picks the group), so w is read once per group
// of columns instead of once per column.
// With 'gate' given, the kernel does nothing unless gate[0] != 0.
constexpr int kDotCols = 8;
__global__ void __launch_bounds__(kB) k_dots(int64_t n, int m, const double* __restrict__ V,
                                             const double* __restrict__ w, double* __restrict__ out,
                                             const int* __restrict__ gate) {
    if (gate && gate[0] == 0) return;  // the same for every thread of the grid
    __shared__ double part[kB / kWave][kDotCols];
    const int c0 = static_cast<int>(blockIdx.y) * kDotCols;
    const int cn = m - c0 < kDotCols ? m - c0 : kDotCols;
    double s[kDotCols];
#pragma unroll
    for (int c = 0; c < kDotCols; ++c) s[c] = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kB) {
        const double wi = w[i];
#pragma unroll
        for (int c = 0; c < kDotCols; ++c)
            if (c < cn) s[c] += V[static_cast<int64_t>(c0 + c) * n + i] * wi;
    }
#pragma unroll
    for (int c = 0; c < kDotCols; ++c)
        for (int o = kWave / 2; o >= 1; o >>= 1) s[c] += __shfl_xor(s[c], o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) {
#pragma unroll
        for (int c = 0; c < kDotCols; ++c) part[threadIdx.x / kWave][c] = s[c];
    }
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < cn) {
        double t = 0;
        for (int q = 0; q < kB / kWave; ++q) t += part[q][threadIdx.x];
        unsafeAtomicAdd(out + c0 + threadIdx.x, t);
    }
}

// Adds the workgroup's sum of v to *out (one atomic per workgroup); every
// thread of the workgroup must call it.
__device__ inline void block_add(double v, double* out) {
    __shared__ double part[kB / kWave];
    for (int o = kWave / 2; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0;
        for (int q = 0; q < kB / kWave; ++q) t += part[q];
        unsafeAtomicAdd(out, t);
    }
}

// w -= sum_c coef[c] * V[c], one row per thread (many wavefronts keep the m
// column loads of each row in flight).  With 'gate' given, the kernel does
// nothing unless gate[0] != 0.
__global__ void __launch_bounds__(kB) k_subtract(int64_t n, int m, const double* __restrict__ V,
                                                 const double* __restrict__ coef, double* __restrict__ w,
                                                 const int* __restrict__ gate) {
    if (gate && gate[0] == 0) return;  // the same for every thread of the grid
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    double s = w[i];
    for (int c = 0; c < m; ++c) s -= coef[c] * V[static_cast<int64_t>(c) * n + i];
    w[i] = s;
}

// At most this many workgroups reduce a sum into one double: their atomic
// adds to that one address are serialised.
constexpr int kNormBlocks = 512;

// The Lanczos three-term recurrence on w = M v1: w -= alpha * v1, and
// w -= beta * v2 when v2 is given (beta = sqrt(*beta2), the norm of the
// direction that became v1); then |w|^2 added to *pre.  Strides over the
// rows with at most kNormBlocks workgroups.
__global__ void __launch_bounds__(kB) k_recurrence(int64_t n, const double* __restrict__ v1,
                                                   const double* __restrict__ alpha, const double* __restrict__ v2,
                                                   const double* __restrict__ beta2, double* __restrict__ w,
                                                   double* __restrict__ pre) {
    const double a = alpha[0];
    const double b = v2 ? sqrt(beta2[0]) : 0.0;
    double sq = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kB) {
        double s = w[i] - a * v1[i];
        if (v2) s -= b * v2[i];
        w[i] = s;
        sq += s * s;
    }
    block_add(sq, pre);
}

// Whether a second Gram-Schmidt pass is needed (the Daniel-Gragg-Kaufman-
// Stewart test): the first pass removed more than 1 - 1/sqrt(2) of |w|.
__global__ void k_reorth_gate(const double* __restrict__ pre, const double* __restrict__ post,
                              int* __restrict__ gate) {
    gate[0] = post[0] < 0.5 * pre[0] ? 1 : 0;
}

// The direction's squared norm after the second pass, where it ran.
__global__ void k_pick_norm(const int* __restrict__ gate, const double* __restrict__ second, double* __restrict__ nrm2) {
    if (gate[0]) nrm2[0] = second[0];
}

// v = w / sqrt(nrm2[0]): the norm stays on the device, so a Lanczos step
// needs no host round trip.
__global__ void k_normalize_into(int64_t n, const double* __restrict__ w, const double* __restrict__ nrm2,
                                 double* __restrict__ v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) v[i] = w[i] * (1.0 / sqrt(nrm2[0]));
}

__global__ void k_fill_start(int64_t n, uint64_t seed, double* __restrict__ v) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i < n) v[i] = static_cast<double>(mix64(seed ^ (static_cast<uint64_t>(i) * kOdd)) >> 11) * 0x1.0p-53 - 0.5;
}

// U[i * d + c] = sum_j V[j][i] * S[j][c]  (Ritz vectors, point-major)
__global__ void k_ritz(int64_t n, int k, int d, const double* __restrict__ V, const double* __restrict__ S,
                       double* __restrict__ U) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    double acc[8];
    for (int c0 = 0; c0 < d; c0 += 8) {
        const int cn = d - c0 < 8 ? d - c0 : 8;
        for (int c = 0; c < cn; ++c) acc[c] = 0;
        for (int j = 0; j < k; ++j) {
            const double vj = V[static_cast<int64_t>(j) * n + i];
            for (int c = 0; c < cn; ++c) acc[c] += vj * S[j * d + c0 + c];
        }
        for (int c = 0; c < cn; ++c) U[i * d + c0 + c] = acc[c];
    }
}

// Entry of largest magnitude per column (value with sign): the magnitude
// bits' maximum (|v| >= 0 orders like its bits) over a grid of at most
// kNormBlocks workgroups striding over the rows, one atomic per workgroup and
// column (the atomics of many workgroups on one address serialise), then a
// second pass that records a signed value.
__global__ void __launch_bounds__(kB) k_col_absmax(int64_t n, int d, const double* __restrict__ U,
                                                   unsigned long long* __restrict__ mag) {
    __shared__ unsigned long long part[kB / kWave];
    for (int c = 0; c < d; ++c) {
        unsigned long long b = 0;
        for (int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x; i < n;
             i += static_cast<int64_t>(gridDim.x) * kB) {
            const double v = fabs(U[i * d + c]);
            unsigned long long t;
            __builtin_memcpy(&t, &v, sizeof(t));
            b = t > b ? t : b;
        }
        for (int o = kWave / 2; o >= 1; o >>= 1) {
            const unsigned long long t = __shfl_xor(b, o, kWave);
            b = t > b ? t : b;
        }
        if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long m = 0;
            for (int q = 0; q < kB / kWave; ++q) m = part[q] > m ? part[q] : m;
            if (m) atomicMax(mag + c, m);
        }
        __syncthreads();  // part is reused for the next column
    }
}

__global__ void k_col_sign(int64_t n, int d, const double* __restrict__ U, const unsigned long long* __restrict__ mag,
                           int* __restrict__ neg) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < d; ++c) {
        const double v = U[i * d + c];
        const double av = fabs(v);
        unsigned long long b;
        __builtin_memcpy(&b, &av, sizeof(b));
        if (b == mag[c]) atomicMax(neg + c, v < 0 ? 1 : 0);  // ties: negative wins, deterministic either way
    }
}

template <typename T>
__global__ void k_emit(int64_t n, int d, const double* __restrict__ U, const double* __restrict__ colmul, int jitter,
                       double jitter_sd, uint64_t seed, T* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    if (i >= n) return;
    for (int c = 0; c < d; ++c) {
        const int64_t q = i * d + c;
        double v = U[q] * colmul[c];
        if (jitter) {
            // Box-Muller on two counter-based uniforms in (0, 1]
            const uint64_t h = mix64(seed + kGold * (static_cast<uint64_t>(q) + 1));
            const double u1 = (static_cast<double>(h >> 11) + 1.0) * 0x1.0p-53;
            const double u2 = static_cast<double>(mix64(h ^ kOdd) >> 11) * 0x1.0p-53;
            v += jitter_sd * sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        }
        out[q] = static_cast<T>(v);
    }
}

}  // namespace

int64_t count_components(int64_t n, const int64_t* rp, const int32_t* idx, hipStream_t s) {
    if (n == 0) return 0;
    DevBuf<int32_t> parent;
    parent.alloc(static_cast<size_t>(n));
    DevBuf<int> changed;
    changed.alloc(1);
    DevBuf<unsigned long long> roots;
    roots.alloc(1);
    UMX_HIP(hipMemsetAsync(changed.p, 0, sizeof(int), s));
    k_check_indices<<<blocks_for(n), kB, 0, s>>>(n, rp, idx, changed.p);
    int bad = 0;
    int64_t nnz = 0;
    UMX_HIP(hipMemcpyAsync(&bad, changed.p, sizeof(int), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipMemcpyAsync(&nnz, rp + n, sizeof(nnz), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    if (bad) throw Error(-2, "neighbour index out of range");
    const int G = lanes_for_degree(std::max<int64_t>(1, (nnz + n - 1) / n));  // from the average degree
    k_cc_init<<<blocks_for(n), kB, 0, s>>>(n, parent.p);
    for (int round = 0; round < 10000; ++round) {
        int h = 0;
        UMX_HIP(hipMemsetAsync(changed.p, 0, sizeof(int), s));
        k_cc_hook<<<blocks_for(n * G), kB, 0, s>>>(n, G, rp, idx, parent.p, changed.p);
        k_cc_compress<<<blocks_for(n), kB, 0, s>>>(n, parent.p);
        UMX_HIP(hipGetLastError());
        UMX_HIP(hipMemcpyAsync(&h, changed.p, sizeof(int), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        if (!h) break;
    }
    UMX_HIP(hipMemsetAsync(roots.p, 0, sizeof(unsigned long long), s));
    k_cc_count<<<blocks_for(n), kB, 0, s>>>(n, parent.p, roots.p);
    UMX_HIP(hipGetLastError());
    unsigned long long r = 0;
    UMX_HIP(hipMemcpyAsync(&r, roots.p, sizeof(r), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    return static_cast<int64_t>(r);
}

template <typename T>
bool spectral_embedding(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, int max_degree, int d,
                        const SpectralParams& prm, T* out, SpectralReport* rep, hipStream_t s) {
    if (rep) *rep = SpectralReport{};
    const int want = d + 1;
    if (n < 2 || want >= n) return false;
    const int64_t comps = count_components(n, rp, idx, s);
    if (rep) rep->components = comps;
    if (comps != 1) return false;
    int64_t nnz = 0;
    UMX_HIP(hipMemcpyAsync(&nnz, rp + n, sizeof(nnz), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));

    // Algorithmic HBM bytes per launch (every array counted once per read,
    // gathered vectors once): the report's roofline figure.
    const double nb = 8.0 * static_cast<double>(n);
    double bytes = 0;

    DevBuf<double> isd, a;
    isd.alloc(static_cast<size_t>(n));
    a.alloc(static_cast<size_t>(nnz));
    k_inv_sqrt_degree<T><<<blocks_for(n), kB, 0, s>>>(n, rp, val, isd.p);
    const int G = lanes_for_degree(std::max<int64_t>(1, (nnz + n - 1) / n));  // from the average degree
    k_norm_values<T><<<blocks_for(n * G), kB, 0, s>>>(n, G, rp, idx, val, isd.p, a.p);
    UMX_HIP(hipGetLastError());
    bytes += (8.0 * (n + 1) + sizeof(T) * nnz + nb) + (8.0 * (n + 1) + (4.0 + sizeof(T) + 8.0) * nnz + nb);
    (void)max_degree;
    const double spmv_bytes = 8.0 * (n + 1) + 12.0 * static_cast<double>(nnz) + 2.0 * nb;

    // One Lanczos run, extended when the Ritz pairs have not converged: the
    // Krylov dimension k doubles (up to max_dim) and the run continues from
    // where it stopped, so no step is computed twice.  alpha_j = <v_j, M v_j>
    // and the squared norm of each orthogonalised direction stay on the
    // device; the host reads them once per convergence check.
    int k = static_cast<int>(std::min<int64_t>(n, std::max(prm.start_dim, 2 * want + 8)));
    const int kmax = static_cast<int>(std::min<int64_t>(n, prm.max_dim));
    const int kcap = std::max(k, kmax);
    DevBuf<double> w, coef, sc;
    DevBuf<int> gates;
    w.alloc(static_cast<size_t>(n));
    coef.alloc(static_cast<size_t>(kcap) + 1);
    sc.alloc(4 * (static_cast<size_t>(kcap) + 1));
    gates.alloc(static_cast<size_t>(kcap) + 1);
    double* const d_alpha = sc.p;                  // alpha_j at [j]
    double* const d_nrm2 = sc.p + (kcap + 1);      // squared norm of direction j at [j]
    double* const d_pre = sc.p + 2 * (kcap + 1);   // |w|^2 after the recurrence, step j
    double* const d_nrm2b = sc.p + 3 * (kcap + 1); // |w|^2 after a second pass, step j
    UMX_HIP(hipMemsetAsync(sc.p, 0, sizeof(double) * sc.n, s));  // each slot is accumulated once
    UMX_HIP(hipMemsetAsync(gates.p, 0, sizeof(int) * gates.n, s));
    DevBuf<double> V;
    V.alloc(static_cast<size_t>(k) * static_cast<size_t>(n));
    // v0: a fixed pseudo-random start (normalised as direction 0)
    k_fill_start<<<blocks_for(n), kB, 0, s>>>(n, prm.start_seed, w.p);
    bytes += nb;
    const unsigned gx = static_cast<unsigned>(std::min<int64_t>((n + kB - 1) / kB, 256));
    const unsigned gnorm = static_cast<unsigned>(std::min<int64_t>((n + kB - 1) / kB, kNormBlocks));
    auto dots = [&](int m, const double* VV, const double* ww, double* to, const int* gate) {
        k_dots<<<dim3(gx, static_cast<unsigned>((m + kDotCols - 1) / kDotCols)), kB, 0, s>>>(n, m, VV, ww, to,
                                                                                             gate);
        if (!gate) bytes += (m + 1) * nb;
    };
    int j = 0;               // the step whose direction w holds
    bool orthogonal = false; // w already orthogonalised against V[0..j) (after an extension)
    int gates_read = 1;      // steps whose second-pass decision is counted in 'bytes'
    std::vector<double> ha, hn, beta;
    std::vector<int> hg;
    while (true) {
        for (;; ++j) {
            if (!orthogonal) {
                if (j == 0) {
                    dots(1, w.p, w.p, d_nrm2, nullptr);
                } else {
                    // w = M v_{j-1}: the three-term recurrence, one classical
                    // Gram-Schmidt pass against V[0..j), and a second pass
                    // only where the first cancelled most of w
                    const double* v1 = V.p + static_cast<int64_t>(j - 1) * n;
                    const double* v2 = j >= 2 ? V.p + static_cast<int64_t>(j - 2) * n : nullptr;
                    k_recurrence<<<gnorm, kB, 0, s>>>(n, v1, d_alpha + (j - 1), v2, d_nrm2 + (j - 1), w.p,
                                                      d_pre + j);
                    bytes += (v2 ? 4 : 3) * nb;
                    for (int pass = 0; pass < 2; ++pass) {
                        const int* gate = pass == 0 ? nullptr : gates.p + j;
                        UMX_HIP(hipMemsetAsync(coef.p, 0, sizeof(double) * static_cast<size_t>(j), s));
                        dots(j, V.p, w.p, coef.p, gate);
                        k_subtract<<<blocks_for(n), kB, 0, s>>>(n, j, V.p, coef.p, w.p, gate);
                        dots(1, w.p, w.p, pass == 0 ? d_nrm2 + j : d_nrm2b + j, gate);
                        if (pass == 0) {
                            bytes += (j + 2) * nb;
                            k_reorth_gate<<<1, 1, 0, s>>>(d_pre + j, d_nrm2 + j, gates.p + j);
                        }
                    }
                    k_pick_norm<<<1, 1, 0, s>>>(gates.p + j, d_nrm2b + j, d_nrm2 + j);
                }
                UMX_HIP(hipGetLastError());
            }
            orthogonal = false;
            if (j == k) break;
            double* vj = V.p + static_cast<int64_t>(j) * n;
            k_normalize_into<<<blocks_for(n), kB, 0, s>>>(n, w.p, d_nrm2 + j, vj);
            // w = M v_j, alpha_j = <v_j, w>
            k_spmv<<<blocks_for(n * G), kB, 0, s>>>(n, G, rp, idx, a.p, vj, w.p);
            UMX_HIP(hipGetLastError());
            bytes += 2 * nb + spmv_bytes;
            dots(1, vj, w.p, d_alpha + j, nullptr);
        }
        ha.resize(static_cast<size_t>(k));
        hn.resize(static_cast<size_t>(k) + 1);
        hg.resize(static_cast<size_t>(k) + 1);
        UMX_HIP(hipMemcpyAsync(ha.data(), d_alpha, sizeof(double) * ha.size(), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipMemcpyAsync(hn.data(), d_nrm2, sizeof(double) * hn.size(), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipMemcpyAsync(hg.data(), gates.p, sizeof(int) * hg.size(), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        for (; gates_read <= k; ++gates_read) {
            if (hg[gates_read]) {
                bytes += (2.0 * gates_read + 5) * nb;
                if (rep) ++rep->second_passes;
            }
        }
        // beta_{j-1} = |w_j|; a vanishing direction means an invariant
        // subspace (T is exact): the steps after it are discarded
        beta.assign(static_cast<size_t>(k), 0.0);
        int kk = k;
        bool broke = false;
        for (int jj = 0; jj <= k; ++jj) {
            const double nrm = std::sqrt(hn[jj]);
            if (jj > 0) beta[jj - 1] = nrm;
            if (jj == k) break;
            if (!(nrm > 1e-300) || (jj > 0 && nrm < 1e-12 * std::fabs(ha[jj - 1]))) {
                broke = true;
                kk = jj;
                break;
            }
        }
        if (rep) ++rep->attempts;
        if (kk < want) return false;
        std::vector<double> dd(ha.begin(), ha.begin() + kk), ee(beta.begin(), beta.begin() + (kk - 1)), z;
        if (!tridiagonal_eigen(kk, dd, ee, z)) return false;
        // the 'want' largest Ritz values and their residuals beta_k * |z[k-1][c]|
        std::vector<int> order(static_cast<size_t>(kk));
        for (int c = 0; c < kk; ++c) order[c] = c;
        std::sort(order.begin(), order.end(), [&](int x, int y) { return dd[x] > dd[y]; });
        const double bk = broke ? 0.0 : beta[kk - 1];
        bool ok = true;
        double worst = 0;
        for (int c = 0; c < want; ++c) {
            const double res = bk * std::fabs(z[static_cast<size_t>(kk - 1) * kk + order[c]]);
            const double rel = res / std::max(std::fabs(dd[order[c]]), 1e-300);
            worst = std::max(worst, rel);
            if (rel > prm.tol) ok = false;
        }
        if (rep) {
            rep->krylov_dim = kk;
            rep->worst_residual = worst;
            rep->bytes = bytes;
            rep->eigenvalues.clear();
            for (int c = 0; c < want; ++c) rep->eigenvalues.push_back(dd[order[c]]);
        }
        if (!ok) {
            if (k >= kmax) return false;
            // extend: room for the next k vectors, then continue at step k,
            // whose direction w is already orthogonalised (norm in d_nrm2[k])
            const int k2 = std::min(kmax, 2 * k);
            DevBuf<double> V2;
            V2.alloc(static_cast<size_t>(k2) * static_cast<size_t>(n));
            UMX_HIP(hipMemcpyAsync(V2.p, V.p, sizeof(double) * static_cast<size_t>(k) * static_cast<size_t>(n),
                                   hipMemcpyDeviceToDevice, s));
            bytes += 2.0 * k * nb;
            V.swap(V2);
            k = k2;
            orthogonal = true;
            continue;
        }
        // Ritz vectors of eigenvalues 2 .. want (drop the largest)
        std::vector<double> S(static_cast<size_t>(kk) * d);
        for (int jj = 0; jj < kk; ++jj)
            for (int c = 0; c < d; ++c) S[static_cast<size_t>(jj) * d + c] = z[static_cast<size_t>(jj) * kk + order[c + 1]];
        DevBuf<double> dS, U;
        dS.alloc(S.size());
        U.alloc(static_cast<size_t>(n) * d);
        UMX_HIP(hipMemcpyAsync(dS.p, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice, s));
        k_ritz<<<blocks_for(n), kB, 0, s>>>(n, kk, d, V.p, dS.p, U.p);
        DevBuf<unsigned long long> mag;
        DevBuf<int> neg;
        mag.alloc(static_cast<size_t>(d));
        neg.alloc(static_cast<size_t>(d));
        UMX_HIP(hipMemsetAsync(mag.p, 0, sizeof(unsigned long long) * d, s));
        UMX_HIP(hipMemsetAsync(neg.p, 0, sizeof(int) * d, s));
        k_col_absmax<<<gnorm, kB, 0, s>>>(n, d, U.p, mag.p);
        k_col_sign<<<blocks_for(n), kB, 0, s>>>(n, d, U.p, mag.p, neg.p);
        UMX_HIP(hipGetLastError());
        std::vector<unsigned long long> hmag(static_cast<size_t>(d));
        std::vector<int> hneg(static_cast<size_t>(d));
        UMX_HIP(hipMemcpyAsync(hmag.data(), mag.p, sizeof(unsigned long long) * d, hipMemcpyDeviceToHost, s));
        UMX_HIP(hipMemcpyAsync(hneg.data(), neg.p, sizeof(int) * d, hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        // spectral_init.hpp:128-131: one expansion for all columns
        double max_val = 0;
        for (int c = 0; c < d; ++c) {
            double v;
            __builtin_memcpy(&v, &hmag[c], sizeof(v));
            max_val = std::max(max_val, v);
        }
        const double expansion = max_val > 0 ? prm.scale / max_val : 1.0;
        std::vector<double> colmul(static_cast<size_t>(d));
        for (int c = 0; c < d; ++c) colmul[c] = hneg[c] ? -expansion : expansion;
        DevBuf<double> dmul;
        dmul.alloc(static_cast<size_t>(d));
        UMX_HIP(hipMemcpyAsync(dmul.p, colmul.data(), sizeof(double) * d, hipMemcpyHostToDevice, s));
        k_emit<T><<<blocks_for(n), kB, 0, s>>>(n, d, U.p, dmul.p, prm.jitter ? 1 : 0, prm.jitter_sd, prm.jitter_seed,
                                              out);
        UMX_HIP(hipGetLastError());
        bytes += (kk + d) * nb + d * nb + d * nb + d * nb + sizeof(T) * d * static_cast<double>(n);
        if (rep) rep->bytes = bytes;
        UMX_HIP(hipStreamSynchronize(s));
        return true;
    }
}

template bool spectral_embedding<float>(int64_t, const int64_t*, const int32_t*, const float*, int, int,
                                        const SpectralParams&, float*, SpectralReport*, hipStream_t);
template bool spectral_embedding<double>(int64_t, const int64_t*, const int32_t*, const double*, int, int,
                                         const SpectralParams&, double*, SpectralReport*, hipStream_t);

}  // namespace umx
