// Exact k-nearest-neighbour search on the GPU: the producer of the
// NeighborList that initialize() consumes (the reference takes it from
// knncolle, initialize.hpp:138-177; knncolle is not part of this library).
//
// Brute force, one thread per query.  Distances are computed in double with
// the textbook order: s = sum over d = 0.. of (q_d - x_d)^2, a separate
// multiply and add per term (no contraction), so squared distances are bit
// for bit those of a sequential double loop; the k smallest (squared
// distance, index) pairs in lexicographic order are kept, i.e. ties go to the
// smaller index, and the query itself is excluded.  Output rows are sorted by
// increasing distance (sqrt of the double sum, rounded to the caller's type).
//
// Work layout.  A workgroup of 256 queries walks the reference points in
// tiles of kTile points staged in LDS (read as broadcasts); each thread holds
// a chunk of kChunk query coordinates in registers and the kTile partial sums
// (32 and 16 for K = 16, less for larger K),
// so each staged coordinate feeds 256 queries and each query coordinate
// kTile reference points.  Partial sums accumulate chunk after chunk in
// increasing d, which keeps the sequential summation order.  The running
// top-k list (capacity K = 16, 32 or 64) stays in registers, sorted, with
// insertion by an unrolled compare-and-swap pass.
#include "umx_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace umx {
namespace {

constexpr int kQ = 256;       // queries per workgroup
constexpr int kLdsDims = 64;  // coordinates of a tile staged at a time

// kTile reference points per LDS tile and kChunk query coordinates in
// registers, smaller for larger K so that the top-k list fits in registers.
template <typename T, int K, int kTile, int kChunk>
__global__ void __launch_bounds__(kQ) k_knn(int64_t n, int dim, const T* __restrict__ x, int k,
                                             int32_t* __restrict__ out_idx, T* __restrict__ out_dist) {
    __shared__ double tile[kLdsDims][kTile];
    const int64_t qi = static_cast<int64_t>(blockIdx.x) * kQ + threadIdx.x;
    const bool live = qi < n;
    double best_d[K];
    int32_t best_i[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        best_d[q] = INFINITY;
        best_i[q] = 0x7fffffff;
    }
    for (int64_t j0 = 0; j0 < n; j0 += kTile) {
        const int tn = n - j0 < kTile ? static_cast<int>(n - j0) : kTile;
        double part[kTile];
#pragma unroll
        for (int p = 0; p < kTile; ++p) part[p] = 0.0;
        for (int d0 = 0; d0 < dim; d0 += kLdsDims) {
            const int dn = dim - d0 < kLdsDims ? dim - d0 : kLdsDims;
            __syncthreads();
            for (int t = threadIdx.x; t < kLdsDims * kTile; t += kQ) {
                const int dd = t / kTile, p = t % kTile;
                tile[dd][p] = (dd < dn && p < tn) ? static_cast<double>(x[(j0 + p) * dim + d0 + dd]) : 0.0;
            }
            __syncthreads();
            for (int c0 = 0; c0 < dn; c0 += kChunk) {
                double qv[kChunk];
#pragma unroll
                for (int c = 0; c < kChunk; ++c)
                    qv[c] = (live && c0 + c < dn) ? static_cast<double>(x[qi * dim + d0 + c0 + c]) : 0.0;
                const int cn = dn - c0 < kChunk ? dn - c0 : kChunk;
                if (cn == kChunk) {
#pragma unroll
                    for (int c = 0; c < kChunk; ++c) {
#pragma unroll
                        for (int p = 0; p < kTile; ++p) {
                            const double t = qv[c] - tile[c0 + c][p];
                            const double sq = t * t;
                            part[p] = part[p] + sq;
                        }
                    }
                } else {
                    for (int c = 0; c < cn; ++c) {
#pragma unroll
                        for (int p = 0; p < kTile; ++p) {
                            const double t = qv[c] - tile[c0 + c][p];
                            const double sq = t * t;
                            part[p] = part[p] + sq;
                        }
                    }
                }
            }
        }
        if (!live) continue;
#pragma unroll
        for (int p = 0; p < kTile; ++p) {
            const int32_t j = static_cast<int32_t>(j0 + p);
            const double s = part[p];
            const bool better = p < tn && j != qi && (s < best_d[K - 1] || (s == best_d[K - 1] && j < best_i[K - 1]));
            if (better) {
                // insert (s, j) into the sorted list: shift larger entries up
                double cd = s;
                int32_t ci = j;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const bool lt = cd < best_d[q] || (cd == best_d[q] && ci < best_i[q]);
                    const double td = best_d[q];
                    const int32_t ti = best_i[q];
                    best_d[q] = lt ? cd : td;
                    best_i[q] = lt ? ci : ti;
                    cd = lt ? td : cd;
                    ci = lt ? ti : ci;
                }
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        if (q < k) {
            out_idx[qi * k + q] = best_i[q];
            out_dist[qi * k + q] = static_cast<T>(sqrt(best_d[q]));
        }
    }
}

// float input with at most kLdsDims coordinates: the workgroup's queries are
// staged in LDS once (as floats: their conversion to double is exact), so
// the per-tile reload of each query's coordinates reads LDS instead of
// global memory, where a lane-strided reload of 256 rows of up to 256 bytes
// each misses the 32 KB L1 on every tile.  Sums, order and ties exactly as
// in k_knn.
template <int K, int kTile, int kChunk>
__global__ void __launch_bounds__(kQ) k_knn_qlds(int64_t n, int dim, const float* __restrict__ x, int k,
                                                 int32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
    __shared__ double tile[kLdsDims][kTile];
    __shared__ float qs[kLdsDims][kQ];
    const int64_t qi = static_cast<int64_t>(blockIdx.x) * kQ + threadIdx.x;
    const bool live = qi < n;
    for (int d = 0; d < dim; ++d) qs[d][threadIdx.x] = live ? x[qi * dim + d] : 0.0f;  // own column only
    double best_d[K];
    int32_t best_i[K];
#pragma unroll
    for (int q = 0; q < K; ++q) {
        best_d[q] = INFINITY;
        best_i[q] = 0x7fffffff;
    }
    for (int64_t j0 = 0; j0 < n; j0 += kTile) {
        const int tn = n - j0 < kTile ? static_cast<int>(n - j0) : kTile;
        double part[kTile];
#pragma unroll
        for (int p = 0; p < kTile; ++p) part[p] = 0.0;
        __syncthreads();
        for (int t = threadIdx.x; t < kLdsDims * kTile; t += kQ) {
            const int dd = t / kTile, p = t % kTile;
            tile[dd][p] = (dd < dim && p < tn) ? static_cast<double>(x[(j0 + p) * dim + dd]) : 0.0;
        }
        __syncthreads();
        for (int c0 = 0; c0 < dim; c0 += kChunk) {
            double qv[kChunk];
#pragma unroll
            for (int c = 0; c < kChunk; ++c)
                qv[c] = c0 + c < dim ? static_cast<double>(qs[c0 + c][threadIdx.x]) : 0.0;
            const int cn = dim - c0 < kChunk ? dim - c0 : kChunk;
            if (cn == kChunk) {
#pragma unroll
                for (int c = 0; c < kChunk; ++c) {
#pragma unroll
                    for (int p = 0; p < kTile; ++p) {
                        const double t = qv[c] - tile[c0 + c][p];
                        const double sq = t * t;
                        part[p] = part[p] + sq;
                    }
                }
            } else {
                for (int c = 0; c < cn; ++c) {
#pragma unroll
                    for (int p = 0; p < kTile; ++p) {
                        const double t = qv[c] - tile[c0 + c][p];
                        const double sq = t * t;
                        part[p] = part[p] + sq;
                    }
                }
            }
        }
        if (!live) continue;
#pragma unroll
        for (int p = 0; p < kTile; ++p) {
            const int32_t j = static_cast<int32_t>(j0 + p);
            const double s = part[p];
            const bool better = p < tn && j != qi && (s < best_d[K - 1] || (s == best_d[K - 1] && j < best_i[K - 1]));
            if (better) {
                double cd = s;
                int32_t ci = j;
#pragma unroll
                for (int q = 0; q < K; ++q) {
                    const bool lt = cd < best_d[q] || (cd == best_d[q] && ci < best_i[q]);
                    const double td = best_d[q];
                    const int32_t ti = best_i[q];
                    best_d[q] = lt ? cd : td;
                    best_i[q] = lt ? ci : ti;
                    cd = lt ? td : cd;
                    ci = lt ? ti : ci;
                }
            }
        }
    }
    if (!live) return;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        if (q < k) {
            out_idx[qi * k + q] = best_i[q];
            out_dist[qi * k + q] = static_cast<float>(sqrt(best_d[q]));
        }
    }
}

template <typename T>
void launch_knn(int64_t n, int dim, const T* x, int k, int32_t* idx, T* dist, hipStream_t s) {
    if constexpr (std::is_same<T, float>::value) {
        if (dim <= kLdsDims && k <= 32) {
            const unsigned grid = static_cast<unsigned>((n + kQ - 1) / kQ);
            if (k <= 16) k_knn_qlds<16, 32, 16><<<grid, kQ, 0, s>>>(n, dim, x, k, idx, dist);
            else k_knn_qlds<32, 16, 8><<<grid, kQ, 0, s>>>(n, dim, x, k, idx, dist);
            UMX_HIP(hipGetLastError());
            return;
        }
    }
    const unsigned grid = static_cast<unsigned>((n + kQ - 1) / kQ);
    if (k <= 16) k_knn<T, 16, 32, 16><<<grid, kQ, 0, s>>>(n, dim, x, k, idx, dist);
    else if (k <= 32) k_knn<T, 32, 16, 8><<<grid, kQ, 0, s>>>(n, dim, x, k, idx, dist);
    else k_knn<T, 64, 8, 4><<<grid, kQ, 0, s>>>(n, dim, x, k, idx, dist);
    UMX_HIP(hipGetLastError());
}

}  // namespace

template <typename T>
void knn_bruteforce(int64_t n, int dim, const T* x, int k, int32_t* idx, T* dist, hipStream_t s) {
    if (n <= 0) return;
    launch_knn<T>(n, dim, x, k, idx, dist, s);
}

template void knn_bruteforce<float>(int64_t, int, const float*, int, int32_t*, float*, hipStream_t);
template void knn_bruteforce<double>(int64_t, int, const double*, int, int32_t*, double*, hipStream_t);

}  // namespace umx
