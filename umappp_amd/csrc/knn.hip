// Exact k-nearest-neighbour search on the GPU: the producer of the
// NeighborList that initialize() consumes (the reference takes it from
// knncolle, initialize.hpp:138-177; knncolle is not part of this library).
//
// The answer is the exact brute force's.  Distances are computed in double
// with the textbook order: s = sum over d = 0.. of (q_d - x_d)^2, a separate
// multiply and add per term (no contraction), so squared distances are bit
// for bit those of a sequential double loop; the k smallest (squared
// distance, index) pairs in lexicographic order are kept, i.e. ties go to the
// smaller index, and the query itself is excluded.  Output rows are sorted by
// increasing distance (sqrt of the double sum, rounded to the caller's type).
//
// Two paths.  Float input of at most 64 coordinates with k <= 16 and at
// least 512 points: candidates from the matrix cores, exact distances for
// the candidates only (knn_mfma, below).  Otherwise every distance is
// computed exactly, one thread per query:
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

// ---------------------------------------------------------------------------
// Float input, at most 64 coordinates, k <= 16: candidates from the matrix
// cores, exact distances for the candidates only.
//
// Filter.  The data are centred on their mean (mu, rounded to float; any
// centre would do, the mean keeps the norms small).  For a query q and a
// reference point r with centred float vectors q', r', the matrix cores
// (v_mfma_f32_16x16x4f32) give G ~ q'.r', and the approximate squared
// distance is a = fma(-2, G, (|q'|^2 + |r'|^2)), the norms computed in double
// and rounded to float.  Against the exact squared distance e (the
// sequential double sum the output must equal), |a - e| <= Delta_q for every
// r, where
//   Delta_q = (S4 + 8) u (|q'|^2 + max_r |r'|^2) (1.01),  u = 2^-24,
// S4 the padded dimension: the dot product's rounding (gamma_S4 of
// sum |q'_i r'_i| <= (|q'|^2 + |r'|^2) / 2, any summation order), the norms'
// rounding to float, the sum and the fused step (2u of the same), and the
// centring (each centred coordinate within u |q_i - mu_i| of the exact
// difference: 4u (|q'|^2 + |r'|^2) on the squared distance).  The window
// below is 2 Delta_q wide; the kernels make it twice that (filter_margin2).
//
// Window.  If tau is the k-th smallest approximate distance over all r != q,
// every true k nearest neighbour p (ties to the smaller index included) has
// a_p <= e_p + Delta <= e_k + Delta <= tau + 2 Delta (the k points of
// smallest approximate distance have exact distances <= tau + Delta).  So the
// k nearest neighbours are among the points with a <= W = tau + 2 Delta, and
// the exact sequential double distances of those, ranked by (distance,
// index), give exactly the brute-force answer.
//
// Collection.  A wavefront takes 16 queries (the B operand, one column each)
// and walks all reference points, 128 at a time staged in LDS; each MFMA
// tile is 16 points (the A operand) by 16 queries, and lane l holds the
// products of query l % 16 with 4 of the tile's points, so each query's
// points are split over 4 lanes.  Each lane keeps the 24 smallest
// approximate distances of its points, sorted, and drops a point once it is
// above its window (its own 17th smallest + 2 Delta: the query itself may be
// among its points and enters the list like any other, and at least 16 of
// the 17 are other points, so that value is never below the k-th smallest
// over all other points, k <= 16, and the window contains W), recording the
// smallest value it ever dropped from below its window.  The ranking skips
// the query itself.  A query whose lanes dropped nothing at or below W has
// every point with a <= W among its candidates; a query with such a drop
// (many points within 2 Delta of one another: ties on a grid, duplicates) is
// searched again exactly by k_knn_one.  Points past the end of the data have
// an infinite norm and never enter a list.
// ---------------------------------------------------------------------------
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kMfmaQ = 64;      // queries per workgroup (16 per wavefront)
constexpr int kMfmaRB = 128;    // reference points per LDS block
constexpr int kMfmaList = 24;   // candidates a lane keeps (> kMfmaK; 32 made the 1M x 50 search 3% slower)
constexpr int kMfmaK = 16;      // largest k of the path (the lane window's rank)
static_assert(kMfmaList > kMfmaK, "a lane's list must hold its window's rank");
constexpr float kUnit = 5.9604644775390625e-08f;  // 2^-24

// Row stride (floats) of the LDS block: the A operand of one MFMA step reads
// rows jj = 0..15 at column g S + s for lane groups g = 0..3, and the stride is
// the smallest one from 4 S + 1 up for which those 64 addresses fall in
// distinct banks of a 64-bank LDS and, within each half-wavefront, of a
// 32-bank one (an odd stride of 4 S + 1 gave 2-way conflicts between lane
// groups).
constexpr bool conflict_free(int rs, int s, int banks, int g0, int g1) {
    for (int a = 0; a < 64; ++a) {
        const int ga = a >> 4, ja = a & 15;
        if (ga < g0 || ga > g1) continue;
        for (int b = a + 1; b < 64; ++b) {
            const int gb = b >> 4, jb = b & 15;
            if (gb < g0 || gb > g1) continue;
            if ((rs * ja + s * ga) % banks == (rs * jb + s * gb) % banks) return false;
        }
    }
    return true;
}
constexpr int lds_row_stride(int s) {
    for (int rs = 4 * s + 1; rs <= 4 * s + 64; ++rs)
        if (conflict_free(rs, s, 64, 0, 3) && conflict_free(rs, s, 32, 0, 1) && conflict_free(rs, s, 32, 2, 3))
            return rs;
    for (int rs = 4 * s + 1; rs <= 4 * s + 64; ++rs)
        if (conflict_free(rs, s, 32, 0, 1) && conflict_free(rs, s, 32, 2, 3)) return rs;
    return 4 * s + 1;
}

// Width of the window above the k-th smallest approximate distance: 2 Delta_q
// as derived above, doubled for safety.
__device__ inline float filter_margin2(int s4, float qn, float nr_max) {
    return 4.0f * (static_cast<float>(s4) + 8.0f) * kUnit * (qn + nr_max) * 1.01f;
}

// Column sums of x in double, blocks of kQ points: partial[b * dim + d].
__global__ void __launch_bounds__(kQ) k_colsum_partial(int64_t n, int dim, const float* __restrict__ x,
                                                       double* __restrict__ partial) {
    __shared__ double red[kQ];
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kQ + threadIdx.x;
    for (int d = 0; d < dim; ++d) {
        red[threadIdx.x] = i < n ? static_cast<double>(x[i * dim + d]) : 0.0;
        __syncthreads();
        for (int h = kQ / 2; h >= 1; h >>= 1) {
            if (static_cast<int>(threadIdx.x) < h) red[threadIdx.x] += red[threadIdx.x + h];
            __syncthreads();
        }
        if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * dim + d] = red[0];
        __syncthreads();
    }
}

// mu[d] = (sum of the partials in order) / n, rounded to float; one thread
// per coordinate.
__global__ void k_colmean(int64_t blocks, int64_t n, int dim, const double* __restrict__ partial,
                          float* __restrict__ mu) {
    const int d = static_cast<int>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (d >= dim) return;
    double s = 0;
    for (int64_t b = 0; b < blocks; ++b) s += partial[b * dim + d];
    mu[d] = static_cast<float>(s / static_cast<double>(n));
}

// Squared norms of the centred points (double, rounded to float) and their
// largest value (as ordered bits: norms are >= 0; a non-finite norm makes the
// largest +inf, which makes every query's margin infinite and sends it to
// the exact search).
__global__ void __launch_bounds__(kQ) k_centred_norms(int64_t n, int dim, const float* __restrict__ x,
                                                      const float* __restrict__ mu, float* __restrict__ nrm,
                                                      unsigned int* __restrict__ nr_max_bits) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kQ + threadIdx.x;
    float v = 0.0f;
    if (i < n) {
        double s = 0;
        for (int d = 0; d < dim; ++d) {
            const double c = static_cast<double>(x[i * dim + d] - mu[d]);
            s += c * c;
        }
        v = static_cast<float>(s);
        if (!(v <= 3.0e38f)) v = INFINITY;
        nrm[i] = v;
    }
    unsigned int b = __float_as_uint(v);
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        const unsigned int o = __shfl_xor(b, m, kWave);
        b = o > b ? o : b;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) atomicMax(nr_max_bits, b);
}

// The candidate collection (see above).  S = padded dimension / 4 (1..16);
// lane group g = l / 16 supplies coordinates g * S .. g * S + S - 1.
template <int S>
__global__ void __launch_bounds__(256, 2) k_knn_mfma(int64_t n, int dim, const float* __restrict__ x,
                                                  const float* __restrict__ mu, const float* __restrict__ nrm,
                                                  const unsigned int* __restrict__ nr_max_bits,
                                                  float* __restrict__ cand_v, int32_t* __restrict__ cand_i,
                                                  float* __restrict__ dropped) {
    constexpr int S4 = 4 * S;
    constexpr int RS = lds_row_stride(S);
    // values of a block each thread stages: the block's rows are one
    // contiguous run of kMfmaRB * dim floats, read flat
    constexpr int kPer = (kMfmaRB * S4 + 255) / 256;
    __shared__ float rt[kMfmaRB * RS + kWave];  // + a spare word per lane for stores past the block
    __shared__ float rn[kMfmaRB];
    __shared__ float mus[4 * 16];
    const int w = threadIdx.x >> 6, l = threadIdx.x & (kWave - 1);
    const int g = l >> 4, jj = l & 15;
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kMfmaQ + w * 16 + jj;
    const bool qlive = q < n;
    if (threadIdx.x < S4) mus[threadIdx.x] = static_cast<int>(threadIdx.x) < dim ? mu[threadIdx.x] : 0.0f;
    // the padding columns (dim .. S4 - 1) of every row stay zero
    for (int t = threadIdx.x; t < kMfmaRB * (S4 - dim); t += 256) {
        const int ri = t / (S4 - dim);
        rt[ri * RS + dim + (t - ri * (S4 - dim))] = 0.0f;
    }
    float bq[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int d = g * S + s;
        bq[s] = (qlive && d < dim) ? x[q * dim + d] - mu[d] : 0.0f;
    }
    const float qn = qlive ? nrm[q] : 0.0f;
    const float margin2 = filter_margin2(S4, qn, __uint_as_float(nr_max_bits[0]));
    float lv[kMfmaList];
    int32_t li[kMfmaList];
#pragma unroll
    for (int m = 0; m < kMfmaList; ++m) {
        lv[m] = INFINITY;
        li[m] = -1;
    }
    float drop = INFINITY, win = INFINITY;
    // Each block's values are loaded into registers while the block before
    // is computed (all loads of a block in flight together), then centred
    // into LDS.  Slot u of a thread holds value threadIdx.x + 256 u of the
    // block.  Loads and stores take no branch: a load reads a clamped address
    // and selects, and a lane whose value lies past the block stores into its
    // own spare word.  (With a guard on each, every load and store was a
    // branch of its own, the centring waited for each mean in turn, and
    // staging a block took as long as its 104 MFMAs per wavefront.)
    const int blk_vals = kMfmaRB * dim;
    const int step_r = 256 / dim, step_d = 256 - (256 / dim) * dim;
    float pre[kPer];
    float pre_n = 0.0f;
    auto load_block = [&](int64_t r0) {
        const float* xb = x + r0 * dim;
        const int avail = static_cast<int>((n - r0 < kMfmaRB ? n - r0 : kMfmaRB) * dim);
        if ((n - r0) * dim >= static_cast<int64_t>(kPer) * 256) {
            // every slot inside x: plain loads at fixed offsets (values past
            // the block are never stored)
#pragma unroll
            for (int u = 0; u < kPer; ++u) pre[u] = xb[threadIdx.x + u * 256];
        } else {
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const int e = static_cast<int>(threadIdx.x) + u * 256;
                const bool ok = e < avail;
                const float v = xb[ok ? e : 0];  // row r0 exists: r0 < n
                pre[u] = ok ? v : 0.0f;
            }
        }
        const int64_t rj = r0 + static_cast<int64_t>(threadIdx.x);
        const bool okn = threadIdx.x < kMfmaRB && rj < n;
        const float vn = nrm[okn ? rj : r0];
        pre_n = okn ? vn : INFINITY;
    };
    load_block(0);
    for (int64_t r0 = 0; r0 < n; r0 += kMfmaRB) {
        __syncthreads();
        {
            // (row, column) of value e = threadIdx.x + 256 u, stepped from u
            // to u + 1 without a division
            int ri = static_cast<int>(threadIdx.x) / dim, d = static_cast<int>(threadIdx.x) - ri * dim;
#pragma unroll
            for (int u = 0; u < kPer; ++u) {
                const bool ok = static_cast<int>(threadIdx.x) + u * 256 < blk_vals;
                rt[ok ? ri * RS + d : kMfmaRB * RS + l] = pre[u] - mus[d];
                ri += step_r;
                d += step_d;
                if (d >= dim) {
                    d -= dim;
                    ++ri;
                }
            }
        }
        if (threadIdx.x < kMfmaRB) rn[threadIdx.x] = pre_n;
        __syncthreads();
        if (r0 + kMfmaRB < n) load_block(r0 + kMfmaRB);
        for (int tg = 0; tg < kMfmaRB; tg += 64) {
            floatx4 acc[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) acc[t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int s = 0; s < S; ++s) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const float a = rt[(tg + t * 16 + jj) * RS + g * S + s];
                    acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, bq[s], acc[t], 0, 0, 0);
                }
            }
            // lane l holds row 4 (l / 16) + v of each tile: point tg + 16 t + 4 g + v.
            // The 16 values are screened first (32-bit block-local indices);
            // the insertion runs only when some lane has a candidate.
            // (points past the end have an infinite norm, so they never
            // pass; the query itself may enter its own lists and is skipped
            // by the ranking)
            float aps[16];
            float amin = INFINITY;
#pragma unroll
            for (int t = 0; t < 4; ++t) {
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int rl = tg + t * 16 + 4 * g + v;
                    const float ap = fmaf(-2.0f, acc[t][v], qn + rn[rl]);
                    aps[t * 4 + v] = ap;
                    amin = fminf(amin, ap);
                }
            }
            if (amin <= win) {
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    // a tile's 4 values with nothing at or below the window
                    // are passed over with one test (the window only shrinks)
                    const float tm = fminf(fminf(aps[4 * t], aps[4 * t + 1]), fminf(aps[4 * t + 2], aps[4 * t + 3]));
                    if (!(tm <= win)) continue;
#pragma unroll
                    for (int v = 0; v < 4; ++v) {
                        const int b = 4 * t + v;
                        const float ap = aps[b];
                        if (ap <= win && ap < INFINITY) {
                            if (ap <= lv[kMfmaList - 1]) {
                                float cv = ap;
                                int32_t ci = static_cast<int32_t>(r0 + tg + t * 16 + 4 * g + v);
#pragma unroll
                                for (int m = 0; m < kMfmaList; ++m) {
                                    const bool lt = cv < lv[m];
                                    const float tv = lv[m];
                                    const int32_t ti = li[m];
                                    lv[m] = lt ? cv : tv;
                                    li[m] = lt ? ci : ti;
                                    cv = lt ? tv : cv;
                                    ci = lt ? ti : ci;
                                }
                                drop = fminf(drop, cv);  // the value pushed out (+inf while the list fills)
                            } else {
                                drop = fminf(drop, ap);
                            }
                            win = lv[kMfmaK] + margin2;
                        }
                    }
                }
            }
        }
    }
    if (!qlive) return;
    const int64_t base = (q * 4 + g) * kMfmaList;
#pragma unroll
    for (int m = 0; m < kMfmaList; ++m) {
        cand_v[base + m] = lv[m];
        cand_i[base + m] = li[m];
    }
    dropped[q * 4 + g] = drop;
}

// Exact distance between points q and r: the sequential double sum, a
// separate multiply and add per term (no contraction).
__device__ inline double exact_sqdist(const float* __restrict__ x, int dim, int64_t q, int64_t r) {
    double s = 0.0;
    for (int d = 0; d < dim; ++d) {
        const double t = static_cast<double>(x[q * dim + d]) - static_cast<double>(x[r * dim + d]);
        const double sq = t * t;
        s = s + sq;
    }
    return s;
}

// Inserts (s, j) into the sorted list (bd, bi) of K entries if it precedes
// the last, in lexicographic (distance, index) order.
template <int K>
__device__ inline void insert_pair(double (&bd)[K], int32_t (&bi)[K], double s, int32_t j) {
    if (!(s < bd[K - 1] || (s == bd[K - 1] && j < bi[K - 1]))) return;
    double cd = s;
    int32_t ci = j;
#pragma unroll
    for (int m = 0; m < K; ++m) {
        const bool lt = cd < bd[m] || (cd == bd[m] && ci < bi[m]);
        const double td = bd[m];
        const int32_t ti = bi[m];
        bd[m] = lt ? cd : td;
        bi[m] = lt ? ci : ti;
        cd = lt ? td : cd;
        ci = lt ? ti : ci;
    }
}

// One thread per query: the window W from the k-th smallest approximate
// distance of the query's 4 candidate lists, the exact distances of the
// candidates at or below W, the k best by (distance, index); a query whose
// lanes dropped a value at or below W goes to the exact search instead.
__global__ void __launch_bounds__(kQ) k_knn_rerank(int64_t n, int dim, int s4, const float* __restrict__ x, int k,
                                                   const float* __restrict__ nrm,
                                                   const unsigned int* __restrict__ nr_max_bits,
                                                   const float* __restrict__ cand_v, const int32_t* __restrict__ cand_i,
                                                   const float* __restrict__ dropped, int32_t* __restrict__ out_idx,
                                                   float* __restrict__ out_dist, int32_t* __restrict__ flagged,
                                                   int* __restrict__ n_flagged) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * kQ + threadIdx.x;
    if (q >= n) return;
    const float* cv = cand_v + q * 4 * kMfmaList;
    const int32_t* ci = cand_i + q * 4 * kMfmaList;
    // k-th smallest of the union: merge the four sorted lists' heads
    int p[4] = {0, 0, 0, 0};
    float tau = INFINITY;
    for (int c = 0; c < k;) {
        int best = -1;
        float bv = INFINITY;
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const float v = p[g] < kMfmaList ? cv[g * kMfmaList + p[g]] : INFINITY;
            if (best < 0 || v < bv) {
                best = g;
                bv = v;
            }
        }
        if (p[best] < kMfmaList && ci[best * kMfmaList + p[best]] == q) {  // the query itself
            ++p[best];
            continue;
        }
        tau = bv;
        ++p[best];
        ++c;
        if (!(bv < INFINITY)) break;
    }
    const float W = tau + filter_margin2(s4, nrm[q], __uint_as_float(nr_max_bits[0]));
    float dmin = dropped[q * 4];
#pragma unroll
    for (int g = 1; g < 4; ++g) dmin = fminf(dmin, dropped[q * 4 + g]);
    if (!(W < INFINITY) || dmin <= W) {
        flagged[atomicAdd(n_flagged, 1)] = static_cast<int32_t>(q);
        return;
    }
    double bd[kMfmaK];
    int32_t bi[kMfmaK];
#pragma unroll
    for (int m = 0; m < kMfmaK; ++m) {
        bd[m] = INFINITY;
        bi[m] = 0x7fffffff;
    }
    for (int g = 0; g < 4; ++g) {
        for (int m = 0; m < kMfmaList; ++m) {
            const float v = cv[g * kMfmaList + m];
            if (!(v <= W)) break;
            const int32_t r = ci[g * kMfmaList + m];
            if (r != q) insert_pair<kMfmaK>(bd, bi, exact_sqdist(x, dim, q, r), r);
        }
    }
    for (int c = 0; c < k; ++c) {
        out_idx[q * k + c] = bi[c];
        out_dist[q * k + c] = static_cast<float>(sqrt(bd[c]));
    }
}

// Exact search of the flagged queries: one workgroup per query (grid-stride
// over the list), each thread the k best of a stride of the points, then a
// tree of pairwise merges in LDS.
__global__ void __launch_bounds__(kQ) k_knn_one(int64_t n, int dim, const float* __restrict__ x, int k,
                                                const int32_t* __restrict__ flagged,
                                                const int* __restrict__ n_flagged, int32_t* __restrict__ out_idx,
                                                float* __restrict__ out_dist) {
    __shared__ double sd[kQ][kMfmaK];
    __shared__ int32_t si[kQ][kMfmaK];
    const int count = n_flagged[0];
    const int tid = threadIdx.x;
    for (int f = blockIdx.x; f < count; f += gridDim.x) {
        const int64_t q = flagged[f];
        double bd[kMfmaK];
        int32_t bi[kMfmaK];
#pragma unroll
        for (int m = 0; m < kMfmaK; ++m) {
            bd[m] = INFINITY;
            bi[m] = 0x7fffffff;
        }
        for (int64_t r = tid; r < n; r += kQ)
            if (r != q) insert_pair<kMfmaK>(bd, bi, exact_sqdist(x, dim, q, r), static_cast<int32_t>(r));
#pragma unroll
        for (int m = 0; m < kMfmaK; ++m) {
            sd[tid][m] = bd[m];
            si[tid][m] = bi[m];
        }
        __syncthreads();
        for (int h = kQ / 2; h >= 1; h >>= 1) {
            if (tid < h) {
                // merge lists tid and tid + h (both sorted) into the first K
                int pa = 0, pb = 0;
                double md[kMfmaK];
                int32_t mi[kMfmaK];
                for (int m = 0; m < kMfmaK; ++m) {
                    const double da = sd[tid][pa], db = sd[tid + h][pb];
                    const int32_t ia = si[tid][pa], ib = si[tid + h][pb];
                    const bool ta = da < db || (da == db && ia < ib);
                    md[m] = ta ? da : db;
                    mi[m] = ta ? ia : ib;
                    pa += ta ? 1 : 0;
                    pb += ta ? 0 : 1;
                }
                for (int m = 0; m < kMfmaK; ++m) {
                    sd[tid][m] = md[m];
                    si[tid][m] = mi[m];
                }
            }
            __syncthreads();
        }
        if (tid < k) {
            out_idx[q * k + tid] = si[0][tid];
            out_dist[q * k + tid] = static_cast<float>(sqrt(sd[0][tid]));
        }
        __syncthreads();
    }
}

template <int S>
void launch_knn_mfma_s(int64_t n, int dim, const float* x, const float* mu, const float* nrm,
                       const unsigned int* nr_max_bits, float* cv, int32_t* ci, float* dropped, hipStream_t s) {
    const unsigned grid = static_cast<unsigned>((n + kMfmaQ - 1) / kMfmaQ);
    k_knn_mfma<S><<<grid, 256, 0, s>>>(n, dim, x, mu, nrm, nr_max_bits, cv, ci, dropped);
}

// The whole float path: mean, norms, candidates, exact ranking, exact search
// of the flagged queries.
void knn_mfma(int64_t n, int dim, const float* x, int k, int32_t* idx, float* dist, hipStream_t s) {
    const int S = (dim + 3) / 4;
    const int64_t blocks = (n + kQ - 1) / kQ;
    DevBuf<double> partial;
    DevBuf<float> mu, nrm, cv, dropped;
    DevBuf<int32_t> ci, flagged;
    DevBuf<int> counters;  // [0] largest norm (float bits), [1] flagged count
    partial.alloc(static_cast<size_t>(blocks * dim));
    mu.alloc(static_cast<size_t>(dim));
    nrm.alloc(static_cast<size_t>(n));
    cv.alloc(static_cast<size_t>(n) * 4 * kMfmaList);
    ci.alloc(static_cast<size_t>(n) * 4 * kMfmaList);
    dropped.alloc(static_cast<size_t>(n) * 4);
    flagged.alloc(static_cast<size_t>(n));
    counters.alloc(2);
    UMX_HIP(hipMemsetAsync(counters.p, 0, 2 * sizeof(int), s));
    unsigned int* nr_max_bits = reinterpret_cast<unsigned int*>(counters.p);
    int* n_flagged = counters.p + 1;
    k_colsum_partial<<<static_cast<unsigned>(blocks), kQ, 0, s>>>(n, dim, x, partial.p);
    k_colmean<<<1, 64, 0, s>>>(blocks, n, dim, partial.p, mu.p);
    k_centred_norms<<<static_cast<unsigned>(blocks), kQ, 0, s>>>(n, dim, x, mu.p, nrm.p, nr_max_bits);
    UMX_HIP(hipGetLastError());
    switch (S) {
#define UMX_KNN_S(v) \
    case v: launch_knn_mfma_s<v>(n, dim, x, mu.p, nrm.p, nr_max_bits, cv.p, ci.p, dropped.p, s); break;
        UMX_KNN_S(1) UMX_KNN_S(2) UMX_KNN_S(3) UMX_KNN_S(4) UMX_KNN_S(5) UMX_KNN_S(6) UMX_KNN_S(7) UMX_KNN_S(8)
        UMX_KNN_S(9) UMX_KNN_S(10) UMX_KNN_S(11) UMX_KNN_S(12) UMX_KNN_S(13) UMX_KNN_S(14) UMX_KNN_S(15)
        UMX_KNN_S(16)
#undef UMX_KNN_S
        default: throw Error(-1, "the matrix-core kNN path takes at most 64 coordinates");
    }
    UMX_HIP(hipGetLastError());
    k_knn_rerank<<<static_cast<unsigned>(blocks), kQ, 0, s>>>(n, dim, 4 * S, x, k, nrm.p, nr_max_bits, cv.p, ci.p,
                                                              dropped.p, idx, dist, flagged.p, n_flagged);
    UMX_HIP(hipGetLastError());
    k_knn_one<<<1024, kQ, 0, s>>>(n, dim, x, k, flagged.p, n_flagged, idx, dist);
    UMX_HIP(hipGetLastError());
    // the buffers are released (stream-ordered) when they go out of scope
    UMX_HIP(hipStreamSynchronize(s));
}

template <typename T>
void launch_knn(int64_t n, int dim, const T* x, int k, int32_t* idx, T* dist, hipStream_t s) {
    if constexpr (std::is_same<T, float>::value) {
        if (dim <= kLdsDims && k <= kMfmaK && n >= 4 * kMfmaRB) {
            knn_mfma(n, dim, x, k, idx, dist, s);
            return;
        }
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
