// Graph-side kernels for gfx950:
//   K1  fuzzy similarities (rho / sigma / weights), FP64 per row
//   K2  fuzzy-union symmetrisation, integer rank sorts + prefix scans
//   K3  similarities -> epochs_per_sample schedule
//
// Rows are processed by "row groups": G consecutive lanes of a 64-lane
// wavefront own one row (G = smallest power of two >= the longest row,
// 4..64), so a k = 15 kNN graph packs four rows per wavefront.  Lane-parallel
// sums use xor-butterfly shuffles inside the group, which leave bit-identical
// totals in every lane, so control flow stays uniform per row.
//
// Built with -ffp-contract=off: the union weights and schedule values must
// round exactly as the reference's separate multiply and add do.
#include "umx_internal.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <limits>

namespace umx {

namespace {

constexpr int kBlock = 256;

// Stream-ordered temporary buffer.
template <typename T>
struct Tmp {
    T* p = nullptr;
    hipStream_t s;
    Tmp(size_t count, hipStream_t st) : s(st) {
        if (count == 0) count = 1;
        prepare_pool();
        UMX_HIP(hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), s));
    }
    ~Tmp() {
        if (p) (void)hipFreeAsync(p, s);
    }
    Tmp(const Tmp&) = delete;
    Tmp& operator=(const Tmp&) = delete;
};

inline unsigned grid_for(int64_t threads) {
    const int64_t b = (threads + kBlock - 1) / kBlock;
    if (b > 0x7fffffffLL) throw Error(-5, "grid too large");
    return static_cast<unsigned>(b < 1 ? 1 : b);
}

template <int G>
__device__ inline double group_sum(double v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
    return v;
}
template <int G>
__device__ inline long long group_min(long long v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) {
        const long long o = __shfl_xor(v, m, G);
        v = o < v ? o : v;
    }
    return v;
}
template <int G>
__device__ inline long long group_max(long long v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) {
        const long long o = __shfl_xor(v, m, G);
        v = o > v ? o : v;
    }
    return v;
}
template <int G>
__device__ inline long long group_add(long long v) {
#pragma unroll
    for (int m = G / 2; m >= 1; m >>= 1) v += __shfl_xor(v, m, G);
    return v;
}
// Value held by entry q of the row (entry q lives in lane q % G, slot q / G).
template <int G, int M>
__device__ inline double group_fetch(const double (&d)[M], long long q) {
    const int src = static_cast<int>(q & (G - 1));
    const int slot = static_cast<int>(q / G);
    double out = 0;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const double x = __shfl(d[m], src, G);
        if (m == slot) out = x;
    }
    return out;
}

// ---------------------------------------------------------------------------
// K1: neighbor_similarities (neighbor_similarities.hpp:53-207), double path.
// ---------------------------------------------------------------------------
struct SimArgs {
    int64_t n;
    const int64_t* rp;
    long long raw;  // truncated local_connectivity
    double interp;  // its fractional part
    double bw;
    double mkds;
    int newton;
    double* rho;
    double* sigma;
};

template <typename T, int G, int M>
__global__ void __launch_bounds__(kBlock) k_similarities(SimArgs a, T* __restrict__ val) {
    const int64_t row = (static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x) / G;
    const int lane = threadIdx.x & (G - 1);
    if (row >= a.n) return;  // uniform per group
    const int64_t start = a.rp[row];
    const long long k = a.rp[row + 1] - start;
    const double nan = __builtin_nan("");
    if (k == 0) {
        if (lane == 0) {
            if (a.rho) a.rho[row] = nan;
            if (a.sigma) a.sigma[row] = nan;
        }
        return;
    }

    double d[M];
    bool ok[M];
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const long long q = lane + static_cast<long long>(G) * m;
        ok[m] = q < k;
        d[m] = ok[m] ? static_cast<double>(val[start + q]) : 0.0;
    }

    // Leading exact zeros (:72-78): position of the first non-zero entry.
    long long first_nz = k;
#pragma unroll
    for (int m = M - 1; m >= 0; --m) {
        const long long q = lane + static_cast<long long>(G) * m;
        if (ok[m] && d[m] != 0.0) first_nz = q;
    }
    const long long nzero = group_min<G>(first_nz);

    auto all_ones = [&](double rho_v) {
#pragma unroll
        for (int m = 0; m < M; ++m)
            if (ok[m]) val[start + lane + static_cast<long long>(G) * m] = static_cast<T>(1);
        if (lane == 0) {
            if (a.rho) a.rho[row] = rho_v;
            if (a.sigma) a.sigma[row] = nan;
        }
    };

    if (k - nzero <= a.raw) {  // :80-88
        all_ones(nan);
        return;
    }
    const long long ci = nzero + a.raw;
    const double lower = ci > 0 ? group_fetch<G, M>(d, ci - 1) : 0.0;
    const double upper = group_fetch<G, M>(d, ci);
    const double rho = lower + a.interp * (upper - lower);  // :92

    // Active deltas (:95-104).
    bool act[M];
    double delta[M];
    long long nle = 0, last_active = -1;
#pragma unroll
    for (int m = 0; m < M; ++m) {
        const long long q = lane + static_cast<long long>(G) * m;
        const bool in = ok[m] && q >= nzero;
        act[m] = in && d[m] > rho;
        delta[m] = act[m] ? d[m] - rho : 0.0;
        if (in && !act[m]) ++nle;
        if (act[m]) last_active = q;
    }
    nle = group_add<G>(nle);
    last_active = group_max<G>(last_active);
    if (last_active < 0) {  // :106-112
        all_ones(rho);
        return;
    }

    const double below = static_cast<double>(nzero + nle);
    double sigma = group_fetch<G, M>(d, last_active) - rho;  // active_delta.back() (:117)
    double lo = 0.0, hi = DBL_MAX;
    const double target = log2(static_cast<double>(k + 1)) * a.bw;  // :127

    for (int it = 0; it < 64; ++it) {  // :129-184
        const double inv = 1.0 / sigma;
        const double inv2 = inv * inv;
        double so = 0.0, sd = 0.0;
#pragma unroll
        for (int m = 0; m < M; ++m) {
            if (act[m]) {
                const double e = exp(-delta[m] * inv);
                so += e;
                sd += delta[m] * e * inv2;
            }
        }
        so = group_sum<G>(so);
        sd = group_sum<G>(sd);
        const double diff = (below + so) - target;
        if (fabs(diff) < 1e-5) break;
        if (diff > 0) hi = sigma;
        else lo = sigma;
        bool stepped = false;
        if (a.newton && sd != 0.0) {
            const double cand = sigma - diff / sd;
            if (cand > lo && cand < hi) {
                sigma = cand;
                stepped = true;
            }
        }
        if (!stepped) {
            if (diff > 0) sigma += (lo - sigma) / 2;
            else if (hi == DBL_MAX) sigma *= 2;
            else sigma += (hi - sigma) / 2;
        }
    }

    double sum = 0.0;  // :187-192
#pragma unroll
    for (int m = 0; m < M; ++m) sum += d[m];
    sum = group_sum<G>(sum);
    const double mean = sum / static_cast<double>(k);
    sigma = fmax(a.mkds * mean, sigma);

    const double inv = 1.0 / sigma;  // :194-202
#pragma unroll
    for (int m = 0; m < M; ++m) {
        if (ok[m]) {
            const double w = d[m] > rho ? exp(-(d[m] - rho) * inv) : 1.0;
            val[start + lane + static_cast<long long>(G) * m] = static_cast<T>(w);
        }
    }
    if (lane == 0) {
        if (a.rho) a.rho[row] = rho;
        if (a.sigma) a.sigma[row] = sigma;
    }
}

// One thread per row, reference summation order; used for rows longer than
// 512 entries, where the register-resident group kernel does not fit.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_similarities_seq(SimArgs a, T* __restrict__ val) {
    const int64_t row = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (row >= a.n) return;
    const int64_t start = a.rp[row];
    const long long k = a.rp[row + 1] - start;
    T* x = val + start;
    const double nan = __builtin_nan("");
    if (a.rho) a.rho[row] = nan;
    if (a.sigma) a.sigma[row] = nan;
    if (k == 0) return;
    long long nzero = 0;
    while (nzero < k && static_cast<double>(x[nzero]) == 0.0) ++nzero;
    if (k - nzero <= a.raw) {
        for (long long q = 0; q < k; ++q) x[q] = static_cast<T>(1);
        return;
    }
    const long long ci = nzero + a.raw;
    const double lower = ci > 0 ? static_cast<double>(x[ci - 1]) : 0.0;
    const double upper = static_cast<double>(x[ci]);
    const double rho = lower + a.interp * (upper - lower);
    if (a.rho) a.rho[row] = rho;
    double below = static_cast<double>(nzero);
    long long last_active = -1;
    for (long long q = nzero; q < k; ++q) {
        if (static_cast<double>(x[q]) > rho) last_active = q;
        else below += 1.0;
    }
    if (last_active < 0) {
        for (long long q = 0; q < k; ++q) x[q] = static_cast<T>(1);
        return;
    }
    double sigma = static_cast<double>(x[last_active]) - rho;
    double lo = 0.0, hi = DBL_MAX;
    const double target = log2(static_cast<double>(k + 1)) * a.bw;
    for (int it = 0; it < 64; ++it) {
        double obs = below, slope = 0.0;
        const double inv = 1.0 / sigma, inv2 = inv * inv;
        for (long long q = nzero; q < k; ++q) {
            const double dq = static_cast<double>(x[q]);
            if (dq > rho) {
                const double dl = dq - rho;
                const double e = exp(-dl * inv);
                obs += e;
                slope += dl * e * inv2;
            }
        }
        const double diff = obs - target;
        if (fabs(diff) < 1e-5) break;
        if (diff > 0) hi = sigma;
        else lo = sigma;
        bool stepped = false;
        if (a.newton && slope != 0.0) {
            const double cand = sigma - diff / slope;
            if (cand > lo && cand < hi) {
                sigma = cand;
                stepped = true;
            }
        }
        if (!stepped) {
            if (diff > 0) sigma += (lo - sigma) / 2;
            else if (hi == DBL_MAX) sigma *= 2;
            else sigma += (hi - sigma) / 2;
        }
    }
    double mean = 0.0;
    for (long long q = 0; q < k; ++q) mean += static_cast<double>(x[q]);
    mean /= static_cast<double>(k);
    sigma = fmax(a.mkds * mean, sigma);
    if (a.sigma) a.sigma[row] = sigma;
    const double inv = 1.0 / sigma;
    for (long long q = 0; q < k; ++q) {
        const double dq = static_cast<double>(x[q]);
        x[q] = static_cast<T>(dq > rho ? exp(-(dq - rho) * inv) : 1.0);
    }
}

template <typename T, int G>
void sim_dispatch_m(const SimArgs& a, T* val, int M, hipStream_t s) {
    const unsigned grid = grid_for(a.n * G);
    switch (M) {
        case 1: k_similarities<T, G, 1><<<grid, kBlock, 0, s>>>(a, val); break;
        case 2: k_similarities<T, G, 2><<<grid, kBlock, 0, s>>>(a, val); break;
        case 4: k_similarities<T, G, 4><<<grid, kBlock, 0, s>>>(a, val); break;
        default: k_similarities<T, G, 8><<<grid, kBlock, 0, s>>>(a, val); break;
    }
}

// ---------------------------------------------------------------------------
// Scan and reductions over int64 row data.
// ---------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;

__global__ void __launch_bounds__(kBlock) k_scan_tiles(const int64_t* __restrict__ in, int64_t* __restrict__ out,
                                                     int64_t count, int64_t* __restrict__ tile_sums) {
    __shared__ int64_t warp_tot[kBlock / kWave];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(threadIdx.x) * kScanItems;
    int64_t v[kScanItems];
    int64_t run = 0;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t idx = base + q;
        v[q] = idx < count ? in[idx] : 0;
        run += v[q];
    }
    // inclusive scan of per-thread totals across the wavefront
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    int64_t incl = run;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const int64_t o = __shfl_up(incl, off, kWave);
        if (lane >= off) incl += o;
    }
    if (lane == kWave - 1) warp_tot[wid] = incl;
    __syncthreads();
    int64_t wave_off = 0;
    for (int w = 0; w < wid; ++w) wave_off += warp_tot[w];
    int64_t acc = wave_off + incl - run;
#pragma unroll
    for (int q = 0; q < kScanItems; ++q) {
        const int64_t idx = base + q;
        if (idx < count) out[idx] = acc;
        acc += v[q];
    }
    if (threadIdx.x == kBlock - 1) {
        int64_t tot = 0;
        for (int w = 0; w < kBlock / kWave; ++w) tot += warp_tot[w];
        tile_sums[blockIdx.x] = tot;
    }
}

__global__ void __launch_bounds__(kBlock) k_add_tile_offsets(int64_t* __restrict__ out, int64_t count,
                                                           const int64_t* __restrict__ tile_offs) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (idx < count) out[idx] += tile_offs[idx / kScanTile];
}

__global__ void k_scan_total(const int64_t* __restrict__ in, int64_t* __restrict__ out, int64_t count) {
    out[count] = count > 0 ? out[count - 1] + in[count - 1] : 0;
}

// Largest value of the workgroup (an integer type), for one atomic per workgroup:
// single-address atomics serialise at one memory channel, so reductions
// below keep their grids to at most kReduceBlocks workgroups.
constexpr int64_t kReduceBlocks = 1024;

template <typename U>
__device__ inline U block_max(U v) {
    __shared__ U part[kBlock / kWave];
#pragma unroll
    for (int m = kWave / 2; m >= 1; m >>= 1) {
        const U o = __shfl_xor(v, m, kWave);
        v = o > v ? o : v;
    }
    if ((threadIdx.x & (kWave - 1)) == 0) part[threadIdx.x / kWave] = v;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kBlock / kWave; ++w) v = part[w] > v ? part[w] : v;
    return v;  // meaningful in thread 0
}

__global__ void __launch_bounds__(kBlock) k_max_row(const int64_t* __restrict__ rp, int64_t n,
                                                  unsigned long long* __restrict__ out) {
    unsigned long long len = 0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kBlock) {
        const unsigned long long l = static_cast<unsigned long long>(rp[i + 1] - rp[i]);
        len = l > len ? l : len;
    }
    len = block_max(len);
    if (threadIdx.x == 0 && len > 0) atomicMax(out, len);
}

// ---------------------------------------------------------------------------
// K2: combine_neighbor_sets (combine_neighbor_sets.hpp:15-100).
//
// Result per unordered pair {i, t}: if both directed entries exist, both
// output entries hold f(a, b) (union (a+b)-ab, product ab, or the mix of the
// two); if only i->t exists, mix = 1 keeps a on i->t and adds t->i = a,
// mix = 0 drops it, other mixes keep and mirror a*mix.  mix = 0 also removes
// every exact zero.  Rows come out sorted by index.
// ---------------------------------------------------------------------------
enum MixMode { MIX_INTERSECT = 0, MIX_UNION = 1, MIX_OTHER = 2 };

// Rank sort of every row by index, with validation of the indices.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sort_rows(int64_t n, int G, const int64_t* __restrict__ rp,
                                                    const int32_t* __restrict__ idx, const T* __restrict__ val,
                                                    int32_t* __restrict__ oidx, T* __restrict__ oval,
                                                    int* __restrict__ err) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (row >= n) return;
    const int64_t s = rp[row], e = rp[row + 1];
    for (int64_t q = s + lane; q < e; q += G) {
        const int32_t key = idx[q];
        if (key < 0 || key >= n || key == row) atomicOr(err, 1);
        int64_t rank = 0;
        for (int64_t f = s; f < e; ++f) {
            const int32_t o = idx[f];
            rank += o < key;
            if (o == key && f != q) atomicOr(err, 2);
        }
        oidx[s + rank] = key;
        oval[s + rank] = val[q];
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_reciprocal(int64_t n, int G, const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ sidx, const T* __restrict__ sval,
                                                     T mix, int mode, T* __restrict__ nval,
                                                     unsigned char* __restrict__ flags,
                                                     int64_t* __restrict__ keep_cnt, int64_t* __restrict__ inc_cnt,
                                                     int32_t* __restrict__ mslot) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (row >= n) return;
    const int64_t s = rp[row], e = rp[row + 1];
    int64_t kept = 0;
    for (int64_t q = s + lane; q < e; q += G) {
        const int32_t t = sidx[q];
        const T a = sval[q];
        // binary search for 'row' in the sorted row t
        int64_t lo = rp[t], hi = rp[t + 1];
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (sidx[mid] < row) lo = mid + 1;
            else hi = mid;
        }
        const bool found = lo < rp[t + 1] && sidx[lo] == row;
        T v;
        unsigned char keep, mirror = 0;
        if (found) {
            const T b = sval[lo];
            const T prod = a * b;
            if (mode == MIX_UNION) v = a + b - prod;
            else if (mode == MIX_INTERSECT) v = prod;
            else v = mix * (a + b - prod) + (1 - mix) * prod;
            keep = mode == MIX_INTERSECT ? (v != 0) : 1;
        } else if (mode == MIX_UNION) {
            v = a;
            keep = 1;
            mirror = 1;
        } else if (mode == MIX_INTERSECT) {
            v = 0;
            keep = 0;
        } else {
            v = a * mix;
            keep = 1;
            mirror = 1;
        }
        nval[q] = v;
        flags[q] = static_cast<unsigned char>(keep | (mirror << 1));
        kept += keep;
        // the count's old value numbers the mirrored entry within row t
        if (mirror)
            mslot[q] = static_cast<int32_t>(atomicAdd(reinterpret_cast<unsigned long long*>(inc_cnt + t), 1ULL));
    }
    // the row's G lanes (aligned, all active) sum their counts; one store
    for (int o = G / 2; o >= 1; o >>= 1) kept += __shfl_xor(kept, o, kWave);
    if (lane == 0) keep_cnt[row] = kept;
}

__global__ void __launch_bounds__(kBlock) k_add_counts(const int64_t* __restrict__ a, const int64_t* __restrict__ b,
                                                     int64_t* __restrict__ out, int64_t n) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) out[i] = a[i] + b[i];
}

// Writes each mirrored entry i -> t as (i, value) into row t's staging
// range, at the place k_reciprocal numbered it (no atomics here).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_stage_mirrors(int64_t n, int G, const int64_t* __restrict__ rp,
                                                        const int32_t* __restrict__ sidx, const T* __restrict__ nval,
                                                        const unsigned char* __restrict__ flags,
                                                        const int64_t* __restrict__ inc_rp,
                                                        const int32_t* __restrict__ mslot,
                                                        int32_t* __restrict__ stage_src, T* __restrict__ stage_val) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (row >= n) return;
    const int64_t s = rp[row], e = rp[row + 1];
    for (int64_t q = s + lane; q < e; q += G) {
        if (flags[q] & 2) {
            const int32_t t = sidx[q];
            const int64_t at = inc_rp[t] + mslot[q];
            stage_src[at] = static_cast<int32_t>(row);
            stage_val[at] = nval[q];
        }
    }
}

// Sorts each row's staged mirrored entries by index (the numbering atomics of
// k_reciprocal leave them in arbitrary order; keys are unique within a row,
// one per source row).
// Rows with at most kSmallMirrors entries: a group of kSmallMirrors lanes per
// row ranks its keys with shuffles.  Longer rows are listed by
// k_list_long_rows for k_sort_mirrors_mid (up to 64 entries) or
// k_sort_mirrors_big (more; an LDS capacity of kBigSortCapSmall entries for
// rows up to that length, kBigSortCap above: the large variant's 64 KB of
// LDS fits two workgroups per CU, the small one's 4 KB eight).
constexpr int kSmallMirrors = 32;
constexpr int kBigSortCapSmall = 512;
constexpr int kBigSortCap = 8192;  // LDS bitonic sort capacity per row
constexpr int kListRowsPerThread = 4;
constexpr int kLongLists = 3;       // mid, big (small LDS), big (large LDS)

// Appends the rows with more than kSmallMirrors staged entries to list 0
// (at most 64 entries), 1 (at most kBigSortCapSmall) or 2 (more), in no
// particular order; counts[l] is list l's length.  A workgroup takes
// kBlock * kListRowsPerThread consecutive rows, numbers its long rows with
// LDS atomics and reserves their list slots with one global atomic per list
// (one global atomic per long row queued on the counters: about 1 ms for
// the hub rows of a 1M-point kNN graph).
__global__ void __launch_bounds__(kBlock) k_list_long_rows(int64_t n, const int64_t* __restrict__ inc_rp,
                                                         int64_t* __restrict__ list0, int64_t* __restrict__ list1,
                                                         int64_t* __restrict__ list2,
                                                         unsigned long long* __restrict__ counts) {
    __shared__ unsigned int cnt[kLongLists];
    __shared__ unsigned long long base[kLongLists];
    if (threadIdx.x < kLongLists) cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t r0 = static_cast<int64_t>(blockIdx.x) * kBlock * kListRowsPerThread + threadIdx.x;
    int cls[kListRowsPerThread];
    unsigned int slot[kListRowsPerThread];
#pragma unroll
    for (int j = 0; j < kListRowsPerThread; ++j) {
        const int64_t row = r0 + static_cast<int64_t>(j) * kBlock;
        const int64_t m = row < n ? inc_rp[row + 1] - inc_rp[row] : 0;
        cls[j] = m <= kSmallMirrors ? -1 : (m <= kWave ? 0 : (m <= kBigSortCapSmall ? 1 : 2));
        slot[j] = cls[j] >= 0 ? atomicAdd(&cnt[cls[j]], 1u) : 0u;
    }
    __syncthreads();
    if (threadIdx.x < kLongLists) {
        const unsigned int c = cnt[threadIdx.x];
        base[threadIdx.x] = c ? atomicAdd(counts + threadIdx.x, static_cast<unsigned long long>(c)) : 0ULL;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < kListRowsPerThread; ++j) {
        if (cls[j] >= 0) {
            const int64_t row = r0 + static_cast<int64_t>(j) * kBlock;
            int64_t* list = cls[j] == 0 ? list0 : (cls[j] == 1 ? list1 : list2);
            list[base[cls[j]] + slot[j]] = row;
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_sort_mirrors_small(int64_t n, const int64_t* __restrict__ inc_rp,
                                                             int32_t* __restrict__ src, T* __restrict__ val) {
    constexpr int G = kSmallMirrors;
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    // every lane of the wavefront runs the shuffles below; rows past the end
    // take part with no entries
    const int64_t ss = row < n ? inc_rp[row] : 0;
    const int64_t m = row < n ? inc_rp[row + 1] - ss : 0;
    const bool small = m <= G;
    const bool mine = small && lane < m;
    const int32_t key = mine ? src[ss + lane] : 0x7fffffff;
    const T v = mine ? val[ss + lane] : static_cast<T>(0);
    int rank = 0;
    const int base = (threadIdx.x & (kWave - 1)) & ~(G - 1);
#pragma unroll
    for (int o = 0; o < G; ++o) {
        const int32_t other = __shfl(key, base + o, kWave);
        rank += other < key;
    }
    if (mine) {
        src[ss + rank] = key;
        val[ss + rank] = v;
    }
}

// Rows with 33 to 64 entries: one wavefront per row, the same shuffle rank
// sort as k_sort_mirrors_small over 64 lanes (every load of the row
// completes before the first store, so the row is sorted in place).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sort_mirrors_mid(const int64_t* __restrict__ mid_rows,
                                                           unsigned long long count,
                                                           const int64_t* __restrict__ inc_rp,
                                                           int32_t* __restrict__ src, T* __restrict__ val) {
    const int64_t w = static_cast<int64_t>(blockIdx.x) * (kBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (w >= static_cast<int64_t>(count)) return;  // uniform per wavefront
    const int64_t row = mid_rows[w];
    const int64_t ss = inc_rp[row];
    const int64_t m = inc_rp[row + 1] - ss;
    const bool mine = lane < m;
    const int32_t key = mine ? src[ss + lane] : 0x7fffffff;
    const T v = mine ? val[ss + lane] : static_cast<T>(0);
    int rank = 0;
#pragma unroll
    for (int o = 0; o < kWave; ++o) rank += __shfl(key, o, kWave) < key;
    if (mine) {
        src[ss + rank] = key;
        val[ss + rank] = v;
    }
}

// One workgroup per long row (more than 64 entries): bitonic sort of
// (key, value) in LDS.  Rows beyond kBigSortCap entries are ranked in global
// memory instead (quadratic, for extreme hubs only).
template <typename T, int CAP>
__global__ void __launch_bounds__(kBlock) k_sort_mirrors_big(const int64_t* __restrict__ big_rows,
                                                           const int64_t* __restrict__ inc_rp,
                                                           int32_t* __restrict__ src, T* __restrict__ val,
                                                           int32_t* __restrict__ tmp_src, T* __restrict__ tmp_val) {
    __shared__ int32_t keys[CAP];
    __shared__ int32_t pos[CAP];
    const int64_t row = big_rows[blockIdx.x];
    const int64_t ss = inc_rp[row];
    const int64_t m = inc_rp[row + 1] - ss;
    if (m <= CAP) {
        int p2 = 1;
        while (p2 < m) p2 <<= 1;
        for (int q = threadIdx.x; q < p2; q += kBlock) {
            keys[q] = q < m ? src[ss + q] : 0x7fffffff;
            pos[q] = q;
        }
        __syncthreads();
        for (int k = 2; k <= p2; k <<= 1) {
            for (int j = k >> 1; j > 0; j >>= 1) {
                for (int q = threadIdx.x; q < p2; q += kBlock) {
                    const int o = q ^ j;
                    if (o > q) {
                        const bool up = (q & k) == 0;
                        const int32_t a = keys[q], b = keys[o];
                        if ((a > b) == up) {
                            keys[q] = b;
                            keys[o] = a;
                            const int32_t t = pos[q];
                            pos[q] = pos[o];
                            pos[o] = t;
                        }
                    }
                }
                __syncthreads();
            }
        }
        // permute values through the scratch copy, then write back
        for (int q = threadIdx.x; q < m; q += kBlock) tmp_val[ss + q] = val[ss + pos[q]];
        __syncthreads();
        for (int q = threadIdx.x; q < m; q += kBlock) {
            src[ss + q] = keys[q];
            val[ss + q] = tmp_val[ss + q];
        }
    } else {
        for (int64_t q = threadIdx.x; q < m; q += kBlock) {
            const int32_t key = src[ss + q];
            int64_t rank = 0;
            for (int64_t f = 0; f < m; ++f) rank += src[ss + f] < key;
            tmp_src[ss + rank] = key;
            tmp_val[ss + rank] = val[ss + q];
        }
        __syncthreads();
        for (int64_t q = threadIdx.x; q < m; q += kBlock) {
            src[ss + q] = tmp_src[ss + q];
            val[ss + q] = tmp_val[ss + q];
        }
    }
}

// Final rows: kept entries (already in index order) merged with the mirrored
// entries (sorted per row above) by rank, so the output order is exactly
// ascending index.  With any mix other than 0 every original entry is kept;
// with mix = 0 there are no mirrors.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_merge_rows(int64_t n, int G, int mode, const int64_t* __restrict__ rp,
                                                     const int32_t* __restrict__ sidx, const T* __restrict__ nval,
                                                     const unsigned char* __restrict__ flags,
                                                     const int64_t* __restrict__ inc_rp,
                                                     const int32_t* __restrict__ stage_src,
                                                     const T* __restrict__ stage_val, const int64_t* __restrict__ out_rp,
                                                     int32_t* __restrict__ oidx, T* __restrict__ oval) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (row >= n) return;
    const int64_t s = rp[row], e = rp[row + 1];
    const int64_t ss = inc_rp[row], se = inc_rp[row + 1];
    const int64_t ob = out_rp[row];
    for (int64_t q = s + lane; q < e; q += G) {
        if (!(flags[q] & 1)) continue;
        const int32_t key = sidx[q];
        int64_t rank;
        if (mode == MIX_INTERSECT) {
            rank = 0;
            for (int64_t f = s; f < q; ++f) rank += flags[f] & 1;
        } else {
            rank = q - s;
            int64_t lo = ss, hi = se;  // mirrors with a smaller index
            while (lo < hi) {
                const int64_t mid = (lo + hi) >> 1;
                if (stage_src[mid] < key) lo = mid + 1;
                else hi = mid;
            }
            rank += lo - ss;
        }
        oidx[ob + rank] = key;
        oval[ob + rank] = nval[q];
    }
    for (int64_t q = ss + lane; q < se; q += G) {
        const int32_t key = stage_src[q];
        int64_t lo = s, hi = e;  // original entries with a smaller index
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (sidx[mid] < key) lo = mid + 1;
            else hi = mid;
        }
        const int64_t rank = (lo - s) + (q - ss);
        oidx[ob + rank] = key;
        oval[ob + rank] = stage_val[q];
    }
}

// ---------------------------------------------------------------------------
// K3: similarities_to_epochs (optimize_layout.hpp:47-92).
// ---------------------------------------------------------------------------
template <typename T>
struct Bits;
template <>
struct Bits<float> {
    using S = int;
};
template <>
struct Bits<double> {
    using S = long long;
};

// Largest non-NaN weight of each instance, starting from 0 like the
// reference's std::max fold.  Non-negative IEEE values order like their bit
// patterns as signed integers; negative values map to negative integers and
// never beat the initial +0.  blockIdx.y is the instance; its entries are
// [rp[pt_off[q]], rp[pt_off[q + 1]]) (all entries when pt_off is null).
template <typename T>
__global__ void __launch_bounds__(kBlock) k_max_weight(const T* __restrict__ val, int64_t nnz,
                                                     const int64_t* __restrict__ rp,
                                                     const int64_t* __restrict__ pt_off,
                                                     typename Bits<T>::S* __restrict__ out) {
    using S = typename Bits<T>::S;
    const int q = blockIdx.y;
    const int64_t lo = pt_off ? rp[pt_off[q]] : 0;
    const int64_t hi = pt_off ? rp[pt_off[q + 1]] : nnz;
    S best = 0;
    for (int64_t e = lo + static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; e < hi;
         e += static_cast<int64_t>(gridDim.x) * kBlock) {
        const T v = val[e];
        if (v == v) {
            S bits;
            __builtin_memcpy(&bits, &v, sizeof(S));
            best = bits > best ? bits : best;
        }
    }
    best = block_max(best);
    if (threadIdx.x == 0 && best > 0) atomicMax(out + q, best);
}

template <typename T>
__device__ inline T bits_to(typename Bits<T>::S b) {
    T v;
    __builtin_memcpy(&v, &b, sizeof(T));
    return v;
}

// Largest weight and survivor limit of the instance that owns 'row'
// (optimize_layout.hpp:49-63, per instance).
template <typename T>
struct RowLimit {
    T maxed, limit;
};
template <typename T>
__device__ inline RowLimit<T> row_limit(int64_t row, const int32_t* __restrict__ of_row,
                                        const typename Bits<T>::S* __restrict__ maxbits, int num_epochs,
                                        const int32_t* __restrict__ epochs_of_inst) {
    const int q = of_row ? of_row[row] : 0;
    const T maxed = bits_to<T>(maxbits[q]);
    const int ne = epochs_of_inst ? epochs_of_inst[q] : num_epochs;
    return RowLimit<T>{maxed, maxed / static_cast<T>(ne)};  // :63
}

// Bits of the G-lane group of the calling lane in a wavefront ballot.
__device__ inline unsigned long long group_bits(unsigned long long ballot, int G) {
    const int base = (threadIdx.x & (kWave - 1)) & ~(G - 1);
    const unsigned long long mask = G >= kWave ? ~0ULL : ((1ULL << G) - 1ULL);
    return (ballot >> base) & mask;
}

// Survivors per row: G lanes per row walk it in chunks of G entries and count
// with ballots.  Every lane of a group runs the same trips (the trip count
// depends on the row only), which the ballots need.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_count_survivors(int64_t n, int G, const int64_t* __restrict__ rp,
                                                          const T* __restrict__ val,
                                                          const typename Bits<T>::S* __restrict__ maxbits,
                                                          int num_epochs, const int32_t* __restrict__ epochs_of_inst,
                                                          const int32_t* __restrict__ of_row,
                                                          int64_t* __restrict__ cnt) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (row >= n) return;
    const T limit = row_limit<T>(row, of_row, maxbits, num_epochs, epochs_of_inst).limit;
    const int64_t s = rp[row], e = rp[row + 1];
    int64_t c = 0;
    for (int64_t b0 = s; b0 < e; b0 += G) {
        const int64_t q = b0 + lane;
        const bool keep = q < e && val[q] >= limit;
        c += __popcll(group_bits(__ballot(keep), G));
    }
    if (lane == 0) cnt[row] = c;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fill_epochs(int64_t n, int G, const int64_t* __restrict__ rp,
                                                      const int32_t* __restrict__ idx, const T* __restrict__ val,
                                                      const typename Bits<T>::S* __restrict__ maxbits,
                                                      int num_epochs, const int32_t* __restrict__ epochs_of_inst,
                                                      const int32_t* __restrict__ of_row, T nsr,
                                                      const int64_t* __restrict__ orp, int32_t* __restrict__ tg,
                                                      T* __restrict__ eps, T* __restrict__ nxt,
                                                      T* __restrict__ nxt_neg) {
    const int64_t gid = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    const int64_t row = gid / G;
    const int lane = static_cast<int>(gid & (G - 1));
    if (row >= n) return;
    const RowLimit<T> rl = row_limit<T>(row, of_row, maxbits, num_epochs, epochs_of_inst);
    const int64_t s = rp[row], e = rp[row + 1];
    int64_t at = orp[row];
    for (int64_t b0 = s; b0 < e; b0 += G) {
        const int64_t q = b0 + lane;
        const T w = q < e ? val[q] : static_cast<T>(0);
        const bool keep = q < e && w >= rl.limit;
        const unsigned long long bits = group_bits(__ballot(keep), G);
        if (keep) {
            const int64_t o = at + __popcll(bits & ((1ULL << lane) - 1ULL));
            const T ep = rl.maxed / w;  // :70
            tg[o] = idx[q];
            eps[o] = ep;
            nxt[o] = ep;  // :77
            nxt_neg[o] = ep / nsr;  // :78-81
        }
        at += __popcll(bits);
    }
}

// ---------------------------------------------------------------------------
// Instances of a batched graph.
// ---------------------------------------------------------------------------
__device__ inline int instance_of(int64_t i, const int64_t* __restrict__ pt_off, int n_inst) {
    int lo = 0, hi = n_inst;  // largest q with pt_off[q] <= i
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (pt_off[mid] <= i) lo = mid;
        else hi = mid;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlock) k_instance_of_row(int64_t n, const int64_t* __restrict__ pt_off, int n_inst,
                                                          int32_t* __restrict__ of_row) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i < n) of_row[i] = instance_of(i, pt_off, n_inst);
}

__global__ void __launch_bounds__(kBlock) k_localize(int64_t n, const int64_t* __restrict__ rp,
                                                   int32_t* __restrict__ idx, const int64_t* __restrict__ pt_off,
                                                   const int32_t* __restrict__ of_row, int* __restrict__ bad) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (i >= n) return;
    const int q = of_row[i];
    const int64_t off = pt_off[q], size = pt_off[q + 1] - off;
    for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
        const int32_t j = idx[e];
        if (j < 0 || j >= size) {
            atomicOr(bad, 1);
            idx[e] = -1;  // rejected again by the symmetrisation's index check
        } else {
            idx[e] = static_cast<int32_t>(off + j);
        }
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------
int lanes_for_degree(int64_t max_degree) {
    int g = 4;
    while (g < 64 && g < max_degree) g <<= 1;
    return g;
}

void exclusive_scan_i64(const int64_t* in, int64_t* out, int64_t count, DevBuf<int64_t>& tmp, hipStream_t s) {
    (void)tmp;
    if (count > 0) {
        const int64_t tiles = (count + kScanTile - 1) / kScanTile;
        Tmp<int64_t> sums(static_cast<size_t>(tiles), s);
        k_scan_tiles<<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(in, out, count, sums.p);
        UMX_HIP(hipGetLastError());
        if (tiles > 1) {
            Tmp<int64_t> offs(static_cast<size_t>(tiles + 1), s);
            DevBuf<int64_t> unused;
            exclusive_scan_i64(sums.p, offs.p, tiles, unused, s);
            k_add_tile_offsets<<<grid_for(count), kBlock, 0, s>>>(out, count, offs.p);
            UMX_HIP(hipGetLastError());
        }
    }
    k_scan_total<<<1, 1, 0, s>>>(in, out, count);
    UMX_HIP(hipGetLastError());
}

int64_t max_row_length(const int64_t* rp, int64_t n, DevBuf<int64_t>& tmp, hipStream_t s) {
    (void)tmp;
    if (n == 0) return 0;
    Tmp<unsigned long long> m(1, s);
    UMX_HIP(hipMemsetAsync(m.p, 0, sizeof(unsigned long long), s));
    k_max_row<<<static_cast<unsigned>(std::min<int64_t>((n + kBlock - 1) / kBlock, kReduceBlocks)), kBlock, 0, s>>>(rp, n, m.p);
    UMX_HIP(hipGetLastError());
    unsigned long long h = 0;
    UMX_HIP(hipMemcpyAsync(&h, m.p, sizeof(h), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    return static_cast<int64_t>(h);
}

template <typename T>
void launch_similarities(int64_t n, const int64_t* rp, T* val, double lc, double bw, double mkds, int newton,
                         double* rho, double* sigma, int64_t max_degree, hipStream_t s) {
    if (n == 0) return;
    SimArgs a;
    a.n = n;
    a.rp = rp;
    a.raw = static_cast<long long>(lc);
    a.interp = lc - static_cast<double>(a.raw);
    a.bw = bw;
    a.mkds = mkds;
    a.newton = newton;
    a.rho = rho;
    a.sigma = sigma;
    if (max_degree > 512) {
        k_similarities_seq<T><<<grid_for(n), kBlock, 0, s>>>(a, val);
    } else {
        const int G = lanes_for_degree(max_degree);
        int M = 1;
        while (static_cast<int64_t>(G) * M < max_degree) M <<= 1;
        switch (G) {
            case 4: sim_dispatch_m<T, 4>(a, val, M, s); break;
            case 8: sim_dispatch_m<T, 8>(a, val, M, s); break;
            case 16: sim_dispatch_m<T, 16>(a, val, M, s); break;
            case 32: sim_dispatch_m<T, 32>(a, val, M, s); break;
            default: sim_dispatch_m<T, 64>(a, val, M, s); break;
        }
    }
    UMX_HIP(hipGetLastError());
}

template <typename T>
void run_combine(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, T mix, int64_t max_degree,
                 DevBuf<int64_t>& out_rp, DevBuf<int32_t>& out_idx, DevBuf<T>& out_val, hipStream_t s) {
    const int64_t nnz_in = [&] {
        int64_t v = 0;
        UMX_HIP(hipMemcpyAsync(&v, rp + n, sizeof(v), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        return v;
    }();
    const int mode = mix == static_cast<T>(1) ? MIX_UNION : (mix == static_cast<T>(0) ? MIX_INTERSECT : MIX_OTHER);
    const int G = lanes_for_degree(max_degree);
    const unsigned grid = grid_for(n * G);

    Tmp<int32_t> sidx(static_cast<size_t>(nnz_in), s);
    Tmp<T> sval(static_cast<size_t>(nnz_in), s);
    Tmp<int> err(1, s);
    UMX_HIP(hipMemsetAsync(err.p, 0, sizeof(int), s));
    k_sort_rows<T><<<grid, kBlock, 0, s>>>(n, G, rp, idx, val, sidx.p, sval.p, err.p);
    UMX_HIP(hipGetLastError());
    int herr = 0;
    UMX_HIP(hipMemcpyAsync(&herr, err.p, sizeof(int), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    if (herr & 1) throw Error(-2, "neighbour index out of range or equal to its own row");
    if (herr & 2) throw Error(-2, "duplicate neighbour index within a row");

    Tmp<T> nval(static_cast<size_t>(nnz_in), s);
    Tmp<unsigned char> flags(static_cast<size_t>(nnz_in), s);
    Tmp<int64_t> cnt(static_cast<size_t>(2 * n), s);  // keep counts, then incoming counts
    int64_t* keep_cnt = cnt.p;
    int64_t* inc_cnt = cnt.p + n;
    UMX_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int64_t) * static_cast<size_t>(2 * n), s));
    Tmp<int32_t> mslot(static_cast<size_t>(nnz_in), s);  // place of each mirrored entry in its target row
    k_reciprocal<T><<<grid, kBlock, 0, s>>>(n, G, rp, sidx.p, sval.p, mix, mode, nval.p, flags.p, keep_cnt, inc_cnt,
                                            mslot.p);
    UMX_HIP(hipGetLastError());

    DevBuf<int64_t> unused;
    Tmp<int64_t> inc_rp(static_cast<size_t>(n + 1), s);
    exclusive_scan_i64(inc_cnt, inc_rp.p, n, unused, s);
    Tmp<int64_t> deg(static_cast<size_t>(n), s);
    k_add_counts<<<grid_for(n), kBlock, 0, s>>>(keep_cnt, inc_cnt, deg.p, n);
    UMX_HIP(hipGetLastError());
    out_rp.alloc(static_cast<size_t>(n + 1));
    exclusive_scan_i64(deg.p, out_rp.p, n, unused, s);

    int64_t nnz_out = 0, n_inc = 0;
    UMX_HIP(hipMemcpyAsync(&nnz_out, out_rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipMemcpyAsync(&n_inc, inc_rp.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));

    Tmp<int32_t> stage_src(static_cast<size_t>(n_inc), s);
    Tmp<T> stage_val(static_cast<size_t>(n_inc), s);
    if (n_inc > 0) {
        k_stage_mirrors<T><<<grid, kBlock, 0, s>>>(n, G, rp, sidx.p, nval.p, flags.p, inc_rp.p, mslot.p,
                                                   stage_src.p, stage_val.p);
        UMX_HIP(hipGetLastError());
    }
    if (n_inc > 0) {
        Tmp<int64_t> mid_rows(static_cast<size_t>(n), s);
        Tmp<int64_t> big_rows(static_cast<size_t>(n), s);
        Tmp<int64_t> huge_rows(static_cast<size_t>(n), s);
        Tmp<unsigned long long> counts(kLongLists, s);  // mid, big, huge
        UMX_HIP(hipMemsetAsync(counts.p, 0, kLongLists * sizeof(unsigned long long), s));
        const int64_t list_rows = static_cast<int64_t>(kBlock) * kListRowsPerThread;
        k_list_long_rows<<<static_cast<unsigned>((n + list_rows - 1) / list_rows), kBlock, 0, s>>>(
            n, inc_rp.p, mid_rows.p, big_rows.p, huge_rows.p, counts.p);
        UMX_HIP(hipGetLastError());
        k_sort_mirrors_small<T><<<grid_for(n * kSmallMirrors), kBlock, 0, s>>>(n, inc_rp.p, stage_src.p,
                                                                               stage_val.p);
        UMX_HIP(hipGetLastError());
        unsigned long long hc[kLongLists] = {0, 0, 0};
        UMX_HIP(hipMemcpyAsync(hc, counts.p, sizeof(hc), hipMemcpyDeviceToHost, s));
        UMX_HIP(hipStreamSynchronize(s));
        const unsigned long long hmid = hc[0], hbig = hc[1], hhuge = hc[2];
        if (hmid > 0) {
            const unsigned mid_grid = static_cast<unsigned>((hmid + kBlock / kWave - 1) / (kBlock / kWave));
            k_sort_mirrors_mid<T><<<mid_grid, kBlock, 0, s>>>(mid_rows.p, hmid, inc_rp.p, stage_src.p, stage_val.p);
            UMX_HIP(hipGetLastError());
        }
        if (hbig + hhuge > 0) {
            Tmp<int32_t> tsrc(static_cast<size_t>(n_inc), s);
            Tmp<T> tval(static_cast<size_t>(n_inc), s);
            if (hbig > 0) {
                k_sort_mirrors_big<T, kBigSortCapSmall><<<static_cast<unsigned>(hbig), kBlock, 0, s>>>(
                    big_rows.p, inc_rp.p, stage_src.p, stage_val.p, tsrc.p, tval.p);
                UMX_HIP(hipGetLastError());
            }
            if (hhuge > 0) {
                k_sort_mirrors_big<T, kBigSortCap><<<static_cast<unsigned>(hhuge), kBlock, 0, s>>>(
                    huge_rows.p, inc_rp.p, stage_src.p, stage_val.p, tsrc.p, tval.p);
                UMX_HIP(hipGetLastError());
            }
        }
    }
    out_idx.alloc(static_cast<size_t>(nnz_out));
    out_val.alloc(static_cast<size_t>(nnz_out));
    k_merge_rows<T><<<grid, kBlock, 0, s>>>(n, G, mode, rp, sidx.p, nval.p, flags.p, inc_rp.p, stage_src.p,
                                            stage_val.p, out_rp.p, out_idx.p, out_val.p);
    UMX_HIP(hipGetLastError());
    UMX_HIP(hipStreamSynchronize(s));
}

void fill_instance_of_row(int64_t n, const Instances& inst, int32_t* of_row, hipStream_t s) {
    if (n <= 0) return;
    k_instance_of_row<<<grid_for(n), kBlock, 0, s>>>(n, inst.pt_off, inst.n_inst, of_row);
    UMX_HIP(hipGetLastError());
}

bool localize_to_global(int64_t n, const int64_t* rp, int32_t* idx, const Instances& inst, hipStream_t s) {
    if (n <= 0) return true;
    Tmp<int> bad(1, s);
    UMX_HIP(hipMemsetAsync(bad.p, 0, sizeof(int), s));
    k_localize<<<grid_for(n), kBlock, 0, s>>>(n, rp, idx, inst.pt_off, inst.of_row, bad.p);
    UMX_HIP(hipGetLastError());
    int h = 0;
    UMX_HIP(hipMemcpyAsync(&h, bad.p, sizeof(int), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    return h == 0;
}

template <typename T>
void run_to_epochs(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, int num_epochs,
                   const int32_t* epochs_of_inst, const Instances& inst, T nsr, int64_t max_degree,
                   DevBuf<int64_t>& out_rp, DevBuf<int32_t>& out_tg, DevBuf<T>& eps, DevBuf<T>& nxt,
                   DevBuf<T>& nxt_neg, hipStream_t s) {
    using S = typename Bits<T>::S;
    int64_t nnz = 0;
    UMX_HIP(hipMemcpyAsync(&nnz, rp + n, sizeof(nnz), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    (void)max_degree;
    // lanes per row from the average degree: long rows loop over chunks
    const int G = lanes_for_degree(n > 0 ? (2 * nnz + n - 1) / n : 1);
    const unsigned grid = grid_for(n * G);
    const int n_inst = inst.n_inst;
    const int32_t* of_row = n_inst > 1 ? inst.of_row : nullptr;

    Tmp<S> maxbits(static_cast<size_t>(n_inst), s);
    UMX_HIP(hipMemsetAsync(maxbits.p, 0, sizeof(S) * static_cast<size_t>(n_inst), s));
    if (nnz > 0) {
        const int64_t want = (nnz / n_inst + kBlock - 1) / kBlock + 1;
        const unsigned gx = static_cast<unsigned>(std::min<int64_t>(want, n_inst > 1 ? 512 : kReduceBlocks));
        k_max_weight<T><<<dim3(gx, static_cast<unsigned>(n_inst)), kBlock, 0, s>>>(
            val, nnz, rp, n_inst > 1 ? inst.pt_off : nullptr, maxbits.p);
        UMX_HIP(hipGetLastError());
    }
    Tmp<int64_t> cnt(static_cast<size_t>(n), s);
    UMX_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int64_t) * static_cast<size_t>(n), s));
    k_count_survivors<T><<<grid, kBlock, 0, s>>>(n, G, rp, val, maxbits.p, num_epochs, epochs_of_inst, of_row,
                                                 cnt.p);
    UMX_HIP(hipGetLastError());
    out_rp.alloc(static_cast<size_t>(n + 1));
    DevBuf<int64_t> unused;
    exclusive_scan_i64(cnt.p, out_rp.p, n, unused, s);
    int64_t kept = 0;
    UMX_HIP(hipMemcpyAsync(&kept, out_rp.p + n, sizeof(kept), hipMemcpyDeviceToHost, s));
    UMX_HIP(hipStreamSynchronize(s));
    out_tg.alloc(static_cast<size_t>(kept));
    eps.alloc(static_cast<size_t>(kept));
    nxt.alloc(static_cast<size_t>(kept));
    nxt_neg.alloc(static_cast<size_t>(kept));
    k_fill_epochs<T><<<grid, kBlock, 0, s>>>(n, G, rp, idx, val, maxbits.p, num_epochs, epochs_of_inst, of_row, nsr,
                                             out_rp.p, out_tg.p, eps.p, nxt.p, nxt_neg.p);
    UMX_HIP(hipGetLastError());
    UMX_HIP(hipStreamSynchronize(s));
}

#define UMX_INST(T)                                                                                              \
    template void launch_similarities<T>(int64_t, const int64_t*, T*, double, double, double, int, double*,     \
                                         double*, int64_t, hipStream_t);                                         \
    template void run_combine<T>(int64_t, const int64_t*, const int32_t*, const T*, T, int64_t, DevBuf<int64_t>&, \
                                 DevBuf<int32_t>&, DevBuf<T>&, hipStream_t);                                     \
    template void run_to_epochs<T>(int64_t, const int64_t*, const int32_t*, const T*, int, const int32_t*,       \
                                   const Instances&, T, int64_t, DevBuf<int64_t>&, DevBuf<int32_t>&, DevBuf<T>&, \
                                   DevBuf<T>&, DevBuf<T>&, hipStream_t);
UMX_INST(float)
UMX_INST(double)
#undef UMX_INST

}  // namespace umx
