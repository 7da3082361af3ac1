// Device helpers shared by the layout kernels (layout_kernels.hip,
// sell_kernels.hip): the counter sampler, the reference's clip and epsilon,
// point loads and stores, tallies and the update trace.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>
#include <cstdint>

#include "umx_internal.hpp"

namespace umx {
namespace dev {

constexpr int kBlock = 256;
constexpr uint64_t kGold = 0x9e3779b97f4a7c15ULL;
constexpr uint64_t kOdd = 0xd1b54a32d192ed03ULL;

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ inline uint64_t edge_key(uint64_t ekey, uint64_t j) { return mix64(ekey ^ (j * kOdd)); }
// 32-bit integer hash with low bias (C. Wellons, "lowbias32").
__host__ __device__ inline uint32_t lowbias32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
// Draw p of an edge with key ej, uniform in [0, n).  Up to 2^24 observations
// the draw hashes the low key word with lowbias32 and maps the 32-bit result
// by a high multiply (relative bias below n / 2^32); beyond, a full 64-bit
// mix64 draw is mapped by a 64-bit high multiply.
constexpr uint32_t kSmallDraw = 1u << 24;
__device__ inline int32_t draw_index_small(uint64_t ej, uint32_t p, uint32_t n) {  // n <= kSmallDraw
    const uint32_t r = lowbias32(static_cast<uint32_t>(ej) + 0x9e3779b9U * (p + 1)) ^ static_cast<uint32_t>(ej >> 32);
    return static_cast<int32_t>(__umulhi(r, n));
}
__device__ inline int32_t draw_index(uint64_t ej, uint32_t p, uint32_t n) {
    if (n <= kSmallDraw) return draw_index_small(ej, p, n);
    const uint64_t r = mix64(ej + kGold * (static_cast<uint64_t>(p) + 1));
    return static_cast<int32_t>(__umul64hi(r, static_cast<uint64_t>(n)));
}

template <typename T>
__device__ inline T clip4(T v) {  // std::min(std::max(v, -4), 4), optimize_layout.hpp:105-110
    const T lo = v < static_cast<T>(-4) ? static_cast<T>(-4) : v;
    return static_cast<T>(4) < lo ? static_cast<T>(4) : lo;
}
// float: the median of (v, -4, 4) in one instruction, equal to the clamp
// above for every non-NaN v.
template <>
__device__ inline float clip4<float>(float v) {
    return __builtin_amdgcn_fmed3f(v, -4.0f, 4.0f);
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

// Quotient for the production kernels: float uses the hardware reciprocal
// (within about 1 ulp), double the IEEE division.
__device__ inline float prod_div(float num, float den) { return num * __builtin_amdgcn_rcpf(den); }
__device__ inline double prod_div(double num, double den) { return num / den; }

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

// Wave-aggregated tally update: one atomic per wavefront per counter, into
// partial-tally slot 'slot' % kTallySlots.
__device__ inline void tally_add(unsigned long long* part, int64_t slot, unsigned long long fires,
                                 unsigned long long draws, unsigned long long self_hits) {
    unsigned long long* t = part + (slot % kTallySlots) * kTallyStride;
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
// The record's first value is kind + 4 * the launch's phase.
template <typename T>
__device__ inline double* trace_slot(const LayoutView<T>& v, int kind, int64_t head, int64_t other) {
    if (!v.trace) return nullptr;
    const unsigned long long slot = atomicAdd(v.trace_len, 1ULL);
    if (slot >= static_cast<unsigned long long>(v.trace_cap)) return nullptr;
    double* rec = v.trace + slot * (3 + 3 * static_cast<long long>(v.dim));
    rec[0] = kind + 4 * v.phase;
    rec[1] = static_cast<double>(head);
    rec[2] = static_cast<double>(other);
    return rec;
}

struct EpochConst {
    int epoch;
    uint64_t ekey;
};

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


// Streaming loads and stores of the per-edge schedule arrays: marked
// non-temporal so that they do not evict the embedding, which every lane
// gathers from at random, from the L2.
template <typename T>
__device__ inline T ld_stream(const T* p) {
    return __builtin_nontemporal_load(p);
}
template <typename T>
__device__ inline void st_stream(T* p, T v) {
    __builtin_nontemporal_store(v, p);
}


}  // namespace dev
}  // namespace umx
