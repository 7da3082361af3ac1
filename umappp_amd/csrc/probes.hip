// Floor probes for the layout epoch's roofline (libumx_probes.so; used by
// bench.py, not by the library).  They time what the epoch kernel cannot do
// without, on arrays of the epoch's own sizes:
//
//  - random point gathers: every lane loads 'per_lane' uniformly random
//    float2 points of an n x 2 array in batches of six independent loads
//    (a firing edge's tail and five negative samples are one such batch in
//    the epoch kernel), and keeps their sum;
//  - the schedule stream: every slot of the sliced-ELL copy read once
//    (target, epochs per sample, next and next-negative epochs: 16 bytes,
//    non-temporal, coalesced), and the next and next-negative epochs of
//    'fires' slots written back.
//
// Launch bounds and indexes: gather indexes are high products of a 32-bit
// hash with n, so always below n; the stream kernel covers slots [0, slots)
// one thread each and writes only indexes below 'fires' <= slots.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kBlock = 256;
constexpr int kBatch = 6;

__device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
__device__ inline uint32_t hash32(uint32_t x) {
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

__global__ void __launch_bounds__(kBlock) k_gather_floor(const float2* __restrict__ pts, uint32_t n, int64_t lanes,
                                                       int batches, uint64_t seed, float* __restrict__ out) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (g >= lanes) return;
    const uint64_t key = mix64(seed ^ (static_cast<uint64_t>(g) * 0xd1b54a32d192ed03ULL));
    float acc = 0.0f;
    for (int t = 0; t < batches; ++t) {
        float2 v[kBatch];
#pragma unroll
        for (int u = 0; u < kBatch; ++u) {
            const uint32_t r = hash32(static_cast<uint32_t>(key) + 0x9e3779b9U * static_cast<uint32_t>(t * kBatch + u + 1)) ^
                               static_cast<uint32_t>(key >> 32);
            v[u] = pts[__umulhi(r, n)];
        }
#pragma unroll
        for (int u = 0; u < kBatch; ++u) acc += v[u].x + v[u].y;
    }
    out[g] = acc;
}

__global__ void __launch_bounds__(kBlock) k_stream_floor(const int32_t* __restrict__ tg, const float* __restrict__ eps,
                                                       float* __restrict__ nx, float* __restrict__ nn, int64_t slots,
                                                       int64_t fires, float* __restrict__ sink) {
    const int64_t s = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (s >= slots) return;
    const int32_t t = __builtin_nontemporal_load(tg + s);
    const float e = __builtin_nontemporal_load(eps + s);
    const float a = __builtin_nontemporal_load(nx + s);
    const float b = __builtin_nontemporal_load(nn + s);
    if (s < fires) {
        __builtin_nontemporal_store(a + e, nx + s);
        __builtin_nontemporal_store(b + static_cast<float>(t), nn + s);
    } else if (a + b + e == -1.0f) {  // never for the probe's inputs; keeps the loads
        sink[s & 255] = static_cast<float>(t);
    }
}

}  // namespace

extern "C" {

// Random float2 gathers from pts (n points): 'lanes' threads, 'batches'
// batches of six gathers each; out[lanes] receives the per-lane sums.
int umxp_gather_floor(const float* pts, uint32_t n, int64_t lanes, int batches, uint64_t seed, float* out,
                      void* stream) {
    if (pts == nullptr || out == nullptr || n == 0 || lanes <= 0 || batches < 0) return -1;
    const unsigned grid = static_cast<unsigned>((lanes + kBlock - 1) / kBlock);
    k_gather_floor<<<grid, kBlock, 0, static_cast<hipStream_t>(stream)>>>(reinterpret_cast<const float2*>(pts), n,
                                                                           lanes, batches, seed, out);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

// The schedule stream: reads tg, eps, nx, nn over 'slots' slots and writes
// nx, nn at the first 'fires' slots (fires <= slots).
int umxp_stream_floor(const int32_t* tg, const float* eps, float* nx, float* nn, int64_t slots, int64_t fires,
                      float* sink, void* stream) {
    if (!tg || !eps || !nx || !nn || !sink || slots <= 0 || fires < 0 || fires > slots) return -1;
    const unsigned grid = static_cast<unsigned>((slots + kBlock - 1) / kBlock);
    k_stream_floor<<<grid, kBlock, 0, static_cast<hipStream_t>(stream)>>>(tg, eps, nx, nn, slots, fires, sink);
    return hipGetLastError() == hipSuccess ? 0 : -3;
}

}  // extern "C"
