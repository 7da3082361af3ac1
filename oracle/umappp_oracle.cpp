// ORACLE — TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of libscran/umappp's graph-and-layout hot path, used only
// as the checker by tests/, by __graft_entry__.smoke() and by the cpu_baseline
// leg of bench.py.  Nothing in the product (umappp_amd/, include/) links,
// loads or calls this file.
//
// Each function cites the reference file:line whose steps it restates.  The
// reference itself cannot be compiled in this image (its headers include
// sanisizer/, aarand/ and knncolle/, none of which exist here), so this
// restatement is pinned instead by the reference's own invariant tests,
// brute-force oracles and known answers, and by closed forms (tests/).
//
// Build: oracle/build.sh (g++ -O2 -std=c++17 -ffp-contract=off, no -march,
// so no FMA contraction can change a rounding).
//
// Parity status per function (details in the tests):
//   orc_neighbor_similarities   pinned (reference invariants + closed-form root check;
//                               rows whose exact sigma follows from the reference's
//                               steps; bit-equal to an independent Python restatement
//                               of :89-202 on rows taking every branch of the solve)
//   orc_combine_neighbor_sets   pinned (brute-force symmetrisation, reference test :47-112)
//   orc_similarities_to_epochs  pinned (closed form eps = max/w, reference Epochs test)
//   orc_layout_* (serial)       pinned (per-edge update vs finite differences of the
//                               UMAP losses; schedule vs closed-form firing counts and
//                               bit-equal to a numpy float32 replay of :131-179 at
//                               non-dyadic negative-sample rates; restart equality)
//   orc_layout_* (parallel)     pinned (bit-equal to serial, reference test :106-125)
//   discrete_uniform (rng kind 0) parity unpinned: aarand::discrete_uniform is not in
//                               the image; this is a documented rejection sampler.
//   orc_loss / orc_loss_grad /  the project's UMAP cross-entropy loss (no reference
//   orc_loss_fd                 counterpart): pinned by central finite differences of
//                               the loss against the closed-form gradient, by the
//                               reference's per-edge update directions
//                               (optimize_layout.hpp:145-176) and by hand-computed
//                               single-term values (tests/test_oracle_loss.py).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <random>
#include <thread>
#include <utility>
#include <vector>

namespace {

template <typename T>
using Row = std::vector<std::pair<int32_t, T>>;
template <typename T>
using Rows = std::vector<Row<T>>;

template <typename T>
Rows<T> csr_to_rows(int64_t n, const int64_t* rp, const int32_t* idx, const T* val) {
    Rows<T> out(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        auto& r = out[static_cast<size_t>(i)];
        r.reserve(static_cast<size_t>(rp[i + 1] - rp[i]));
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) r.emplace_back(idx ? idx[e] : 0, val[e]);
    }
    return out;
}

// ---------------------------------------------------------------------------
// neighbor_similarities: include/umappp/neighbor_similarities.hpp:53-207
// Works on one row at a time; 'd' holds the row's distances in neighbour
// order and is overwritten with fuzzy weights.
// ---------------------------------------------------------------------------
template <typename T>
void similarities_row(T* d, int64_t k, int64_t raw_connect, T interp, T bandwidth, T min_k_dist_scale,
                      bool newton, std::vector<T>& active, double* rho_out, double* sigma_out) {
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (rho_out) *rho_out = nan;
    if (sigma_out) *sigma_out = nan;
    if (k == 0) return;  // :66-68

    // Leading exact-zero distances (:72-78).
    int64_t nzero = 0;
    while (nzero < k && !d[nzero]) ++nzero;

    if (k - nzero <= raw_connect) {  // :80-88
        for (int64_t q = 0; q < k; ++q) d[q] = 1;
        return;
    }
    const int64_t ci = nzero + raw_connect;  // :89
    const T lower = ci > 0 ? d[ci - 1] : static_cast<T>(0);
    const T upper = d[ci];
    const T rho = lower + interp * (upper - lower);  // :92
    if (rho_out) *rho_out = static_cast<double>(rho);

    active.clear();  // :95-104
    T below = static_cast<T>(nzero);
    for (int64_t q = nzero; q < k; ++q) {
        if (d[q] > rho) active.push_back(d[q] - rho);
        else below += 1;
    }
    if (active.empty()) {  // :106-112
        for (int64_t q = 0; q < k; ++q) d[q] = 1;
        return;
    }

    T sigma = active.back();  // :115-121
    T lo = 0;
    const T top = std::numeric_limits<T>::max();
    T hi = top;
    // log2 of an integer is taken in double, scaled, then stored as T (:127).
    const T target = static_cast<T>(std::log2(static_cast<double>(k + 1)) * bandwidth);

    for (int it = 0; it < 64; ++it) {  // :129-184
        T obs = below;
        T slope = 0;
        const T inv = 1 / sigma;
        const T inv2 = inv * inv;
        for (const T x : active) {
            const T e = std::exp(-x * inv);
            obs += e;
            slope += x * e * inv2;
        }
        const T diff = obs - target;
        if (std::abs(diff) < static_cast<T>(1e-5)) break;
        if (diff > 0) hi = sigma;
        else lo = sigma;

        bool took_newton = false;
        if (newton && slope) {
            const T cand = sigma - diff / slope;
            if (cand > lo && cand < hi) {
                sigma = cand;
                took_newton = true;
            }
        }
        if (!took_newton) {
            if (diff > 0) sigma += (lo - sigma) / 2;
            else if (hi == top) sigma *= 2;
            else sigma += (hi - sigma) / 2;
        }
    }

    T mean = 0;  // :187-192, zeros included
    for (int64_t q = 0; q < k; ++q) mean += d[q];
    mean /= static_cast<T>(k);
    sigma = std::max(min_k_dist_scale * mean, sigma);
    if (sigma_out) *sigma_out = static_cast<double>(sigma);

    const T inv = 1 / sigma;  // :194-202
    for (int64_t q = 0; q < k; ++q) d[q] = d[q] > rho ? std::exp(-(d[q] - rho) * inv) : static_cast<T>(1);
}

template <typename T>
int similarities(int64_t n, const int64_t* rp, T* val, double lc_in, double bw_in, double mkds_in, int newton,
                 double* rho, double* sigma, int nthreads) {
    // Options arrive as double and are stored as T, as initialize.hpp:71-74 does.
    const T lc = static_cast<T>(lc_in), bw = static_cast<T>(bw_in), mkds = static_cast<T>(mkds_in);
    if (!(lc >= 0) || !std::isfinite(lc)) return -1;
    const int64_t raw = static_cast<int64_t>(lc);  // truncation, as sanisizer::from_float does
    const T interp = lc - static_cast<T>(raw);
    if (nthreads < 1) nthreads = 1;
    auto work = [&](int64_t b, int64_t e) {
        std::vector<T> active;
        for (int64_t i = b; i < e; ++i)
            similarities_row<T>(val + rp[i], rp[i + 1] - rp[i], raw, interp, bw, mkds, newton != 0, active,
                                rho ? rho + i : nullptr, sigma ? sigma + i : nullptr);
    };
    if (nthreads == 1) {
        work(0, n);
    } else {
        std::vector<std::thread> pool;
        const int64_t chunk = (n + nthreads - 1) / nthreads;
        for (int t = 0; t < nthreads; ++t) {
            const int64_t b = std::min<int64_t>(n, t * chunk), e = std::min<int64_t>(n, b + chunk);
            pool.emplace_back(work, b, e);
        }
        for (auto& th : pool) th.join();
    }
    return 0;
}

// ---------------------------------------------------------------------------
// combine_neighbor_sets: include/umappp/combine_neighbor_sets.hpp:15-100
// Restated on a vector-of-rows so that appends behave as the reference's do.
// ---------------------------------------------------------------------------
template <typename T>
void combine(Rows<T>& x, const T mix) {
    const size_t n = x.size();
    std::vector<size_t> cursor(n, 0), orig(n, 0);
    for (size_t i = 0; i < n; ++i) {  // :21-25 (pairs sort by index first)
        std::sort(x[i].begin(), x[i].end());
        orig[i] = x[i].size();
    }
    for (size_t i = 0; i < n; ++i) {  // :27-72
        // Index-based walk: rows other than i may grow while row i is walked,
        // row i itself never does (no self entries).
        for (size_t q = 0; q < x[i].size(); ++q) {
            const int32_t t = x[i][q].first;
            auto& other = x[static_cast<size_t>(t)];
            size_t& c = cursor[static_cast<size_t>(t)];
            const size_t lim = orig[static_cast<size_t>(t)];
            while (c < lim && other[c].first < static_cast<int32_t>(i)) ++c;
            if (c < lim && other[c].first == static_cast<int32_t>(i)) {
                if (static_cast<int64_t>(i) < t) {
                    const T a = x[i][q].second, b = other[c].second;
                    const T prod = a * b;
                    T merged;
                    if (mix == 1) merged = a + b - prod;
                    else if (mix == 0) merged = prod;
                    else merged = mix * (a + b - prod) + (1 - mix) * prod;
                    x[i][q].second = merged;
                    other[c].second = merged;
                }
            } else {
                if (mix == 1) {
                    other.emplace_back(static_cast<int32_t>(i), x[i][q].second);
                } else if (mix == 0) {
                    x[i][q].second = 0;
                } else {
                    x[i][q].second *= mix;
                    other.emplace_back(static_cast<int32_t>(i), x[i][q].second);
                }
            }
        }
    }
    if (mix == 0) {  // :74-90
        for (auto& r : x) r.erase(std::remove_if(r.begin(), r.end(), [](const std::pair<int32_t, T>& p) { return p.second == 0; }), r.end());
    }
    for (auto& r : x) std::sort(r.begin(), r.end());  // :92-96
}

template <typename T>
int64_t combine_csr(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, double mix, int64_t cap,
                    int64_t* orp, int32_t* oidx, T* oval) {
    for (int64_t i = 0; i < n; ++i)
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e)
            if (idx[e] < 0 || idx[e] >= n || idx[e] == i) return -1;
    auto rows = csr_to_rows<T>(n, rp, idx, val);
    combine<T>(rows, static_cast<T>(mix));  // initialize.hpp:77 casts mix_ratio to Float_
    int64_t total = 0;
    for (auto& r : rows) total += static_cast<int64_t>(r.size());
    if (total > cap) return -2;
    int64_t pos = 0;
    orp[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (auto& p : rows[static_cast<size_t>(i)]) {
            oidx[pos] = p.first;
            oval[pos] = p.second;
            ++pos;
        }
        orp[i + 1] = pos;
    }
    return total;
}

// ---------------------------------------------------------------------------
// similarities_to_epochs: include/umappp/optimize_layout.hpp:47-92
// ---------------------------------------------------------------------------
template <typename T>
int64_t to_epochs(int64_t n, const int64_t* rp, const int32_t* idx, const T* val, int num_epochs, double nsr_in,
                  int64_t* orp, int32_t* otg, T* eps, T* nxt, T* nxt_neg, T* maxed_out) {
    const T nsr = static_cast<T>(nsr_in);
    T maxed = 0;
    for (int64_t e = 0; e < rp[n]; ++e) maxed = std::max(maxed, val[e]);
    const T limit = maxed / static_cast<T>(num_epochs);
    int64_t pos = 0;
    orp[0] = 0;
    for (int64_t i = 0; i < n; ++i) {
        for (int64_t e = rp[i]; e < rp[i + 1]; ++e) {
            if (val[e] >= limit) {
                otg[pos] = idx[e];
                eps[pos] = maxed / val[e];
                ++pos;
            }
        }
        orp[i + 1] = pos;
    }
    for (int64_t e = 0; e < pos; ++e) {
        nxt[e] = eps[e];
        nxt_neg[e] = eps[e] / nsr;
    }
    if (maxed_out) *maxed_out = maxed;
    // The cast check of num_epochs * nsr (:89): reject if it does not fit an int.
    const T prod = static_cast<T>(num_epochs) * nsr;
    if (!(prod <= static_cast<T>(std::numeric_limits<int>::max()))) return -3;
    return pos;
}

// ---------------------------------------------------------------------------
// Negative-sample sources.
// ---------------------------------------------------------------------------

// Rejection sampler over std::mt19937_64 (stand-in for aarand::discrete_uniform,
// optimize_layout.hpp:164,465; NOT upstream-identical).  Values at or above the
// largest multiple of n below 2^64 are redrawn; the accepted value is taken mod n.
inline uint64_t discrete_uniform(std::mt19937_64& eng, uint64_t n) {
    const uint64_t rem = (std::numeric_limits<uint64_t>::max() % n + 1) % n;  // 2^64 mod n
    const uint64_t last_ok = std::numeric_limits<uint64_t>::max() - rem;
    uint64_t x = eng();
    while (x > last_ok) x = eng();
    return x % n;
}

// The project's counter-based sampler.  The draw for firing edge j (local
// index), epoch n, draw number p is a pure function of (seed, n, j, p):
//   ek = mix64(seed + GOLD * (n + 1))
//   ej = mix64(ek ^ (j * ODD))
//   N <= 2^24: r = lowbias32(lo32(ej) + 0x9e3779b9 * (p + 1)) ^ hi32(ej),
//              k = high 32 bits of r * N
//   N >  2^24: r = mix64(ej + GOLD * (p + 1)), k = high 64 bits of r * N
// mix64 is the splitmix64 output finaliser (Steele, Lea, Flood 2014) and
// lowbias32 C. Wellons' 32-bit integer hash.
inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}
constexpr uint64_t kGold = 0x9e3779b97f4a7c15ULL;
constexpr uint64_t kOdd = 0xd1b54a32d192ed03ULL;
inline uint64_t counter_epoch_key(uint64_t seed, int epoch) { return mix64(seed + kGold * (static_cast<uint64_t>(epoch) + 1)); }
inline uint64_t counter_edge_key(uint64_t ek, uint64_t j) { return mix64(ek ^ (j * kOdd)); }
inline uint32_t lowbias32(uint32_t x) {  // C. Wellons' 32-bit hash
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}
inline uint32_t counter_draw(uint64_t ej, uint32_t p, uint32_t n) {
    if (n <= (1u << 24)) {
        const uint32_t r = lowbias32(static_cast<uint32_t>(ej) + 0x9e3779b9U * (p + 1)) ^ static_cast<uint32_t>(ej >> 32);
        return static_cast<uint32_t>((static_cast<uint64_t>(r) * n) >> 32);
    }
    const uint64_t r = mix64(ej + kGold * (static_cast<uint64_t>(p) + 1));
    return static_cast<uint32_t>((static_cast<unsigned __int128>(r) * n) >> 64);
}

// ---------------------------------------------------------------------------
// Layout state (EpochData + engine), optimize_layout.hpp:27-45, Status.hpp:42-45.
// ---------------------------------------------------------------------------
enum RngKind { RNG_MT = 0, RNG_COUNTER = 1, RNG_REPLAY = 2 };

template <typename T>
struct Layout {
    int64_t n = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> tg;
    std::vector<T> eps, nxt, nxt_neg;
    T nsr = 5;
    int total_epochs = 0;
    int current_epoch = 0;
    int rng_kind = RNG_MT;
    uint64_t seed = 0;
    std::mt19937_64 engine;
    // replay input (consumed in draw order) and optional recording output
    std::vector<int32_t> replay;
    size_t replay_pos = 0;
    int32_t* rec = nullptr;
    int64_t rec_cap = 0, rec_len = 0;
    // per-epoch tallies: firing edges, draws, self-hits
    std::vector<int64_t> fires, draws, selfhits;
    // optional update trace: (kind 0 = attractive / 1 = repulsive, head, other)
    int64_t* trace = nullptr;
    int64_t trace_cap = 0, trace_len = 0;
    void note(int kind, int64_t head, int64_t other) {
        if (trace && trace_len < trace_cap) {
            trace[3 * trace_len] = kind;
            trace[3 * trace_len + 1] = head;
            trace[3 * trace_len + 2] = other;
        }
        ++trace_len;
    }

    uint64_t ek = 0;
    int32_t draw(int64_t edge, uint32_t p) {
        int32_t s;
        if (rng_kind == RNG_MT) s = static_cast<int32_t>(discrete_uniform(engine, static_cast<uint64_t>(n)));
        else if (rng_kind == RNG_COUNTER) s = static_cast<int32_t>(counter_draw(counter_edge_key(ek, static_cast<uint64_t>(edge)), p, static_cast<uint32_t>(n)));
        else s = replay_pos < replay.size() ? replay[replay_pos++] : 0;
        if (rec && rec_len < rec_cap) rec[rec_len] = s;
        ++rec_len;
        return s;
    }
};

template <typename T>
inline T sqdist(const T* l, const T* r, int D) {  // quick_squared_distance, :94-103
    T s = 0;
    for (int d = 0; d < D; ++d) {
        const T q = l[d] - r[d];
        s += q * q;
    }
    return std::max(std::numeric_limits<T>::epsilon(), s);
}
template <typename T>
inline T clip4(T v) { return std::min(std::max(v, static_cast<T>(-4)), static_cast<T>(4)); }  // :105-110

// Attractive coefficient with the reference's mixed precision: numerator in T,
// denominator promoted to double by the 1.0 literal (:148-149).
template <typename T>
inline T attract_coef(T d2, T a, T b) {
    const T pd2b = std::pow(d2, b);
    return static_cast<T>((-2 * a * b * pd2b) / (d2 * (a * pd2b + 1.0)));
}
template <typename T>
inline T repulse_coef(T d2, T a, T b, T gamma) {  // :171
    return static_cast<T>(2 * gamma * b / ((0.001 + d2) * (a * std::pow(d2, b) + 1.0)));
}
template <typename T>
inline T epoch_alpha(T alpha0, int n, int total) {  // :132-133
    const T ep = static_cast<T>(n);
    return static_cast<T>(alpha0 * (1.0 - ep / static_cast<T>(total)));
}

// Serial epochs: optimize_layout.hpp:116-185.
template <typename T>
void run_serial(Layout<T>& L, int D, T* emb, T a, T b, T gamma, T alpha0, int limit) {
    for (; L.current_epoch < limit; ++L.current_epoch) {
        const int n = L.current_epoch;
        const T ep = static_cast<T>(n);
        const T alpha = epoch_alpha<T>(alpha0, n, L.total_epochs);
        if (L.rng_kind == RNG_COUNTER) L.ek = counter_epoch_key(L.seed, n);
        int64_t f = 0, dr = 0, sh = 0;
        for (int64_t i = 0; i < L.n; ++i) {
            T* left = emb + i * D;
            for (int64_t j = L.rp[i]; j < L.rp[i + 1]; ++j) {
                if (L.nxt[j] > ep) continue;
                ++f;
                T* right = emb + static_cast<int64_t>(L.tg[j]) * D;
                L.note(0, i, L.tg[j]);
                const T gc = attract_coef<T>(sqdist<T>(left, right, D), a, b);
                for (int d = 0; d < D; ++d) {
                    const T g = alpha * clip4<T>(gc * (left[d] - right[d]));
                    left[d] += g;
                    right[d] -= g;
                }
                const T epns = L.eps[j] / L.nsr;
                const int nneg = static_cast<int>((ep - L.nxt_neg[j]) / epns);
                for (int p = 0; p < nneg; ++p) {
                    const int32_t s = L.draw(j, static_cast<uint32_t>(p));
                    ++dr;
                    if (s == i) {
                        ++sh;
                        continue;
                    }
                    L.note(1, i, s);
                    const T* r2 = emb + static_cast<int64_t>(s) * D;
                    const T rc = repulse_coef<T>(sqdist<T>(left, r2, D), a, b, gamma);
                    for (int d = 0; d < D; ++d) left[d] += alpha * clip4<T>(rc * (left[d] - r2[d]));
                }
                L.nxt[j] += L.eps[j];
                L.nxt_neg[j] += nneg * epns;
            }
        }
        L.fires.push_back(f);
        L.draws.push_back(dr);
        L.selfhits.push_back(sh);
    }
}

// Conflict-scheduled parallel epochs: optimize_layout.hpp:191-560.
// The dispatcher draws every sample serially, tracks which observations the
// in-flight jobs read or write, and only hands a job to a spinning worker when
// it cannot interact with any job still running; the result equals the serial
// one bit for bit.
template <typename T>
struct Job {
    int64_t obs = 0;
    int64_t first_edge = 0;
    T alpha = 0;
    std::vector<int32_t> samples;  // accepted samples (self-hits dropped)
    std::vector<int> per_edge;     // accepted samples per edge, -1 = edge not firing
};

template <typename T>
void apply_job(const Job<T>& job, int D, T* emb, const Layout<T>& L, T a, T b, T gamma, T* self) {
    T* src = emb + job.obs * D;
    std::copy(src, src + D, self);  // :220-221
    size_t pos = 0;
    for (size_t q = 0; q < job.per_edge.size(); ++q) {
        const int cnt = job.per_edge[q];
        if (cnt < 0) continue;
        const int64_t j = job.first_edge + static_cast<int64_t>(q);
        T* right = emb + static_cast<int64_t>(L.tg[j]) * D;
        const T gc = attract_coef<T>(sqdist<T>(self, right, D), a, b);
        for (int d = 0; d < D; ++d) {
            const T g = job.alpha * clip4<T>(gc * (self[d] - right[d]));
            self[d] += g;
            right[d] -= g;
        }
        for (int c = 0; c < cnt; ++c, ++pos) {
            const T* r2 = emb + static_cast<int64_t>(job.samples[pos]) * D;
            const T rc = repulse_coef<T>(sqdist<T>(self, r2, D), a, b, gamma);
            for (int d = 0; d < D; ++d) self[d] += job.alpha * clip4<T>(rc * (self[d] - r2[d]));
        }
    }
    std::copy(self, self + D, src);  // :265
}

template <typename T>
struct Spinner {
    std::atomic<bool> go{false};
    bool quit = false;
    Job<T>* job = nullptr;
};

template <typename T>
void run_parallel(Layout<T>& L, int D, T* emb, T a, T b, T gamma, T alpha0, int limit, int nthreads) {
    const int nworkers = nthreads - 1;
    std::vector<std::unique_ptr<Spinner<T>>> spin;
    std::vector<std::thread> workers;
    for (int w = 0; w < nworkers; ++w) spin.emplace_back(new Spinner<T>());
    for (int w = 0; w < nworkers; ++w) {
        workers.emplace_back([&, w]() {
            std::vector<T> self(static_cast<size_t>(D));
            Spinner<T>& s = *spin[static_cast<size_t>(w)];
            for (;;) {
                while (!s.go.load(std::memory_order_acquire)) {
                }
                if (s.quit) break;
                apply_job<T>(*s.job, D, emb, L, a, b, gamma, self.data());
                s.go.store(false, std::memory_order_release);
            }
        });
    }
    auto wait_for = [&](int w) {
        while (spin[static_cast<size_t>(w)]->go.load(std::memory_order_acquire)) {
        }
    };

    std::vector<Job<T>> slots(static_cast<size_t>(nthreads));
    std::vector<Job<T>*> worker_job(static_cast<size_t>(nworkers));
    for (int w = 0; w < nworkers; ++w) worker_job[static_cast<size_t>(w)] = &slots[static_cast<size_t>(w)];
    Job<T>* pending = &slots.back();
    std::vector<T> main_self(static_cast<size_t>(D));

    // Per observation: 1-based id of the last job that touched it (0 = none in
    // this epoch) and whether that touch was a write.
    std::vector<int64_t> stamp(static_cast<size_t>(L.n));
    std::vector<unsigned char> wrote(static_cast<size_t>(L.n));

    for (; L.current_epoch < limit; ++L.current_epoch) {
        const int n = L.current_epoch;
        const T ep = static_cast<T>(n);
        const T alpha = epoch_alpha<T>(alpha0, n, L.total_epochs);
        if (L.rng_kind == RNG_COUNTER) L.ek = counter_epoch_key(L.seed, n);
        std::fill(stamp.begin(), stamp.end(), 0);
        int64_t base = 1;  // jobs with id >= base may still be running
        int busy = 0;
        int64_t i = 0;
        int64_t f = 0, dr = 0, sh = 0;

        while (i < L.n) {
            bool clear = true;
            for (int t = busy; t < nthreads; ++t) {
                Job<T>& job = *pending;
                job.alpha = alpha;
                job.obs = i;
                job.samples.clear();
                job.per_edge.clear();
                const int64_t me = i + 1;

                if (stamp[static_cast<size_t>(i)] >= base) clear = false;  // self is written
                stamp[static_cast<size_t>(i)] = me;
                wrote[static_cast<size_t>(i)] = 1;

                job.first_edge = L.rp[i];
                for (int64_t j = L.rp[i]; j < L.rp[i + 1]; ++j) {
                    if (L.nxt[j] > ep) {
                        job.per_edge.push_back(-1);
                        continue;
                    }
                    ++f;
                    const size_t tgt = static_cast<size_t>(L.tg[j]);
                    if (stamp[tgt] >= base && stamp[tgt] != me) clear = false;  // tail is written
                    stamp[tgt] = me;
                    wrote[tgt] = 1;

                    const size_t before = job.samples.size();
                    const T epns = L.eps[j] / L.nsr;
                    const int nneg = static_cast<int>((ep - L.nxt_neg[j]) / epns);
                    for (int p = 0; p < nneg; ++p) {
                        const int32_t s = L.draw(j, static_cast<uint32_t>(p));
                        ++dr;
                        if (s == i) {
                            ++sh;
                            continue;
                        }
                        job.samples.push_back(s);
                        const size_t su = static_cast<size_t>(s);
                        if (stamp[su] >= base) {
                            if (stamp[su] != me && wrote[su]) clear = false;  // read after another job's write
                        } else {
                            stamp[su] = me;
                            wrote[su] = 0;
                        }
                    }
                    job.per_edge.push_back(static_cast<int>(job.samples.size() - before));
                    L.nxt[j] += L.eps[j];
                    L.nxt_neg[j] += nneg * epns;
                }

                if (!clear) {
                    // This job waits for everything in flight and then runs first
                    // in the next round, so its reads must be recorded as its own.
                    for (int32_t s : job.samples) {
                        const size_t su = static_cast<size_t>(s);
                        if (stamp[su] != me) {
                            stamp[su] = me;
                            wrote[su] = 0;
                        }
                    }
                    break;
                }
                if (t + 1 == nthreads) {
                    apply_job<T>(job, D, emb, L, a, b, gamma, main_self.data());
                } else {
                    std::swap(worker_job[static_cast<size_t>(t)], pending);
                    spin[static_cast<size_t>(t)]->job = worker_job[static_cast<size_t>(t)];
                    spin[static_cast<size_t>(t)]->go.store(true, std::memory_order_release);
                    ++busy;
                }
                ++i;
                if (i == L.n) break;
            }
            for (int w = 0; w < busy; ++w) wait_for(w);
            base = i + 1;
            if (!clear) {
                std::swap(worker_job[0], pending);
                spin[0]->job = worker_job[0];
                spin[0]->go.store(true, std::memory_order_release);
                busy = 1;
                ++i;
            } else {
                busy = 0;
            }
        }
        for (int w = 0; w < busy; ++w) wait_for(w);
        L.fires.push_back(f);
        L.draws.push_back(dr);
        L.selfhits.push_back(sh);
    }

    for (int w = 0; w < nworkers; ++w) {
        wait_for(w);
        spin[static_cast<size_t>(w)]->quit = true;
        spin[static_cast<size_t>(w)]->go.store(true, std::memory_order_release);
    }
    for (auto& th : workers) th.join();
}

// ---------------------------------------------------------------------------
// Test-data helpers (reference test fixtures are regenerated from their seeds).
// ---------------------------------------------------------------------------
void gen_normal(uint64_t seed, int64_t count, double* out) {
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> dist(0, 1);
    for (int64_t q = 0; q < count; ++q) out[q] = dist(rng);
}
void gen_uniform01(uint64_t seed, int64_t count, double* out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> dist(0, 1);
    for (int64_t q = 0; q < count; ++q) out[q] = dist(rng);
}

// Exact Euclidean kNN by brute force, self excluded; distances are the square
// root of a sequential sum of squared differences; ties broken by index.
// Queries q0 .. q1 - 1 against all n points; row i - q0 of the outputs.
void knn_bruteforce(int64_t n, int dim, const double* x, int k, int64_t q0, int64_t q1, int nthreads,
                    int32_t* out_idx, double* out_dist) {
    auto work = [&](int64_t b, int64_t e) {
        std::vector<std::pair<double, int32_t>> cand(static_cast<size_t>(n));
        for (int64_t i = b; i < e; ++i) {
            size_t c = 0;
            for (int64_t j = 0; j < n; ++j) {
                if (j == i) continue;
                double s = 0;
                for (int d = 0; d < dim; ++d) {
                    const double q = x[i * dim + d] - x[j * dim + d];
                    s += q * q;
                }
                cand[c++] = {s, static_cast<int32_t>(j)};
            }
            const size_t kk = std::min<size_t>(static_cast<size_t>(k), c);
            std::partial_sort(cand.begin(), cand.begin() + static_cast<long>(kk), cand.begin() + static_cast<long>(c));
            for (size_t q = 0; q < kk; ++q) {
                out_idx[(i - q0) * k + static_cast<int64_t>(q)] = cand[q].second;
                out_dist[(i - q0) * k + static_cast<int64_t>(q)] = std::sqrt(cand[q].first);
            }
        }
    };
    if (nthreads < 1) nthreads = 1;
    std::vector<std::thread> pool;
    const int64_t chunk = (q1 - q0 + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; ++t) {
        const int64_t b = std::min<int64_t>(q1, q0 + t * chunk), e = std::min<int64_t>(q1, b + chunk);
        pool.emplace_back(work, b, e);
    }
    for (auto& th : pool) th.join();
}

}  // namespace

template <typename T>
Layout<T>* make_layout(int64_t n, const int64_t* rp, const int32_t* tg, const T* eps, const T* nxt,
                              const T* nxt_neg, double nsr, int total_epochs, int current_epoch, int rng_kind,
                              uint64_t seed) {
    auto* L = new Layout<T>();
    L->n = n;
    L->rp.assign(rp, rp + n + 1);
    const int64_t E = rp[n];
    L->tg.assign(tg, tg + E);
    L->eps.assign(eps, eps + E);
    L->nxt.assign(nxt, nxt + E);
    L->nxt_neg.assign(nxt_neg, nxt_neg + E);
    L->nsr = static_cast<T>(nsr);
    L->total_epochs = total_epochs;
    L->current_epoch = current_epoch;
    L->rng_kind = rng_kind;
    L->seed = seed;
    L->engine.seed(seed);
    return L;
}

// One update's move of the head, as a pure function of the two points'
// coordinates before the step (optimize_layout.hpp:145-158 and :169-175):
// kind 0 = attractive (the tail moves by the negative), 1 = repulsive.
template <typename T>
void step_delta(int kind, int D, const T* head, const T* other, T a, T b, T gamma, T alpha, T* delta) {
    const T d2 = sqdist<T>(head, other, D);
    const T c = kind == 0 ? attract_coef<T>(d2, a, b) : repulse_coef<T>(d2, a, b, gamma);
    for (int d = 0; d < D; ++d) delta[d] = alpha * clip4<T>(c * (head[d] - other[d]));
}
struct OrcLayout {
    int dtype;
    void* impl;
};

// ---------------------------------------------------------------------------
// C entry points (dtype: 0 = float, 1 = double)
// ---------------------------------------------------------------------------
namespace {
// ---------------------------------------------------------------------------
// UMAP cross-entropy loss over an edge list (the project's autograd loss).
//   L = sum_e c_e [ log(1 + a s^b)
//                   + gamma * sum_{p < m_e, k != head} (log(1 + a s_k^b) - log a - b log s_k) ]
// s = max(FLT_EPSILON, |y_head - y_other|^2) as in optimize_layout.hpp:94-103.
// The attractive term's gradient with respect to the head is minus the
// reference's attractive step direction (optimize_layout.hpp:145-158); the
// repulsive term's is that of :160-176 with s in place of 0.001 + s.
// Negative p of edge e is the counter sampler's draw for (seed, step, key_e,
// p), key_e = keys[e] (or e); a draw equal to the head is skipped.
// ---------------------------------------------------------------------------
struct LossIn {
    int64_t n;
    int dim;
    const double* emb;
    int64_t nnz;
    const int32_t* heads;
    const int32_t* tails;
    const int64_t* keys;
    const double* coef;
    const int32_t* nneg;
    int m;
    double a, b, gamma;
    uint64_t seed;
    int step;
};

double loss_sqd(const LossIn& L, int64_t i, int64_t j) {
    double s = 0;
    for (int d = 0; d < L.dim; ++d) {
        const double q = L.emb[i * L.dim + d] - L.emb[j * L.dim + d];
        s += q * q;
    }
    return std::max(s, static_cast<double>(std::numeric_limits<float>::epsilon()));
}
double loss_attr(const LossIn& L, double s) { return std::log1p(L.a * std::pow(s, L.b)); }
double loss_rep(const LossIn& L, double s) {
    return L.gamma * (std::log1p(L.a * std::pow(s, L.b)) - std::log(L.a) - L.b * std::log(s));
}
int64_t loss_draw(const LossIn& L, int64_t e, int p) {
    const uint64_t key = L.keys ? static_cast<uint64_t>(L.keys[e]) : static_cast<uint64_t>(e);
    return counter_draw(counter_edge_key(counter_epoch_key(L.seed, L.step), key), static_cast<uint32_t>(p),
                        static_cast<uint32_t>(L.n));
}
int loss_m(const LossIn& L, int64_t e) { return L.nneg ? L.nneg[e] : L.m; }
double loss_c(const LossIn& L, int64_t e) { return L.coef ? L.coef[e] : 1.0; }

double loss_total(const LossIn& L) {
    double tot = 0;
    for (int64_t e = 0; e < L.nnz; ++e) {
        const int64_t i = L.heads[e];
        double t = loss_attr(L, loss_sqd(L, i, L.tails[e]));
        for (int p = 0; p < loss_m(L, e); ++p) {
            const int64_t k = loss_draw(L, e, p);
            if (k == i) continue;
            t += loss_rep(L, loss_sqd(L, i, k));
        }
        tot += loss_c(L, e) * t;
    }
    return tot;
}

LossIn loss_in(int64_t n, int dim, const double* emb, int64_t nnz, const int32_t* heads, const int32_t* tails,
               const int64_t* keys, const double* coef, const int32_t* nneg, int m, double a, double b,
               double gamma, uint64_t seed, int step) {
    return LossIn{n, dim, emb, nnz, heads, tails, keys, coef, nneg, m, a, b, gamma, seed, step};
}

}  // namespace

extern "C" {

int orc_neighbor_similarities(int dtype, int64_t n, const int64_t* rp, void* val, double lc, double bw, double mkds,
                              int newton, double* rho, double* sigma, int nthreads) {
    if (dtype == 0) return similarities<float>(n, rp, static_cast<float*>(val), lc, bw, mkds, newton, rho, sigma, nthreads);
    return similarities<double>(n, rp, static_cast<double*>(val), lc, bw, mkds, newton, rho, sigma, nthreads);
}

int64_t orc_combine_neighbor_sets(int dtype, int64_t n, const int64_t* rp, const int32_t* idx, const void* val,
                                  double mix, int64_t cap, int64_t* orp, int32_t* oidx, void* oval) {
    if (dtype == 0) return combine_csr<float>(n, rp, idx, static_cast<const float*>(val), mix, cap, orp, oidx, static_cast<float*>(oval));
    return combine_csr<double>(n, rp, idx, static_cast<const double*>(val), mix, cap, orp, oidx, static_cast<double*>(oval));
}

int64_t orc_similarities_to_epochs(int dtype, int64_t n, const int64_t* rp, const int32_t* idx, const void* val,
                                   int num_epochs, double nsr, int64_t* orp, int32_t* otg, void* eps, void* nxt,
                                   void* nxt_neg, void* maxed) {
    if (dtype == 0)
        return to_epochs<float>(n, rp, idx, static_cast<const float*>(val), num_epochs, nsr, orp, otg, static_cast<float*>(eps),
                                static_cast<float*>(nxt), static_cast<float*>(nxt_neg), static_cast<float*>(maxed));
    return to_epochs<double>(n, rp, idx, static_cast<const double*>(val), num_epochs, nsr, orp, otg, static_cast<double*>(eps),
                             static_cast<double*>(nxt), static_cast<double*>(nxt_neg), static_cast<double*>(maxed));
}

void* orc_layout_create(int dtype, int64_t n, const int64_t* rp, const int32_t* tg, const void* eps, const void* nxt,
                        const void* nxt_neg, double nsr, int total_epochs, int current_epoch, int rng_kind,
                        uint64_t seed) {
    auto* h = new OrcLayout{dtype, nullptr};
    if (dtype == 0)
        h->impl = make_layout<float>(n, rp, tg, static_cast<const float*>(eps), static_cast<const float*>(nxt),
                                     static_cast<const float*>(nxt_neg), nsr, total_epochs, current_epoch, rng_kind, seed);
    else
        h->impl = make_layout<double>(n, rp, tg, static_cast<const double*>(eps), static_cast<const double*>(nxt),
                                      static_cast<const double*>(nxt_neg), nsr, total_epochs, current_epoch, rng_kind, seed);
    return h;
}

void orc_layout_destroy(void* hp) {
    auto* h = static_cast<OrcLayout*>(hp);
    if (!h) return;
    if (h->dtype == 0) delete static_cast<Layout<float>*>(h->impl);
    else delete static_cast<Layout<double>*>(h->impl);
    delete h;
}

int orc_layout_set_replay(void* hp, const int32_t* samples, int64_t count) {
    auto* h = static_cast<OrcLayout*>(hp);
    if (h->dtype == 0) {
        auto* L = static_cast<Layout<float>*>(h->impl);
        L->replay.assign(samples, samples + count);
        L->replay_pos = 0;
    } else {
        auto* L = static_cast<Layout<double>*>(h->impl);
        L->replay.assign(samples, samples + count);
        L->replay_pos = 0;
    }
    return 0;
}

int orc_layout_set_record(void* hp, int32_t* buf, int64_t cap) {
    auto* h = static_cast<OrcLayout*>(hp);
    if (h->dtype == 0) {
        auto* L = static_cast<Layout<float>*>(h->impl);
        L->rec = buf, L->rec_cap = cap, L->rec_len = 0;
    } else {
        auto* L = static_cast<Layout<double>*>(h->impl);
        L->rec = buf, L->rec_cap = cap, L->rec_len = 0;
    }
    return 0;
}

int64_t orc_layout_recorded(void* hp) {
    auto* h = static_cast<OrcLayout*>(hp);
    return h->dtype == 0 ? static_cast<Layout<float>*>(h->impl)->rec_len : static_cast<Layout<double>*>(h->impl)->rec_len;
}

int orc_layout_run(void* hp, int D, void* emb, double a, double b, double gamma, double alpha0, int epoch_limit,
                   int nthreads) {
    auto* h = static_cast<OrcLayout*>(hp);
    if (D < 1) return -1;
    if (h->dtype == 0) {
        auto* L = static_cast<Layout<float>*>(h->impl);
        if (epoch_limit > L->total_epochs || epoch_limit < L->current_epoch) return -2;
        if (nthreads <= 1) run_serial<float>(*L, D, static_cast<float*>(emb), static_cast<float>(a), static_cast<float>(b), static_cast<float>(gamma), static_cast<float>(alpha0), epoch_limit);
        else run_parallel<float>(*L, D, static_cast<float*>(emb), static_cast<float>(a), static_cast<float>(b), static_cast<float>(gamma), static_cast<float>(alpha0), epoch_limit, nthreads);
    } else {
        auto* L = static_cast<Layout<double>*>(h->impl);
        if (epoch_limit > L->total_epochs || epoch_limit < L->current_epoch) return -2;
        if (nthreads <= 1) run_serial<double>(*L, D, static_cast<double*>(emb), a, b, gamma, alpha0, epoch_limit);
        else run_parallel<double>(*L, D, static_cast<double*>(emb), a, b, gamma, alpha0, epoch_limit, nthreads);
    }
    return 0;
}

// Exports the schedule state and per-epoch tallies.  'tallies' holds
// 3 * epochs_run int64 values: fires, draws, self-hits per epoch run so far.
int orc_layout_state(void* hp, void* nxt, void* nxt_neg, int* current_epoch, int64_t* tallies, int64_t tally_cap) {
    auto* h = static_cast<OrcLayout*>(hp);
    auto dump = [&](auto* L, auto* a, auto* b) {
        if (nxt) std::copy(L->nxt.begin(), L->nxt.end(), a);
        if (nxt_neg) std::copy(L->nxt_neg.begin(), L->nxt_neg.end(), b);
        if (current_epoch) *current_epoch = L->current_epoch;
        if (tallies) {
            const int64_t m = static_cast<int64_t>(L->fires.size());
            for (int64_t e = 0; e < m && 3 * e + 2 < tally_cap; ++e) {
                tallies[3 * e] = L->fires[static_cast<size_t>(e)];
                tallies[3 * e + 1] = L->draws[static_cast<size_t>(e)];
                tallies[3 * e + 2] = L->selfhits[static_cast<size_t>(e)];
            }
        }
    };
    if (h->dtype == 0) dump(static_cast<Layout<float>*>(h->impl), static_cast<float*>(nxt), static_cast<float*>(nxt_neg));
    else dump(static_cast<Layout<double>*>(h->impl), static_cast<double*>(nxt), static_cast<double*>(nxt_neg));
    return 0;
}

// Serial-run trace of (kind, head, other) triples; returns via orc_layout_trace_len.
int orc_layout_set_trace(void* hp, int64_t* buf, int64_t cap) {
    auto* h = static_cast<OrcLayout*>(hp);
    if (h->dtype == 0) {
        auto* L = static_cast<Layout<float>*>(h->impl);
        L->trace = buf, L->trace_cap = cap, L->trace_len = 0;
    } else {
        auto* L = static_cast<Layout<double>*>(h->impl);
        L->trace = buf, L->trace_cap = cap, L->trace_len = 0;
    }
    return 0;
}
int64_t orc_layout_trace_len(void* hp) {
    auto* h = static_cast<OrcLayout*>(hp);
    return h->dtype == 0 ? static_cast<Layout<float>*>(h->impl)->trace_len : static_cast<Layout<double>*>(h->impl)->trace_len;
}

int orc_step_delta(int dtype, int kind, int D, const void* head, const void* other, double a, double b, double gamma,
                   double alpha, void* delta) {
    if (dtype == 0)
        step_delta<float>(kind, D, static_cast<const float*>(head), static_cast<const float*>(other), static_cast<float>(a),
                          static_cast<float>(b), static_cast<float>(gamma), static_cast<float>(alpha), static_cast<float*>(delta));
    else
        step_delta<double>(kind, D, static_cast<const double*>(head), static_cast<const double*>(other), a, b, gamma, alpha,
                           static_cast<double*>(delta));
    return 0;
}
// step_delta over many steps: kinds[q], head row q and other row q of the
// point-major arrays (D values each) into row q of delta.
int orc_step_delta_many(int dtype, int64_t count, const int32_t* kinds, int D, const void* head, const void* other,
                        double a, double b, double gamma, double alpha, void* delta) {
    for (int64_t q = 0; q < count; ++q) {
        if (dtype == 0)
            step_delta<float>(kinds[q], D, static_cast<const float*>(head) + q * D,
                              static_cast<const float*>(other) + q * D, static_cast<float>(a), static_cast<float>(b),
                              static_cast<float>(gamma), static_cast<float>(alpha), static_cast<float*>(delta) + q * D);
        else
            step_delta<double>(kinds[q], D, static_cast<const double*>(head) + q * D,
                               static_cast<const double*>(other) + q * D, a, b, gamma, alpha,
                               static_cast<double*>(delta) + q * D);
    }
    return 0;
}
double orc_epoch_alpha(int dtype, double alpha0, int n, int total) {
    return dtype == 0 ? static_cast<double>(epoch_alpha<float>(static_cast<float>(alpha0), n, total))
                      : epoch_alpha<double>(alpha0, n, total);
}

// Counter sampler exposed for known-answer tests.
uint32_t orc_counter_draw(uint64_t seed, int epoch, uint64_t edge, uint32_t p, uint32_t n) {
    return counter_draw(counter_edge_key(counter_epoch_key(seed, epoch), edge), p, n);
}
uint64_t orc_mt19937_64_nth(uint64_t seed, int64_t nth) {
    std::mt19937_64 eng(seed);
    uint64_t v = 0;
    for (int64_t q = 0; q < nth; ++q) v = eng();
    return v;
}
int orc_discrete_uniform_stream(uint64_t seed, uint64_t n, int64_t count, int64_t* out) {
    std::mt19937_64 eng(seed);
    for (int64_t q = 0; q < count; ++q) out[q] = static_cast<int64_t>(discrete_uniform(eng, n));
    return 0;
}

void orc_gen_normal(uint64_t seed, int64_t count, double* out) { gen_normal(seed, count, out); }
void orc_gen_uniform01(uint64_t seed, int64_t count, double* out) { gen_uniform01(seed, count, out); }
void orc_knn_bruteforce(int64_t n, int dim, const double* x, int k, int nthreads, int32_t* idx, double* dist) {
    knn_bruteforce(n, dim, x, k, 0, n, nthreads, idx, dist);
}
void orc_knn_bruteforce_range(int64_t n, int dim, const double* x, int k, int64_t q0, int64_t q1, int nthreads,
                              int32_t* idx, double* dist) {
    knn_bruteforce(n, dim, x, k, q0, q1, nthreads, idx, dist);
}

// Loss value (double).
double orc_loss(int64_t n, int dim, const double* emb, int64_t nnz, const int32_t* heads, const int32_t* tails,
                const int64_t* keys, const double* coef, const int32_t* nneg, int m, double a, double b, double gamma,
                uint64_t seed, int step) {
    return loss_total(loss_in(n, dim, emb, nnz, heads, tails, keys, coef, nneg, m, a, b, gamma, seed, step));
}

// Closed-form gradient of orc_loss with respect to the embedding (grad,
// n * dim) and to (a, b) (gab[2]).  A clamped distance (s = FLT_EPSILON)
// contributes nothing to the embedding gradient.  clip > 0: every term's
// gradient with respect to its head is clipped to [-clip, clip] per
// coordinate before the sum, the other point receiving minus the clipped
// value (per-sample clipping; clip = 4 is the reference's clip,
// optimize_layout.hpp:105-110).
void orc_loss_grad(int64_t n, int dim, const double* emb, int64_t nnz, const int32_t* heads, const int32_t* tails,
                   const int64_t* keys, const double* coef, const int32_t* nneg, int m, double a, double b,
                   double gamma, uint64_t seed, int step, double clip, double* grad, double* gab) {
    const LossIn L = loss_in(n, dim, emb, nnz, heads, tails, keys, coef, nneg, m, a, b, gamma, seed, step);
    std::fill(grad, grad + n * dim, 0.0);
    gab[0] = gab[1] = 0;
    const double feps = std::numeric_limits<float>::epsilon();
    auto pair = [&](int64_t i, int64_t j, double c, bool rep) {
        double s = 0;
        for (int d = 0; d < dim; ++d) {
            const double q = emb[i * dim + d] - emb[j * dim + d];
            s += q * q;
        }
        const bool clamped = !(s > feps);
        if (clamped) s = feps;
        const double sb = std::pow(s, b), den = 1.0 + a * sb, ls = std::log(s);
        double dlds;
        if (!rep) {
            dlds = c * a * b * sb / (s * den);
            gab[0] += c * sb / den;
            gab[1] += c * a * sb * ls / den;
        } else {
            dlds = -c * gamma * b / (s * den);
            gab[0] += c * gamma * (sb / den - 1.0 / a);
            gab[1] += -c * gamma * ls / den;
        }
        if (clamped) return;
        for (int d = 0; d < dim; ++d) {
            double g = 2.0 * dlds * (emb[i * dim + d] - emb[j * dim + d]);
            if (clip > 0) g = std::min(std::max(g, -clip), clip);
            grad[i * dim + d] += g;
            grad[j * dim + d] -= g;
        }
    };
    for (int64_t e = 0; e < nnz; ++e) {
        const int64_t i = heads[e];
        const double c = loss_c(L, e);
        pair(i, tails[e], c, false);
        for (int p = 0; p < loss_m(L, e); ++p) {
            const int64_t k = loss_draw(L, e, p);
            if (k != i) pair(i, k, c, true);
        }
    }
}

// Central finite differences of orc_loss, step h, with respect to
// emb[points[q] * dim + dims[q]] (out[q]).  Only the terms that involve the
// perturbed observation change, so each difference sums those terms (found
// once through per-observation incidence lists) at y + h and y - h.
void orc_loss_fd(int64_t n, int dim, const double* emb, int64_t nnz, const int32_t* heads, const int32_t* tails,
                 const int64_t* keys, const double* coef, const int32_t* nneg, int m, double a, double b,
                 double gamma, uint64_t seed, int step, int64_t count, const int64_t* points, const int32_t* dims,
                 double h, double* out) {
    const LossIn L0 = loss_in(n, dim, emb, nnz, heads, tails, keys, coef, nneg, m, a, b, gamma, seed, step);
    // incidence: attractive terms (edge e) and repulsive terms (edge e, draw p)
    std::vector<int64_t> cnt(static_cast<size_t>(n) + 1, 0);
    std::vector<std::pair<int64_t, int32_t>> terms;  // (edge, -1 attractive | draw p)
    std::vector<int64_t> owner;
    for (int64_t e = 0; e < nnz; ++e) {
        const int64_t i = heads[e];
        terms.emplace_back(e, -1);
        owner.push_back(i);
        terms.emplace_back(e, -1);
        owner.push_back(tails[e]);
        for (int p = 0; p < loss_m(L0, e); ++p) {
            const int64_t k = loss_draw(L0, e, p);
            if (k == i) continue;
            terms.emplace_back(e, p);
            owner.push_back(i);
            terms.emplace_back(e, p);
            owner.push_back(k);
        }
    }
    for (int64_t o : owner) ++cnt[o + 1];
    for (int64_t i = 0; i < n; ++i) cnt[i + 1] += cnt[i];
    std::vector<std::pair<int64_t, int32_t>> by(terms.size());
    {
        std::vector<int64_t> at(cnt.begin(), cnt.end() - 1);
        for (size_t t = 0; t < terms.size(); ++t) by[at[owner[t]]++] = terms[t];
    }
    std::vector<double> y(emb, emb + n * dim);
    LossIn L = L0;
    L.emb = y.data();
    auto local = [&](int64_t pt) {
        double tot = 0;
        for (int64_t t = cnt[pt]; t < cnt[pt + 1]; ++t) {
            const int64_t e = by[t].first;
            const int p = by[t].second;
            const int64_t i = heads[e];
            const double c = loss_c(L, e);
            if (p < 0) tot += c * loss_attr(L, loss_sqd(L, i, tails[e]));
            else tot += c * loss_rep(L, loss_sqd(L, i, loss_draw(L, e, p)));
        }
        return tot;
    };
    for (int64_t q = 0; q < count; ++q) {
        const int64_t pt = points[q];
        const int64_t at = pt * dim + dims[q];
        const double y0 = y[at];
        y[at] = y0 + h;
        const double up = local(pt);
        y[at] = y0 - h;
        const double dn = local(pt);
        y[at] = y0;
        out[q] = (up - dn) / (2.0 * h);
    }
}

}  // extern "C"
