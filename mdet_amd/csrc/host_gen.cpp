// Synthetic metagenome generator for the TNF path.
//
// The reference ships no data; these inputs are defined here.  Rules (all
// draws from per-genome / per-contig xoshiro256** streams seeded through
// splitmix64, so output is independent of thread count and machine):
//   genome g: gc ~ U[gc_lo, gc_hi]; for each of the 64 order-3 contexts, the
//     4 emission weights are Exp(1) draws (-ln U), i.e. Dirichlet(1) after
//     normalising; then weights of C and G are multiplied by gc and those of
//     A and T by (1 - gc), and the row is renormalised.
//   contig c: genome ~ U{0..n_genomes-1}; length = round(exp(ln(median) +
//     sigma*Z)), Z standard normal (Box-Muller), clipped to [len_min, len_max];
//     the first 3 bases are uniform, the rest follow the genome's chain.
//   N runs: after each run (and at the contig start) the gap to the next run
//     start is Geometric(q) with q = n_frac / E[run length]; run lengths are
//     U{1..n_run_max}; a run is cut at the contig end.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "host_common.h"

namespace {

inline uint64_t splitmix64(uint64_t& x) {
    uint64_t z = (x += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct Xoshiro {
    uint64_t s[4];
    explicit Xoshiro(uint64_t seed) {
        uint64_t x = seed;
        for (int i = 0; i < 4; ++i) s[i] = splitmix64(x);
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    // uniform in (0, 1)
    inline double uniform() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
};

// Stream ids keep the genome, contig-header and contig-body streams apart.
inline uint64_t stream_seed(uint64_t seed, uint64_t kind, uint64_t id) {
    uint64_t x = seed ^ (kind * 0xD1B54A32D192ED03ull);
    uint64_t a = splitmix64(x);
    uint64_t y = a ^ (id * 0x9E3779B97F4A7C15ull + kind);
    return splitmix64(y);
}

struct Genome {
    // cumulative emission thresholds per context (64 contexts x 3 cut points)
    double cut[64][3];
    double gc;
};

Genome make_genome(const mdet_gen_config& c, uint64_t g) {
    Xoshiro r(stream_seed(c.seed, 1, g));
    Genome G;
    G.gc = c.gc_lo + (c.gc_hi - c.gc_lo) * r.uniform();
    for (int ctx = 0; ctx < 64; ++ctx) {
        double w[4];
        for (int b = 0; b < 4; ++b) w[b] = -std::log(r.uniform());
        // A=0 C=1 G=2 T=3
        w[0] *= (1.0 - G.gc);
        w[3] *= (1.0 - G.gc);
        w[1] *= G.gc;
        w[2] *= G.gc;
        double s = w[0] + w[1] + w[2] + w[3];
        double acc = 0;
        for (int b = 0; b < 3; ++b) {
            acc += w[b] / s;
            G.cut[ctx][b] = acc;
        }
    }
    return G;
}

struct ContigHeader {
    uint64_t len;
    uint32_t genome;
};

ContigHeader make_header(const mdet_gen_config& c, uint64_t id) {
    Xoshiro r(stream_seed(c.seed, 2, id));
    ContigHeader h;
    h.genome = (uint32_t)std::min<uint64_t>((uint64_t)(r.uniform() * c.n_genomes), c.n_genomes - 1);
    double u1 = r.uniform(), u2 = r.uniform();
    double z = std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    double L = std::exp(std::log(c.len_median) + c.len_sigma * z);
    L = std::min(std::max(L, c.len_min), c.len_max);
    h.len = (uint64_t)std::llround(L);
    return h;
}

int check_config(const mdet_gen_config* c) {
    if (!c) MDET_FAIL(MDET_E_INVALID, "generator: null config");
    if (c->n_genomes == 0) MDET_FAIL(MDET_E_INVALID, "generator: n_genomes must be > 0");
    if (!(c->len_median > 0) || !(c->len_min >= 1) || !(c->len_max >= c->len_min))
        MDET_FAIL(MDET_E_INVALID, "generator: bad length parameters");
    if (!(c->n_frac >= 0 && c->n_frac < 0.5)) MDET_FAIL(MDET_E_INVALID, "generator: n_frac must be in [0, 0.5)");
    if (c->n_frac > 0 && c->n_run_max == 0) MDET_FAIL(MDET_E_INVALID, "generator: n_run_max must be > 0");
    if (!(c->gc_lo >= 0 && c->gc_hi <= 1 && c->gc_lo <= c->gc_hi))
        MDET_FAIL(MDET_E_INVALID, "generator: bad GC range");
    return MDET_OK;
}

// Emit the bases of contig id: callback(pos, code) for code in 0..3, or
// code 4 for an N.  Bases are produced in order.
template <class Emit>
void gen_contig(const mdet_gen_config& c, const std::vector<Genome>& genomes, uint64_t id,
                const ContigHeader& h, Emit&& emit) {
    Xoshiro r(stream_seed(c.seed, 3, id));
    const Genome& G = genomes[h.genome];
    double q = 0;
    if (c.n_frac > 0) q = c.n_frac / (0.5 * (1.0 + c.n_run_max));
    auto next_gap = [&]() -> uint64_t {
        if (q <= 0) return UINT64_MAX;
        double u = r.uniform();
        double g = std::floor(std::log(u) / std::log1p(-q));
        return g > 1e18 ? UINT64_MAX : (uint64_t)g;
    };
    uint64_t next_run = next_gap();
    uint64_t run_left = 0;
    unsigned ctx = 0;
    for (uint64_t p = 0; p < h.len; ++p) {
        unsigned b;
        if (p < 3) {
            b = (unsigned)(r.next() >> 62);
        } else {
            double u = r.uniform();
            const double* cut = G.cut[ctx];
            b = (u >= cut[0]) + (u >= cut[1]) + (u >= cut[2]);
        }
        ctx = ((ctx << 2) | b) & 63;
        if (run_left == 0 && p == next_run) {
            run_left = 1 + (uint64_t)(r.next() % c.n_run_max);
        }
        if (run_left) {
            emit(p, 4u);
            if (--run_left == 0) next_run = p + 1 + next_gap();
        } else {
            emit(p, b);
        }
    }
}

std::vector<Genome> make_genomes(const mdet_gen_config& c) {
    std::vector<Genome> g(c.n_genomes);
    for (uint32_t i = 0; i < c.n_genomes; ++i) g[i] = make_genome(c, i);
    return g;
}

}  // namespace

extern "C" {

int mdet_gen_lengths(const mdet_gen_config* cfg, uint64_t n, uint64_t* lengths, uint32_t* genome) {
    if (int e = check_config(cfg)) return e;
    if (n && !lengths) MDET_FAIL(MDET_E_INVALID, "mdet_gen_lengths: null lengths");
    for (uint64_t i = 0; i < n; ++i) {
        ContigHeader h = make_header(*cfg, i);
        lengths[i] = h.len;
        if (genome) genome[i] = h.genome;
    }
    return MDET_OK;
}

int mdet_gen_packed(const mdet_gen_config* cfg, uint64_t n, const uint64_t* base_off, uint32_t* packed,
                    uint64_t packed_words, uint64_t* runs, uint64_t runs_cap, uint64_t* n_runs,
                    int threads) {
    if (int e = check_config(cfg)) return e;
    if (!base_off || !n_runs) MDET_FAIL(MDET_E_INVALID, "mdet_gen_packed: null argument");
    const uint64_t total = base_off[n];
    if ((total + 15) / 16 > packed_words) MDET_FAIL(MDET_E_CAPACITY, "mdet_gen_packed: packed buffer too small");
    if (total && !packed) MDET_FAIL(MDET_E_INVALID, "mdet_gen_packed: null packed buffer");
    std::vector<Genome> genomes = make_genomes(*cfg);
    for (uint64_t i = 0; i < n; ++i) {
        ContigHeader h = make_header(*cfg, i);
        if (base_off[i + 1] - base_off[i] != h.len)
            MDET_FAIL(MDET_E_INVALID, "mdet_gen_packed: base_off[%llu] does not match the generated length",
                      (unsigned long long)i);
    }
    memset(packed, 0, sizeof(uint32_t) * ((total + 15) / 16));
    unsigned nt = threads > 0 ? (unsigned)threads : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 64) nt = 1;
    std::vector<std::vector<uint64_t>> tr(nt);
    auto work = [&](unsigned t) {
        uint64_t c0 = n * t / nt, c1 = n * (t + 1) / nt;
        std::vector<uint64_t>& rl = tr[t];
        for (uint64_t id = c0; id < c1; ++id) {
            ContigHeader h = make_header(*cfg, id);
            const uint64_t off = base_off[id];
            uint32_t word = 0;
            uint64_t wi = off / 16;
            bool in_run = false;
            uint64_t run_s = 0;
            auto flush = [&](bool edge) {
                if (edge) __atomic_fetch_or(&packed[wi], word, __ATOMIC_RELAXED);
                else packed[wi] = word;
            };
            gen_contig(*cfg, genomes, id, h, [&](uint64_t p, unsigned b) {
                uint64_t g = off + p;
                uint64_t w = g / 16;
                if (w != wi) {
                    flush(wi == off / 16);
                    wi = w;
                    word = 0;
                }
                if (b > 3) {
                    if (!in_run) { in_run = true; run_s = g; }
                    b = 0;
                } else if (in_run) {
                    rl.push_back(run_s);
                    rl.push_back(g);
                    in_run = false;
                }
                word |= (uint32_t)b << (2 * (g % 16));
            });
            if (h.len) flush(true);
            if (in_run) {
                rl.push_back(run_s);
                rl.push_back(off + h.len);
            }
        }
    };
    std::vector<std::thread> th;
    for (unsigned t = 1; t < nt; ++t) th.emplace_back(work, t);
    work(0);
    for (auto& x : th) x.join();
    uint64_t cnt = 0;
    for (unsigned t = 0; t < nt; ++t) {
        for (size_t k = 0; k < tr[t].size(); k += 2) {
            uint64_t s = tr[t][k], e = tr[t][k + 1];
            if (cnt && cnt <= runs_cap && runs[2 * (cnt - 1) + 1] == s) {
                runs[2 * (cnt - 1) + 1] = e;  // runs of adjacent contigs touch
                continue;
            }
            if (cnt < runs_cap) {
                runs[2 * cnt] = s;
                runs[2 * cnt + 1] = e;
            }
            ++cnt;
        }
    }
    *n_runs = cnt;
    if (cnt > runs_cap)
        MDET_FAIL(MDET_E_CAPACITY, "mdet_gen_packed: %llu N runs, capacity %llu", (unsigned long long)cnt,
                  (unsigned long long)runs_cap);
    return MDET_OK;
}

int mdet_gen_ascii(const mdet_gen_config* cfg, uint64_t contig, char* out, uint64_t len) {
    if (int e = check_config(cfg)) return e;
    ContigHeader h = make_header(*cfg, contig);
    if (len != h.len) MDET_FAIL(MDET_E_INVALID, "mdet_gen_ascii: contig %llu has length %llu, not %llu",
                                (unsigned long long)contig, (unsigned long long)h.len, (unsigned long long)len);
    if (!out && len) MDET_FAIL(MDET_E_INVALID, "mdet_gen_ascii: null output");
    // the genome tables depend on (seed, n_genomes, gc range) only; callers
    // write a FASTA one contig at a time, so keep the last set per thread
    thread_local struct {
        uint64_t seed = 0;
        uint32_t n_genomes = 0;
        double gc_lo = 0, gc_hi = 0;
        std::vector<Genome> g;
    } cache;
    if (cache.g.empty() || cache.seed != cfg->seed || cache.n_genomes != cfg->n_genomes || cache.gc_lo != cfg->gc_lo ||
        cache.gc_hi != cfg->gc_hi) {
        cache.g = make_genomes(*cfg);
        cache.seed = cfg->seed;
        cache.n_genomes = cfg->n_genomes;
        cache.gc_lo = cfg->gc_lo;
        cache.gc_hi = cfg->gc_hi;
    }
    const std::vector<Genome>& genomes = cache.g;
    static const char B[5] = {'A', 'C', 'G', 'T', 'N'};
    gen_contig(*cfg, genomes, contig, h, [&](uint64_t p, unsigned b) { out[p] = B[b]; });
    return MDET_OK;
}

}  // extern "C"
