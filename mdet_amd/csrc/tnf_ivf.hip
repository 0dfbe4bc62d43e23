// Cluster index of the learned proposal filter (no reference counterpart: the
// reference evaluates every pair).
//
// The filter's embedding puts a contig's partners close to it, but the
// embeddings do not form tiles compact enough to skip 32x32 blocks of pairs
// by bounds.  They do cluster: k-means with one centroid per ~192 contigs puts
// most partner pairs in the same cluster, and almost all of the rest in one of
// the few clusters whose centroids are nearest.  So the learned path evaluates
// its embedding Gram only on the pairs of probed cluster pairs:
//   1. k-means (Lloyd) on a sample of the rows, 4-8 iterations; centroid sums
//      in 2^-20 fixed point (integer atomics: the result does not depend on
//      the order of the additions), then one assignment of every row;
//   2. rows sorted by cluster (radix sort, stable), cluster offsets;
//   3. per cluster the n_probe nearest non-empty centroids (itself first);
//   4. on the host: the distinct unordered cluster pairs (a, b) with b in a's
//      probes or a in b's, each turned into prefilter items over a's rows x
//      b's rows (pairs i < j within one cluster).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "ivf.h"

namespace {

constexpr int E_D = 32;                 // embedding dimensions
constexpr uint32_t KM_MAX_K = 8192;
constexpr double KM_FIX = 1048576.0;    // fixed-point scale of the centroid sums
// Lloyd iterations x sampled rows per centroid, by the number of centroids.
// Measured: at 1M contigs (5208 centroids, 32 probes) the edge recall is
// 0.9983 with 4 x 16, 0.9992 with 8 x 32, 0.9993 with 8 x 64; at 200k
// contigs (1042 centroids) 4 x 16 keeps as many partner pairs as 8 x 32, and
// the sampled passes (bound by the centroid-sum atomics) cost 1.3 ms more.
constexpr uint32_t KM_FEW_CLUSTERS = 2048;
inline int km_iters(uint32_t k) { return k <= KM_FEW_CLUSTERS ? 4 : 8; }
inline uint32_t km_sample_per_cluster(uint32_t k) { return k <= KM_FEW_CLUSTERS ? 16 : 32; }

// centroid c <- row c * n / k (rows are in input order)
__global__ void k_km_init(const float* __restrict__ emb, uint32_t n, uint32_t k, float* __restrict__ cent) {
    const uint32_t c = blockIdx.x, d = threadIdx.x;
    if (c >= k || d >= (uint32_t)E_D) return;
    const uint64_t src = (uint64_t)c * n / k;
    cent[(size_t)c * E_D + d] = emb[src * E_D + d];
}

__global__ void k_km_norms(const float* __restrict__ cent, uint32_t k, float* __restrict__ cnorm) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= k) return;
    float s = 0.f;
#pragma unroll
    for (int d = 0; d < E_D; ++d) s = fmaf(cent[(size_t)c * E_D + d], cent[(size_t)c * E_D + d], s);
    cnorm[c] = s;
}

// Nearest centroid (smallest |c|^2 - 2 x.c, lowest index on ties) of the
// points q * stride, q < m.  Block = 64 points x KM_WAVES waves: lane l of
// every wave holds point l's 32 coordinates in registers, wave w scans the
// centroids c = w, w + KM_WAVES, ... (their coordinates are wave-uniform:
// scalar loads, FMAs from SGPRs; two centroids per iteration on independent
// accumulator chains), and the partial minima are merged in LDS.  More waves
// per point for the small sampled passes, whose latency would otherwise be
// exposed.  mode 1: counts[best] += 1;
// mode 2: also the point's coordinates into the fixed-point sums of its
// centroid.
template <int KM_WAVES>
__global__ __launch_bounds__(64 * KM_WAVES) void k_km_assign(const float* __restrict__ emb, uint32_t m,
                                                            uint32_t stride, const float* __restrict__ cent,
                                                            const float* __restrict__ cnorm, uint32_t k,
                                                            uint32_t* __restrict__ assign,
                                                            uint32_t* __restrict__ counts,
                                                            unsigned long long* __restrict__ sums, int mode) {
    __shared__ float sb[KM_WAVES][64];
    __shared__ uint32_t si[KM_WAVES][64];
    // wave-uniform by construction; readfirstlane makes it an SGPR value, so
    // the centroid addresses are uniform and load through the scalar unit
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t q = blockIdx.x * 64 + lane;
    const bool valid = q < m;
    const size_t p = (size_t)(valid ? q : 0) * stride;
    float x[E_D];
#pragma unroll
    for (int d = 0; d < E_D; ++d) x[d] = emb[p * E_D + d];
    float best = __builtin_inff();
    uint32_t bi = 0xffffffffu;
    uint32_t c = wave;
    for (; c + KM_WAVES < k; c += 2 * KM_WAVES) {
        const float* c0 = cent + (size_t)c * E_D;
        const float* c1 = c0 + KM_WAVES * E_D;
        float u0 = 0.f, u1 = 0.f, v0 = 0.f, v1 = 0.f;
#pragma unroll
        for (int d = 0; d < E_D; d += 2) {
            u0 = fmaf(x[d], c0[d], u0);
            u1 = fmaf(x[d + 1], c0[d + 1], u1);
            v0 = fmaf(x[d], c1[d], v0);
            v1 = fmaf(x[d + 1], c1[d + 1], v1);
        }
        const float du = fmaf(-2.f, u0 + u1, cnorm[c]);
        const float dv = fmaf(-2.f, v0 + v1, cnorm[c + KM_WAVES]);
        if (du < best) {
            best = du;
            bi = c;
        }
        if (dv < best) {
            best = dv;
            bi = c + KM_WAVES;
        }
    }
    if (c < k) {
        const float* c0 = cent + (size_t)c * E_D;
        float u0 = 0.f, u1 = 0.f;
#pragma unroll
        for (int d = 0; d < E_D; d += 2) {
            u0 = fmaf(x[d], c0[d], u0);
            u1 = fmaf(x[d + 1], c0[d + 1], u1);
        }
        const float du = fmaf(-2.f, u0 + u1, cnorm[c]);
        if (du < best) {
            best = du;
            bi = c;
        }
    }
    sb[wave][lane] = best;
    si[wave][lane] = bi;
    __syncthreads();
    if (wave != 0 || !valid) return;
#pragma unroll
    for (int w = 1; w < KM_WAVES; ++w) {
        const float b2 = sb[w][lane];
        const uint32_t i2 = si[w][lane];
        if (b2 < best || (b2 == best && i2 < bi)) {
            best = b2;
            bi = i2;
        }
    }
    if (bi >= k) bi = 0;  // k >= 1; only reachable with non-finite coordinates
    if (assign) assign[q] = bi;
    if (mode >= 1) atomicAdd(&counts[bi], 1u);
    if (mode >= 2)
#pragma unroll
        for (int d = 0; d < E_D; ++d)
            atomicAdd(&sums[(size_t)bi * E_D + d], (unsigned long long)llrint((double)x[d] * KM_FIX));
}

// centroid <- mean of its points (unchanged if it has none); clears the sums
// and counts for the next iteration.  32 threads (one per dimension) per
// centroid, |c|^2 by a half-wave reduction.
__global__ __launch_bounds__(256) void k_km_update(unsigned long long* __restrict__ sums, uint32_t* __restrict__ counts,
                                                   uint32_t k, float* __restrict__ cent, float* __restrict__ cnorm) {
    const uint32_t c = blockIdx.x * (blockDim.x / E_D) + threadIdx.x / E_D;
    const uint32_t d = threadIdx.x % E_D;
    const bool ok = c < k;
    const uint32_t cnt = ok ? counts[c] : 0u;
    float cv = 0.f;
    if (ok) {
        const size_t e = (size_t)c * E_D + d;
        const long long v = (long long)sums[e];
        sums[e] = 0ull;
        cv = cnt ? (float)((double)v / KM_FIX / (double)cnt) : cent[e];
        cent[e] = cv;
    }
    float s2 = cv * cv;
#pragma unroll
    for (int off = E_D / 2; off >= 1; off >>= 1) s2 += __shfl_xor(s2, off, E_D);
    __syncthreads();  // every lane has read counts[c] before it is cleared
    if (ok && d == 0) {
        cnorm[c] = s2;
        counts[c] = 0;
    }
}

// Per cluster a (one wave; 4 clusters per block): its n_probe (<= KM_MAX_PROBE)
// nearest non-empty centroids, itself first.  Each lane keeps the n_probe
// best of the centroids b = lane, lane + 64, ... in registers (insertion into
// a sorted list of (distance, index)), then the wave sorts the 64 x n_probe
// candidates in LDS and the first n_probe are the answer.
constexpr uint32_t KM_MAX_PROBE = 32;

__device__ inline void probe_wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(256) void k_km_probe(const float* __restrict__ cent, const uint32_t* __restrict__ counts,
                                                  uint32_t k, uint32_t n_probe, uint32_t* __restrict__ probes) {
    __shared__ float dk[4][64 * KM_MAX_PROBE];
    __shared__ uint32_t di[4][64 * KM_MAX_PROBE];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t a = blockIdx.x * 4 + wave;
    if (a >= k) return;  // no block-wide synchronisation below
    float bd[KM_MAX_PROBE];
    uint32_t bx[KM_MAX_PROBE];
#pragma unroll
    for (uint32_t t = 0; t < KM_MAX_PROBE; ++t) {
        bd[t] = __builtin_inff();
        bx[t] = 0xffffffffu;
    }
    for (uint32_t b = lane; b < k; b += 64) {
        if (counts[b] == 0) continue;
        float dist = -1.f;
        if (b != a) {
            float s = 0.f;
#pragma unroll
            for (int d = 0; d < E_D; ++d) {
                const float t = cent[(size_t)a * E_D + d] - cent[(size_t)b * E_D + d];
                s = fmaf(t, t, s);
            }
            dist = s;
        }
        // insert (dist, b) into the sorted list (b increases: ties keep the earlier)
        if (dist < bd[KM_MAX_PROBE - 1]) {
            float cd = dist;
            uint32_t cx = b;
#pragma unroll
            for (uint32_t t = 0; t < KM_MAX_PROBE; ++t) {
                if (cd < bd[t]) {
                    const float td = bd[t];
                    const uint32_t tx = bx[t];
                    bd[t] = cd;
                    bx[t] = cx;
                    cd = td;
                    cx = tx;
                }
            }
        }
    }
    float* sd = dk[wave];
    uint32_t* sx = di[wave];
    const uint32_t sz = 64 * KM_MAX_PROBE;
#pragma unroll
    for (uint32_t t = 0; t < KM_MAX_PROBE; ++t) {
        sd[lane * KM_MAX_PROBE + t] = bd[t];
        sx[lane * KM_MAX_PROBE + t] = bx[t];
    }
    probe_wave_sync();
    for (uint32_t s2 = 2; s2 <= sz; s2 <<= 1)
        for (uint32_t j = s2 >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < sz; t += 64) {
                const uint32_t u = t ^ j;
                if (u > t) {
                    const bool up = (t & s2) == 0;
                    const bool before = sd[u] < sd[t] || (sd[u] == sd[t] && sx[u] < sx[t]);
                    if (before == up) {
                        const float td = sd[t];
                        sd[t] = sd[u];
                        sd[u] = td;
                        const uint32_t tx = sx[t];
                        sx[t] = sx[u];
                        sx[u] = tx;
                    }
                }
            }
            probe_wave_sync();
        }
    for (uint32_t t = lane; t < n_probe; t += 64)
        probes[(size_t)a * n_probe + t] = sd[t] < __builtin_inff() ? sx[t] : 0xffffffffu;
}

__global__ void k_iota(uint32_t* __restrict__ v, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = i;
}

}  // namespace

namespace mdet {

constexpr int IVF_EVENT = 8;  // context event index marking the host copies

int ivf_launch(mdet_ctx* ctx, const float* emb, uint32_t n, uint32_t k, uint32_t n_probe, IvfIndex* out) {
    if (!out || !emb || n == 0 || n_probe == 0) MDET_FAIL(MDET_E_INVALID, "ivf_build: bad arguments");
    if (k == 0) k = std::max<uint32_t>(1, (n + 96) / 192);
    k = std::min<uint32_t>(std::min<uint32_t>(k, KM_MAX_K), n);
    if (n_probe > KM_MAX_PROBE)
        MDET_FAIL(MDET_E_RANGE, "mdet_tnf_edges_learned: n_probe %u above %u", n_probe, KM_MAX_PROBE);
    n_probe = std::min(n_probe, k);
    hipStream_t s = ctx->stream;
    float *cent, *cnorm;
    uint32_t *counts, *assign, *keys_sorted, *idx, *perm, *off, *probes;
    unsigned long long* sums;
    MDET_TRY(ctx_buf(ctx, "ivf.cent", (size_t)k * E_D, &cent));
    MDET_TRY(ctx_buf(ctx, "ivf.cnorm", (size_t)k, &cnorm));
    MDET_TRY(ctx_buf(ctx, "ivf.counts", (size_t)k + 1, &counts));
    MDET_TRY(ctx_buf(ctx, "ivf.sums", (size_t)k * E_D, &sums));
    MDET_TRY(ctx_buf(ctx, "ivf.assign", (size_t)n, &assign));
    MDET_TRY(ctx_buf(ctx, "ivf.keys", (size_t)n, &keys_sorted));
    MDET_TRY(ctx_buf(ctx, "ivf.idx", (size_t)n, &idx));
    MDET_TRY(ctx_buf(ctx, "ivf.perm", (size_t)n, &perm));
    MDET_TRY(ctx_buf(ctx, "ivf.off", (size_t)k + 1, &off));
    MDET_TRY(ctx_buf(ctx, "ivf.probes", (size_t)k * n_probe, &probes));

    // 1. k-means on a sample (every stride-th row), then every row
    hipLaunchKernelGGL(k_km_init, dim3(k), dim3(E_D), 0, s, emb, n, k, cent);
    hipLaunchKernelGGL(k_km_norms, dim3((k + 255) / 256), dim3(256), 0, s, cent, k, cnorm);
    MDET_HIP(hipMemsetAsync(sums, 0, (size_t)k * E_D * 8, s));
    MDET_HIP(hipMemsetAsync(counts, 0, ((size_t)k + 1) * 4, s));
    const uint32_t stride = std::max<uint32_t>(1, n / std::max<uint32_t>(1, k * km_sample_per_cluster(k)));
    const uint32_t m = (n - 1) / stride + 1;
    for (int it = 0; it < km_iters(k); ++it) {
        hipLaunchKernelGGL(k_km_assign<16>, dim3((m + 63) / 64), dim3(64 * 16), 0, s, emb, m, stride, cent, cnorm, k,
                           (uint32_t*)nullptr, counts, sums, 2);
        hipLaunchKernelGGL(k_km_update, dim3((k + 7) / 8), dim3(256), 0, s, sums, counts, k, cent, cnorm);
    }
    hipLaunchKernelGGL(k_km_assign<4>, dim3((n + 63) / 64), dim3(64 * 4), 0, s, emb, n, 1u, cent, cnorm, k, assign,
                       counts, (unsigned long long*)nullptr, 1);
    MDET_HIP(hipGetLastError());

    // 2. rows sorted by cluster (stable), cluster offsets
    hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, idx, n);
    int end_bit = 1;
    while ((1u << end_bit) < k) ++end_bit;
    size_t tmp = 0;
    MDET_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp, assign, keys_sorted, idx, perm, (int)n, 0, end_bit, s));
    void* t;
    MDET_TRY(ctx_buf(ctx, "ivf.sorttmp", tmp, &t));
    MDET_HIP(hipcub::DeviceRadixSort::SortPairs(t, tmp, assign, keys_sorted, idx, perm, (int)n, 0, end_bit, s));
    size_t tmp2 = 0;
    MDET_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp2, counts, off, (int)k + 1, s));
    void* t2;
    MDET_TRY(ctx_buf(ctx, "ivf.scantmp", tmp2, &t2));
    MDET_HIP(hipcub::DeviceScan::ExclusiveSum(t2, tmp2, counts, off, (int)k + 1, s));

    // 3. probe lists
    hipLaunchKernelGGL(k_km_probe, dim3((k + 3) / 4), dim3(256), 0, s, cent, counts, k, n_probe, probes);
    MDET_HIP(hipGetLastError());

    // 4. copies for the host part
    void* hb;
    MDET_TRY(ctx_host_buf(ctx, "ivf.h_index", ((size_t)k + 1 + (size_t)k * n_probe) * 4, &hb));
    uint32_t* h_off = static_cast<uint32_t*>(hb);
    uint32_t* h_probes = h_off + k + 1;
    MDET_HIP(hipMemcpyAsync(h_off, off, ((size_t)k + 1) * 4, hipMemcpyDeviceToHost, s));
    MDET_HIP(hipMemcpyAsync(h_probes, probes, (size_t)k * n_probe * 4, hipMemcpyDeviceToHost, s));
    hipEvent_t ev;
    MDET_TRY(ctx_event(ctx, IVF_EVENT, &ev));
    MDET_HIP(hipEventRecord(ev, s));
    out->h_off = h_off;
    out->h_probes = h_probes;
    out->n_probe = n_probe;
    out->perm = perm;
    out->k = k;
    return MDET_OK;
}

int ivf_finish(mdet_ctx* ctx, IvfIndex* out) {
    hipEvent_t ev;
    MDET_TRY(ctx_event(ctx, IVF_EVENT, &ev));
    MDET_HIP(hipEventSynchronize(ev));
    const uint32_t k = out->k, n_probe = out->n_probe;
    const uint32_t* offh = out->h_off;
    const uint32_t* probeh = out->h_probes;
    std::vector<uint64_t> cp;
    cp.reserve((size_t)k * n_probe);
    for (uint32_t a = 0; a < k; ++a) {
        if (offh[a + 1] == offh[a]) continue;
        for (uint32_t p = 0; p < n_probe; ++p) {
            const uint32_t b = probeh[(size_t)a * n_probe + p];
            if (b == 0xffffffffu || b >= k || offh[b + 1] == offh[b]) continue;
            cp.push_back(((uint64_t)std::min(a, b) << 32) | std::max(a, b));
        }
    }
    std::sort(cp.begin(), cp.end());
    cp.erase(std::unique(cp.begin(), cp.end()), cp.end());
    std::vector<uint4> items;
    items.reserve(cp.size() * 2);
    uint64_t pairs = 0;
    for (uint64_t key : cp) {
        const uint32_t a = (uint32_t)(key >> 32), b = (uint32_t)key;
        const uint32_t a0 = offh[a], a1 = offh[a + 1], b0 = offh[b], b1 = offh[b + 1];
        for (uint32_t rs = a0; rs < a1;) {
            const uint32_t re = std::min<uint32_t>((rs & ~31u) + ITEM_ROWS, a1);
            if (a == b) {
                if (rs + 1 < a1) items.push_back(make_uint4(rs, re, rs, a1 | ITEM_TRI));
                // pairs i < j, i in [rs, re), j in [rs, a1): sum of (a1 - 1 - i)
                const uint64_t cnt = re - rs;
                pairs += cnt * (a1 - 1) - (cnt * (2 * (uint64_t)rs + cnt - 1)) / 2;
            } else {
                items.push_back(make_uint4(rs, re, b0, b1));
                pairs += (uint64_t)(re - rs) * (b1 - b0);
            }
            rs = re;
        }
    }
    out->pairs = pairs;
    out->probe_pairs = (uint32_t)cp.size();
    out->nitems = (uint32_t)items.size();
    MDET_TRY(ctx_buf(ctx, "ivf.items", std::max<size_t>(items.size(), 1), &out->items_dev));
    if (!items.empty()) {
        // the previous call's upload from this buffer finished before that call returned
        void* hb;
        MDET_TRY(ctx_host_buf(ctx, "ivf.h_items", items.size() * sizeof(uint4), &hb));
        memcpy(hb, items.data(), items.size() * sizeof(uint4));
        MDET_HIP(hipMemcpyAsync(out->items_dev, hb, items.size() * sizeof(uint4), hipMemcpyHostToDevice, ctx->stream));
    }
    return MDET_OK;
}

}  // namespace mdet
