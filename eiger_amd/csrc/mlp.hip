// mlp.hip -- learned strategy dispatch: sample features, a small MLP and its
// training, all as HIP kernels (no reference counterpart: the reference has
// only the rule auto_select_groupby_algo, operators/groupby/groupby.cu:68-119,
// which stays the fallback).
//
// features_kernel   one workgroup per query: a strided sample of up to 4096
//                   keys goes into an LDS hash table of (key, count); the
//                   workgroup derives 16 features (log rows, key width, sample
//                   distinct count, Chao1 and GEE distinct estimates, top-8
//                   frequency mass, entropy, singleton/doubleton fractions,
//                   key span, sortedness, ...).
// mlp               16 -> 64 -> 64 -> 4 (ReLU), fp32: outputs 0..2 are the
//                   logits of {HASH, SORT, PARTITION}, output 3 the predicted
//                   log2(group count) / 32.  Parameters are one flat array
//                   [W1 64x16 | b1 64 | W2 64x64 | b2 64 | W3 4x64 | b3 4].
// mlp_fwd_kernel    inference, 64 samples per workgroup, activations in LDS.
// mlp_grad_kernel   forward + backward of 64 samples per workgroup; writes the
//                   workgroup's partial gradient (no atomics: deterministic).
// mlp_adam_kernel   sums the partial gradients in a fixed order and applies
//                   one Adam step.
// Loss per sample: cross-entropy against a soft strategy target (the trace's
// per-strategy times turned into a distribution) plus lambda * squared error
// of the log-cardinality head; averaged over the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "eg_device.hpp"
#include "eg_internal.hpp"

namespace eg {
namespace {

constexpr int kIn = 16, kH = 64, kOut = 4;
constexpr int kSelOut = kOut + 1;  // select_kernel: the four outputs, then the sample's singleton fraction
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH, kW3 = kB2 + kH,
              kB3 = kW3 + kOut * kH, kParams = kB3 + kOut;  // 5508
static_assert(kParams == EG_MLP_PARAMS, "parameter layout");
constexpr int kS = 64;           // samples per workgroup
constexpr int kMB = 256;         // threads per MLP workgroup
constexpr int kSample = 4096;    // rows sampled per query
constexpr int kFSlots = 8192;    // LDS hash slots of the feature kernel

struct FQuery {
    const void* keys;
    int64_t rows;
    int64_t group_bound;
    int32_t key_bytes;
    int32_t nops;
    int32_t value_bytes;
    int32_t pad;
};

__device__ __forceinline__ float block_sum(float v, float* red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    float t = 0.f;
    for (int i = 0; i < nw; ++i) t += red[i];
    return t;
}

__device__ __forceinline__ uint32_t block_max_u32(uint32_t v, uint32_t* red)
{
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_xor(v, off);
        v = o > v ? o : v;
    }
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    uint32_t t = 0;
    for (int i = 0; i < nw; ++i) t = red[i] > t ? red[i] : t;
    return t;
}

constexpr int kFB = 1024;  // threads of the feature workgroup

struct FeatLds {
    unsigned long long skey[kFSlots];
    uint32_t scnt[kFSlots];
    float redf[4][kFB / 64];
    uint32_t redu[kFB / 64];
    uint32_t top[kFB / 64][8];  // each wave's eight largest (count, slot) words
    uint32_t s_asc;
    unsigned long long s_min, s_max;
};

// The wave's eight largest of the values v[0..7] of its lanes (distinct
// values), largest first, in every lane.
__device__ __forceinline__ void wave_top8(uint32_t (&v)[8], uint32_t (&best)[8])
{
#pragma unroll
    for (int r = 0; r < 8; ++r) {
        uint32_t mine = 0u;
#pragma unroll
        for (int i = 0; i < 8; ++i) mine = v[i] > mine ? v[i] : mine;
        uint32_t m = mine;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const uint32_t o = __shfl_xor(m, off);
            m = o > m ? o : m;
        }
        best[r] = m;
        // the one lane holding m drops it (values are distinct or zero)
#pragma unroll
        for (int i = 0; i < 8; ++i)
            if (m != 0u && v[i] == m) v[i] = 0u;
    }
}

// The 16 features of query q, written by thread 0 to o (kFB threads).
__device__ void compute_features(const FQuery& q, FeatLds& L, float* o)
{
    unsigned long long* skey = L.skey;
    uint32_t* scnt = L.scnt;
    uint32_t* redu = L.redu;
    uint32_t& s_asc = L.s_asc;
    unsigned long long& s_min = L.s_min;
    unsigned long long& s_max = L.s_max;
    const int t = threadIdx.x;
    for (int i = t; i < kFSlots; i += blockDim.x) {
        skey[i] = 0ull;
        scnt[i] = 0u;
    }
    if (t == 0) {
        s_asc = 0u;
        s_min = ~0ull;
        s_max = 0ull;
    }
    __syncthreads();
    const int64_t n = q.rows;
    const int ns = (int)(n < kSample ? n : kSample);
    const int64_t stride = n <= kSample ? 1 : n / kSample;
    // a key k is stored as k + 1 so that 0 marks an empty slot (the one key
    // with k + 1 == 0 takes a fixed stand-in tag)
    uint32_t asc = 0;
    // the sample's keys (and each one's successor) are loaded before any is
    // inserted, so the loads overlap
    constexpr int kPer = kSample / kFB;
    uint64_t ks[kPer], kn[kPer];
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = t + j * kFB;
        ks[j] = kn[j] = 0ull;
        if (i < ns) {
            const int64_t r = (int64_t)i * stride;
            const uint64_t raw = q.key_bytes == 4 ? (uint64_t)((const uint32_t*)q.keys)[r] : ((const uint64_t*)q.keys)[r];
            ks[j] = q.key_bytes == 4 ? (uint64_t)(int64_t)(int32_t)(uint32_t)raw : raw;
            if (i + 1 < ns) {
                const uint64_t raw2 = q.key_bytes == 4 ? (uint64_t)((const uint32_t*)q.keys)[r + stride]
                                                       : ((const uint64_t*)q.keys)[r + stride];
                kn[j] = q.key_bytes == 4 ? (uint64_t)(int64_t)(int32_t)(uint32_t)raw2 : raw2;
            }
        }
    }
    // span: a wave's minimum and maximum first, one LDS atomic each per wave
    unsigned long long wmin = ~0ull, wmax = 0ull;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = t + j * kFB;
        if (i >= ns) break;
        const unsigned long long ek = ks[j] ^ 0x8000000000000000ull;  // order-preserving for the span
        wmin = ek < wmin ? ek : wmin;
        wmax = ek > wmax ? ek : wmax;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        const unsigned long long a = __shfl_xor(wmin, off), b = __shfl_xor(wmax, off);
        wmin = a < wmin ? a : wmin;
        wmax = b > wmax ? b : wmax;
    }
    if ((t & 63) == 0 && wmax >= wmin) {
        atomicMin(&s_min, wmin);
        atomicMax(&s_max, wmax);
    }
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int i = t + j * kFB;
        if (i >= ns) break;
        const uint64_t k = ks[j];
        if (i + 1 < ns) asc += ((int64_t)kn[j] >= (int64_t)k) ? 1u : 0u;
        const uint64_t tag = k + 1 == 0 ? 0x5bd1e9955bd1e995ull : k + 1;  // never 0
        uint32_t h = (uint32_t)(dev::fmix64(k) >> 40) & (kFSlots - 1);
        for (int p = 0; p < kFSlots; ++p) {
            const unsigned long long cur = skey[h];
            if (cur == tag) break;
            if (cur == 0ull) {
                const unsigned long long prev = atomicCAS(&skey[h], 0ull, (unsigned long long)tag);
                if (prev == 0ull || prev == tag) break;
            }
            h = (h + 1) & (kFSlots - 1);
        }
        atomicAdd(&scnt[h], 1u);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) asc += __shfl_xor(asc, off);
    if ((t & 63) == 0 && asc != 0u) atomicAdd(&s_asc, asc);
    __syncthreads();

    // per-slot statistics
    float ndv = 0.f, f1 = 0.f, f2 = 0.f, ent = 0.f;
    uint32_t mx = 0;
    const float inv_ns = ns > 0 ? 1.f / (float)ns : 0.f;
    for (int i = t; i < kFSlots; i += blockDim.x) {
        const uint32_t c = scnt[i];
        if (c == 0u) continue;
        ndv += 1.f;
        f1 += c == 1u ? 1.f : 0.f;
        f2 += c == 2u ? 1.f : 0.f;
        const float p = (float)c * inv_ns;
        ent -= p * log2f(p);
        mx = c > mx ? c : mx;
    }
    // the four sums and the maximum in one exchange (the same shuffle tree
    // and wave order as block_sum, so the same floats)
    {
        float f[4] = {ndv, f1, f2, ent};
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
            for (int a = 0; a < 4; ++a) f[a] += __shfl_xor(f[a], off);
            const uint32_t o = __shfl_xor(mx, off);
            mx = o > mx ? o : mx;
        }
        const int w = t >> 6, nw = blockDim.x >> 6;
        if ((t & 63) == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a) L.redf[a][w] = f[a];
            redu[w] = mx;
        }
        // top-8 frequency mass: each wave's eight largest (count, slot) words
        // (count << 13 | (8191 - slot), distinct per slot), then the eight
        // largest of those
        constexpr int kPerT = kFSlots / kFB;
        static_assert(kPerT == 8, "eight slots a thread");
        uint32_t v[8], best[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int sl = t + i * kFB;
            const uint32_t c = scnt[sl];
            v[i] = c ? (c << 13) | (uint32_t)(kFSlots - 1 - sl) : 0u;
        }
        wave_top8(v, best);
        if ((t & 63) == 0) {
#pragma unroll
            for (int r = 0; r < 8; ++r) L.top[w][r] = best[r];
        }
        __syncthreads();
        for (int a = 0; a < 4; ++a) {
            float tot = 0.f;
            for (int i = 0; i < nw; ++i) tot += L.redf[a][i];
            f[a] = tot;
        }
        ndv = f[0];
        f1 = f[1];
        f2 = f[2];
        ent = f[3];
        uint32_t m = 0;
        for (int i = 0; i < nw; ++i) m = redu[i] > m ? redu[i] : m;
        mx = m;
    }
    if (t >= 64) return;
    // wave 0: the eight largest of the waves' candidates (two per lane)
    float top8 = 0.f;
    {
        const int nw = blockDim.x >> 6;
        uint32_t v[8], best[8];
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const int c = t + i * 64;  // candidate c: wave c / 8, rank c % 8
            v[i] = c < nw * 8 ? L.top[c >> 3][c & 7] : 0u;
        }
        wave_top8(v, best);
#pragma unroll
        for (int r = 0; r < 8; ++r) top8 += (float)(best[r] >> 13);
    }
    if (t != 0) return;
    const float rows = (float)n;
    const float nsf = (float)ns;
    const float chao = f2 > 0.f ? ndv + f1 * f1 / (2.f * f2) : ndv + f1 * (f1 - 1.f) * 0.5f;
    const float gee = (ns > 0 ? sqrtf(rows / nsf) : 0.f) * f1 + (ndv - f1);
    const float span = ns > 0 ? (float)(s_max - s_min) + 1.f : 0.f;
    const int64_t bnd = (q.group_bound > 0 && q.group_bound < n) ? q.group_bound : n;
    o[0] = log2f(rows + 1.f) / 32.f;
    o[1] = (float)q.key_bytes / 8.f;
    o[2] = log2f(ndv + 1.f) / 12.f;
    o[3] = log2f(fminf(chao, rows) + 1.f) / 32.f;
    o[4] = log2f(fminf(gee, rows) + 1.f) / 32.f;
    o[5] = top8 * inv_ns;
    o[6] = ns > 1 ? ent / log2f(nsf) : 0.f;
    o[7] = f1 * inv_ns;
    o[8] = f2 * inv_ns;
    o[9] = (float)mx * inv_ns;
    o[10] = log2f(span + 1.f) / 64.f;
    o[11] = log2f((float)bnd + 1.f) / 32.f;
    o[12] = (float)q.nops / 16.f;
    o[13] = (float)q.value_bytes / 32.f;
    o[14] = ns > 1 ? (float)s_asc / (nsf - 1.f) : 1.f;
    o[15] = nsf / (float)kSample;
}

// kFB threads per query.
__global__ __launch_bounds__(kFB) void features_kernel(const FQuery* __restrict__ qs, float* __restrict__ out)
{
    __shared__ FeatLds L;
    const FQuery q = qs[blockIdx.x];
    compute_features(q, L, out + (size_t)blockIdx.x * kIn);
}

// Features of one query and the MLP's four outputs in one launch (the
// dispatcher's path: one kernel, one 16-byte copy back).
__global__ __launch_bounds__(kFB) void select_kernel(FQuery q, const float* __restrict__ P, float* __restrict__ y)
{
    __shared__ FeatLds L;
    __shared__ float x[kIn], h1[kH], h2[kH];
    __shared__ float W[kParams];
    // the weights reach LDS while the sample is read (every thread's loads
    // are issued before compute_features' first barrier)
    for (int i = threadIdx.x; i < kParams; i += kFB) W[i] = P[i];
    compute_features(q, L, x);
    __syncthreads();
    const int u = threadIdx.x;
    if (u < kH) {
        float z = W[kB1 + u];
        for (int f = 0; f < kIn; ++f) z = fmaf(W[kW1 + u * kIn + f], x[f], z);
        h1[u] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    if (u < kH) {
        float z = W[kB2 + u];
        for (int v = 0; v < kH; ++v) z = fmaf(W[kW2 + u * kH + v], h1[v], z);
        h2[u] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    if (u < kOut) {
        float z = W[kB3 + u];
        for (int v = 0; v < kH; ++v) z = fmaf(W[kW3 + u * kH + v], h2[v], z);
        y[u] = z;
    }
    if (u == kOut) y[kOut] = x[7];  // the sample's singleton fraction
}

// ---- MLP kernels (256 threads, 64 samples per workgroup) ---------------------------
struct Act {
    float x[kS][kIn];
    float h1[kS][kH];
    float h2[kS][kH];
    float o[kS][kOut];
};

__device__ void mlp_forward_tile(const float* __restrict__ P, const float* __restrict__ X, int s0, int ns, Act& a)
{
    const int t = threadIdx.x;
    for (int i = t; i < kS * kIn; i += kMB) {
        const int s = i / kIn, f = i % kIn;
        a.x[s][f] = s < ns ? X[(size_t)(s0 + s) * kIn + f] : 0.f;
    }
    __syncthreads();
    for (int i = t; i < kS * kH; i += kMB) {
        const int s = i / kH, u = i % kH;
        float z = P[kB1 + u];
#pragma unroll
        for (int f = 0; f < kIn; ++f) z = fmaf(P[kW1 + u * kIn + f], a.x[s][f], z);
        a.h1[s][u] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    for (int i = t; i < kS * kH; i += kMB) {
        const int s = i / kH, u = i % kH;
        float z = P[kB2 + u];
#pragma unroll 8
        for (int v = 0; v < kH; ++v) z = fmaf(P[kW2 + u * kH + v], a.h1[s][v], z);
        a.h2[s][u] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    {
        const int s = t / kOut, o = t % kOut;  // 256 threads = 64 samples x 4 outputs
        float z = P[kB3 + o];
#pragma unroll 8
        for (int v = 0; v < kH; ++v) z = fmaf(P[kW3 + o * kH + v], a.h2[s][v], z);
        a.o[s][o] = z;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kMB) void mlp_fwd_kernel(const float* __restrict__ P, const float* __restrict__ X, int n,
                                                      float* __restrict__ Y)
{
    __shared__ Act a;
    const int s0 = blockIdx.x * kS;
    const int ns = min(kS, n - s0);
    mlp_forward_tile(P, X, s0, ns, a);
    const int s = threadIdx.x / kOut, o = threadIdx.x % kOut;
    if (s < ns) Y[(size_t)(s0 + s) * kOut + o] = a.o[s][o];
}

// target: [n][3] strategy distribution; card: [n] log2(groups)/32.
__global__ __launch_bounds__(kMB) void mlp_grad_kernel(const float* __restrict__ P, const float* __restrict__ X,
                                                       const float* __restrict__ target, const float* __restrict__ card,
                                                       const int* __restrict__ idx, int n, float lambda, float inv_batch,
                                                       float* __restrict__ partial, float* __restrict__ loss_part)
{
    __shared__ Act a;
    __shared__ float dz3[kS][kOut];
    __shared__ float dh2[kS][kH];
    __shared__ float red[4];
    const int t = threadIdx.x;
    const int s0 = blockIdx.x * kS;
    const int ns = min(kS, n - s0);
    // gather this tile's samples (idx: the shuffled minibatch order)
    for (int i = t; i < kS * kIn; i += kMB) {
        const int s = i / kIn, f = i % kIn;
        a.x[s][f] = s < ns ? X[(size_t)idx[s0 + s] * kIn + f] : 0.f;
    }
    __syncthreads();
    for (int i = t; i < kS * kH; i += kMB) {
        const int s = i / kH, u = i % kH;
        float z = P[kB1 + u];
#pragma unroll
        for (int f = 0; f < kIn; ++f) z = fmaf(P[kW1 + u * kIn + f], a.x[s][f], z);
        a.h1[s][u] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    for (int i = t; i < kS * kH; i += kMB) {
        const int s = i / kH, u = i % kH;
        float z = P[kB2 + u];
#pragma unroll 8
        for (int v = 0; v < kH; ++v) z = fmaf(P[kW2 + u * kH + v], a.h1[s][v], z);
        a.h2[s][u] = z > 0.f ? z : 0.f;
    }
    __syncthreads();
    {
        const int s = t / kOut, o = t % kOut;
        float z = P[kB3 + o];
#pragma unroll 8
        for (int v = 0; v < kH; ++v) z = fmaf(P[kW3 + o * kH + v], a.h2[s][v], z);
        a.o[s][o] = z;
    }
    __syncthreads();
    // output gradients and loss (one thread per sample)
    float l = 0.f;
    if (t < kS) {
        const int s = t;
        if (s < ns) {
            const int g = idx[s0 + s];
            const float z0 = a.o[s][0], z1 = a.o[s][1], z2 = a.o[s][2];
            const float m = fmaxf(z0, fmaxf(z1, z2));
            const float e0 = expf(z0 - m), e1 = expf(z1 - m), e2 = expf(z2 - m);
            const float se = e0 + e1 + e2;
            const float lse = m + logf(se);
            const float p0 = target[(size_t)g * 3 + 0], p1 = target[(size_t)g * 3 + 1], p2 = target[(size_t)g * 3 + 2];
            l = -(p0 * (z0 - lse) + p1 * (z1 - lse) + p2 * (z2 - lse));
            const float d = a.o[s][3] - card[g];
            l += lambda * d * d;
            dz3[s][0] = (e0 / se - p0) * inv_batch;
            dz3[s][1] = (e1 / se - p1) * inv_batch;
            dz3[s][2] = (e2 / se - p2) * inv_batch;
            dz3[s][3] = 2.f * lambda * d * inv_batch;
        } else {
            dz3[s][0] = dz3[s][1] = dz3[s][2] = dz3[s][3] = 0.f;
        }
    }
    l = block_sum(l, red);
    if (t == 0) loss_part[blockIdx.x] = l * inv_batch;
    __syncthreads();
    float* G = partial + (size_t)blockIdx.x * kParams;
    // layer 3
    {
        const int o = t / kH, v = t % kH;  // 256 = 4 x 64
        float gw = 0.f;
        for (int s = 0; s < kS; ++s) gw = fmaf(dz3[s][o], a.h2[s][v], gw);
        G[kW3 + o * kH + v] = gw;
        if (t < kOut) {
            float gb = 0.f;
            for (int s = 0; s < kS; ++s) gb += dz3[s][t];
            G[kB3 + t] = gb;
        }
    }
    for (int i = t; i < kS * kH; i += kMB) {
        const int s = i / kH, u = i % kH;
        float g = 0.f;
#pragma unroll
        for (int o = 0; o < kOut; ++o) g = fmaf(dz3[s][o], P[kW3 + o * kH + u], g);
        dh2[s][u] = a.h2[s][u] > 0.f ? g : 0.f;
    }
    __syncthreads();
    // layer 2: dW2[u][v] = sum_s dh2[s][u] h1[s][v]; dh1 into h2's buffer afterwards
    for (int i = t; i < kH * kH; i += kMB) {
        const int u = i / kH, v = i % kH;
        float g = 0.f;
        for (int s = 0; s < kS; ++s) g = fmaf(dh2[s][u], a.h1[s][v], g);
        G[kW2 + u * kH + v] = g;
    }
    if (t < kH) {
        float g = 0.f;
        for (int s = 0; s < kS; ++s) g += dh2[s][t];
        G[kB2 + t] = g;
    }
    __syncthreads();
    for (int i = t; i < kS * kH; i += kMB) {
        const int s = i / kH, v = i % kH;
        float g = 0.f;
#pragma unroll 8
        for (int u = 0; u < kH; ++u) g = fmaf(dh2[s][u], P[kW2 + u * kH + v], g);
        a.h2[s][v] = a.h1[s][v] > 0.f ? g : 0.f;  // dh1
    }
    __syncthreads();
    // layer 1
    for (int i = t; i < kH * kIn; i += kMB) {
        const int v = i / kIn, f = i % kIn;
        float g = 0.f;
        for (int s = 0; s < kS; ++s) g = fmaf(a.h2[s][v], a.x[s][f], g);
        G[kW1 + v * kIn + f] = g;
    }
    if (t < kH) {
        float g = 0.f;
        for (int s = 0; s < kS; ++s) g += a.h2[s][t];
        G[kB1 + t] = g;
    }
}

// Sums `parts` partial gradients (fixed order) into grad; with adam != 0 also
// applies one Adam step (step count `step`, 1-based).
__global__ void mlp_reduce_adam_kernel(float* __restrict__ P, float* __restrict__ M, float* __restrict__ V,
                                       const float* __restrict__ partial, int parts, float* __restrict__ grad,
                                       int adam, float lr, float b1, float b2, float eps, int step, float wd)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= kParams) return;
    float g = 0.f;
    for (int p = 0; p < parts; ++p) g += partial[(size_t)p * kParams + i];
    if (grad) grad[i] = g;
    if (!adam) return;
    g += wd * P[i];
    const float m = b1 * M[i] + (1.f - b1) * g;
    const float v = b2 * V[i] + (1.f - b2) * g * g;
    M[i] = m;
    V[i] = v;
    const float mh = m / (1.f - powf(b1, (float)step));
    const float vh = v / (1.f - powf(b2, (float)step));
    P[i] -= lr * mh / (sqrtf(vh) + eps);
}

__global__ void sum_kernel(const float* __restrict__ x, int n, float* __restrict__ out)
{
    __shared__ float red[4];
    float v = 0.f;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += x[i];
    v = block_sum(v, red);
    if (threadIdx.x == 0) *out = v;
}

}  // namespace

// ---- host side -----------------------------------------------------------------------
void mlp_features(const FQuery* dq, int64_t nq, float* dout, hipStream_t s)
{
    if (nq <= 0) return;
    KernelTimer timer("mlp_features", s);
    hipLaunchKernelGGL(features_kernel, dim3((unsigned)nq), dim3(kFB), 0, s, dq, dout);
    EG_LAUNCH_CHECK();
}

// Device-resident model used by the dispatcher (one per process; the
// dispatcher runs on the current device).
struct LoadedModel {
    std::mutex mu;
    DevBuf params;
    DevBuf out;                 // 4 floats
    float* host_out = nullptr;  // pinned, kSelOut floats
    bool loaded = false;
};
LoadedModel& loaded_model()
{
    static LoadedModel m;
    return m;
}

// Inputs below this many rows go to hash without sampling: the call's fixed
// costs dominate there (hash: one aggregation launch; sort and partition:
// three launches per radix pass), and the sample's features and inference
// (~0.08 ms, one synchronisation) would cost more than any choice saves.
constexpr int64_t kSmallRows = (int64_t)1 << 17;

// Strategy from the model's outputs o[0..3] (class scores of hash / sort /
// partition, log2 group count / 32).  A large input whose estimated group
// count fits one LDS table goes to hash: one streaming pass, where every other
// strategy makes radix passes (in the training and held-out traces hash was
// the fastest for all such queries: rows >= 2^24, estimate <= 2048).
int algo_of(const float* o, int64_t rows)
{
    if (rows < kSmallRows) return EG_ALGO_HASH;
    const double card = std::exp2(std::min(62.0, std::max(0.0, (double)o[3] * 32.0)));
    if (rows >= ((int64_t)1 << 24) && card <= 2048.0) return EG_ALGO_HASH;
    const int best = (o[1] > o[0] && o[1] >= o[2]) ? 1 : (o[2] > o[0] && o[2] > o[1]) ? 2 : 0;
    return best == 0 ? EG_ALGO_HASH : best == 1 ? EG_ALGO_SORT : EG_ALGO_PARTITION;
}

// A large input whose caller-given group bound already fits one LDS table
// goes to hash without sampling: the bound caps the true group count, so
// algo_of's rule would pick hash for any estimate the model could give.
bool bound_picks_hash(int64_t rows, int64_t group_bound)
{
    return rows >= ((int64_t)1 << 24) && group_bound > 0 && group_bound <= 2048;
}

// Learned strategy choice for one key column; false when no model is loaded.
bool mlp_select(const void* keys, int32_t key_type, int64_t rows, int64_t group_bound, int nops, int value_bytes,
                hipStream_t s, int32_t* algo, int64_t* card_est, float* singletons)
{
    LoadedModel& m = loaded_model();
    std::lock_guard<std::mutex> lk(m.mu);
    if (!m.loaded) return false;
    if (rows < kSmallRows || bound_picks_hash(rows, group_bound)) {
        // hash, sized by the caller's bound (or the row count)
        *algo = EG_ALGO_HASH;
        if (card_est) *card_est = rows < kSmallRows ? rows : group_bound;
        if (singletons) *singletons = 0.f;
        return true;
    }
    FQuery q{keys, rows, group_bound, type_size(key_type), nops, value_bytes, 0};
    {
        KernelTimer timer("mlp_select", s);
        hipLaunchKernelGGL(select_kernel, dim3(1), dim3(kFB), 0, s, q, (const float*)m.params.get<float>(),
                           m.out.get<float>());
        EG_LAUNCH_CHECK();
    }
    EG_HIP(hipMemcpyAsync(m.host_out, m.out.get(), kSelOut * 4, hipMemcpyDeviceToHost, s));
    EG_HIP(hipStreamSynchronize(s));
    float h[kSelOut];
    std::memcpy(h, m.host_out, sizeof h);
    for (float v : h)
        if (!std::isfinite(v)) return false;  // fall back to the rule
    *algo = algo_of(h, rows);
    if (card_est) {
        const double lg = std::min(62.0, std::max(0.0, (double)h[3] * 32.0));
        *card_est = (int64_t)std::min<double>((double)rows, std::exp2(lg));
    }
    if (singletons) *singletons = h[kOut];
    return true;
}

}  // namespace eg

extern "C" {

int eg_mlp_features(const eg_mlp_query* queries, int64_t num_queries, float* out_features, void* stream)
{
    return eg::guard([&] {
        if (num_queries < 0 || (num_queries > 0 && (!queries || !out_features))) eg::raise(EG_ERR_INVALID, "bad arguments");
        if (num_queries == 0) return;
        if (num_queries >= ((int64_t)1 << 31)) eg::raise(EG_ERR_UNSUPPORTED, "too many queries");
        hipStream_t s = (hipStream_t)stream;
        std::vector<eg::FQuery> h((size_t)num_queries);
        for (int64_t i = 0; i < num_queries; ++i) {
            const eg_mlp_query& u = queries[i];
            const int kb = eg::type_size(u.key_type);
            if (!(u.key_type == EG_INT32 || u.key_type == EG_INT64 || u.key_type == EG_UINT32 || u.key_type == EG_UINT64))
                eg::raise(EG_ERR_UNSUPPORTED, "features need 4- or 8-byte integer keys");
            if (u.rows < 0 || (u.rows > 0 && !u.keys)) eg::raise(EG_ERR_INVALID, "bad key column");
            h[(size_t)i] = eg::FQuery{u.keys, u.rows, u.group_bound, kb, u.nops, u.value_bytes, 0};
        }
        eg::DevBuf dq(h.size() * sizeof(eg::FQuery), s);
        EG_HIP(hipMemcpyAsync(dq.get(), h.data(), h.size() * sizeof(eg::FQuery), hipMemcpyHostToDevice, s));
        eg::mlp_features(dq.get<eg::FQuery>(), num_queries, out_features, s);
        EG_HIP(hipStreamSynchronize(s));
    });
}

int eg_mlp_forward(const float* params, const float* x, int64_t n, float* y, void* stream)
{
    return eg::guard([&] {
        if (n < 0 || (n > 0 && (!params || !x || !y))) eg::raise(EG_ERR_INVALID, "bad arguments");
        if (n == 0) return;
        if (n >= ((int64_t)1 << 31)) eg::raise(EG_ERR_UNSUPPORTED, "too many samples");
        hipStream_t s = (hipStream_t)stream;
        const unsigned g = (unsigned)((n + eg::kS - 1) / eg::kS);
        hipLaunchKernelGGL(eg::mlp_fwd_kernel, dim3(g), dim3(eg::kMB), 0, s, params, x, (int)n, y);
        EG_LAUNCH_CHECK();
        EG_HIP(hipStreamSynchronize(s));
    });
}

int eg_mlp_loss_grad(const float* params, const float* x, const float* target, const float* card, int64_t n,
                     float lambda, float* grad, float* loss, void* stream)
{
    return eg::guard([&] {
        if (n <= 0 || !params || !x || !target || !card || !grad) eg::raise(EG_ERR_INVALID, "bad arguments");
        if (n >= ((int64_t)1 << 24)) eg::raise(EG_ERR_UNSUPPORTED, "batch too large");
        hipStream_t s = (hipStream_t)stream;
        const int parts = (int)((n + eg::kS - 1) / eg::kS);
        std::vector<int> idx((size_t)n);
        for (int64_t i = 0; i < n; ++i) idx[(size_t)i] = (int)i;
        eg::DevBuf didx((size_t)n * 4, s), part((size_t)parts * eg::kParams * 4, s), lp((size_t)parts * 4, s), lsum(4, s);
        EG_HIP(hipMemcpyAsync(didx.get(), idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(eg::mlp_grad_kernel, dim3(parts), dim3(eg::kMB), 0, s, params, x, target, card,
                           (const int*)didx.get<int>(), (int)n, lambda, 1.f / (float)n, part.get<float>(), lp.get<float>());
        EG_LAUNCH_CHECK();
        hipLaunchKernelGGL(eg::mlp_reduce_adam_kernel, dim3((eg::kParams + 255) / 256), dim3(256), 0, s, (float*)nullptr,
                           (float*)nullptr, (float*)nullptr, (const float*)part.get<float>(), parts, grad, 0, 0.f, 0.f,
                           0.f, 0.f, 0, 0.f);
        EG_LAUNCH_CHECK();
        hipLaunchKernelGGL(eg::sum_kernel, dim3(1), dim3(256), 0, s, (const float*)lp.get<float>(), parts, lsum.get<float>());
        EG_LAUNCH_CHECK();
        float h = 0.f;
        EG_HIP(hipMemcpyAsync(&h, lsum.get(), 4, hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        if (loss) *loss = h;
    });
}

int eg_mlp_train(float* params, const float* x, const float* target, const float* card, int64_t n,
                 const eg_mlp_train_config* cfg, float* loss_per_epoch, void* stream)
{
    return eg::guard([&] {
        if (n <= 0 || !params || !x || !target || !card || !cfg) eg::raise(EG_ERR_INVALID, "bad arguments");
        if (n >= ((int64_t)1 << 30)) eg::raise(EG_ERR_UNSUPPORTED, "too many samples");
        if (cfg->batch < eg::kS || cfg->batch % eg::kS != 0 || cfg->epochs <= 0)
            eg::raise(EG_ERR_INVALID, "batch must be a positive multiple of 64 and epochs positive");
        hipStream_t s = (hipStream_t)stream;
        const int B = cfg->batch;
        const int parts = B / eg::kS;
        eg::DevBuf M((size_t)eg::kParams * 4, s), V((size_t)eg::kParams * 4, s);
        eg::fill_zero(M.get(), (size_t)eg::kParams * 4, s);
        eg::fill_zero(V.get(), (size_t)eg::kParams * 4, s);
        eg::DevBuf part((size_t)parts * eg::kParams * 4, s), lp((size_t)parts * 4, s), lsum(4, s), didx((size_t)n * 4, s);
        std::vector<int> idx((size_t)n);
        uint64_t rng = cfg->seed * 0x9E3779B97F4A7C15ull + 1;
        auto next = [&]() {
            rng ^= rng << 13;
            rng ^= rng >> 7;
            rng ^= rng << 17;
            return rng;
        };
        int step = 0;
        for (int e = 0; e < cfg->epochs; ++e) {
            for (int64_t i = 0; i < n; ++i) idx[(size_t)i] = (int)i;
            for (int64_t i = n - 1; i > 0; --i) std::swap(idx[(size_t)i], idx[(size_t)(next() % (uint64_t)(i + 1))]);
            EG_HIP(hipMemcpyAsync(didx.get(), idx.data(), (size_t)n * 4, hipMemcpyHostToDevice, s));
            double epoch_loss = 0.0;
            int nb = 0;
            for (int64_t b0 = 0; b0 + B <= n; b0 += B) {
                ++step;
                hipLaunchKernelGGL(eg::mlp_grad_kernel, dim3(parts), dim3(eg::kMB), 0, s, (const float*)params, x, target,
                                   card, (const int*)(didx.get<int>() + b0), B, cfg->lambda, 1.f / (float)B,
                                   part.get<float>(), lp.get<float>());
                EG_LAUNCH_CHECK();
                hipLaunchKernelGGL(eg::mlp_reduce_adam_kernel, dim3((eg::kParams + 255) / 256), dim3(256), 0, s, params,
                                   M.get<float>(), V.get<float>(), (const float*)part.get<float>(), parts, (float*)nullptr,
                                   1, cfg->lr, 0.9f, 0.999f, 1e-8f, step, cfg->weight_decay);
                EG_LAUNCH_CHECK();
                if (loss_per_epoch) {
                    hipLaunchKernelGGL(eg::sum_kernel, dim3(1), dim3(256), 0, s, (const float*)lp.get<float>(), parts,
                                       lsum.get<float>());
                    EG_LAUNCH_CHECK();
                    float h = 0.f;
                    EG_HIP(hipMemcpyAsync(&h, lsum.get(), 4, hipMemcpyDeviceToHost, s));
                    EG_HIP(hipStreamSynchronize(s));
                    epoch_loss += h;
                }
                ++nb;
            }
            if (loss_per_epoch) loss_per_epoch[e] = nb ? (float)(epoch_loss / nb) : 0.f;
            EG_HIP(hipStreamSynchronize(s));  // idx is rewritten next epoch
        }
    });
}

int eg_mlp_load(const float* host_params, int64_t n)
{
    return eg::guard([&] {
        eg::LoadedModel& m = eg::loaded_model();
        std::lock_guard<std::mutex> lk(m.mu);
        if (!host_params) {
            m.loaded = false;
            return;
        }
        if (n != eg::kParams) eg::raise(EG_ERR_INVALID, "wrong parameter count");
        for (int64_t i = 0; i < n; ++i)
            if (!std::isfinite(host_params[i])) eg::raise(EG_ERR_INVALID, "non-finite parameter");
        m.params = eg::DevBuf((size_t)n * 4, nullptr);
        EG_HIP(hipMemcpy(m.params.get(), host_params, (size_t)n * 4, hipMemcpyHostToDevice));
        m.out = eg::DevBuf(eg::kSelOut * 4, nullptr);
        if (!m.host_out) EG_HIP(hipHostMalloc((void**)&m.host_out, eg::kSelOut * 4, hipHostMallocDefault));
        m.loaded = true;
    });
}

int eg_mlp_select(eg_column_view key, int64_t group_bound, int32_t nops, int32_t value_bytes, void* stream,
                  int32_t* algo, int64_t* card_est)
{
    return eg::guard([&] {
        if (!algo) eg::raise(EG_ERR_INVALID, "null output");
        const int kb = eg::type_size(key.type);
        if (!(key.type == EG_INT32 || key.type == EG_INT64 || key.type == EG_UINT32 || key.type == EG_UINT64))
            eg::raise(EG_ERR_UNSUPPORTED, "features need 4- or 8-byte integer keys");
        if (key.size <= 0) eg::raise(EG_ERR_INVALID, "empty key column");
        const void* data = (const char*)key.data + key.offset * kb;
        if (!eg::mlp_select(data, key.type, key.size, group_bound, nops, value_bytes, (hipStream_t)stream, algo, card_est, nullptr))
            eg::raise(EG_ERR_INVALID, "no dispatch model is loaded");
    });
}

int eg_mlp_select_batch(const eg_mlp_query* queries, int64_t num, int32_t* algos, int64_t* card_est, void* stream)
{
    return eg::guard([&] {
        if (num < 0 || (num > 0 && (!queries || !algos))) eg::raise(EG_ERR_INVALID, "bad arguments");
        if (num == 0) return;
        if (num >= ((int64_t)1 << 31)) eg::raise(EG_ERR_UNSUPPORTED, "too many queries");
        eg::LoadedModel& m = eg::loaded_model();
        std::lock_guard<std::mutex> lk(m.mu);
        if (!m.loaded) eg::raise(EG_ERR_INVALID, "no dispatch model is loaded");
        hipStream_t s = (hipStream_t)stream;
        std::vector<eg::FQuery> h((size_t)num);
        for (int64_t i = 0; i < num; ++i) {
            const eg_mlp_query& u = queries[i];
            if (!(u.key_type == EG_INT32 || u.key_type == EG_INT64 || u.key_type == EG_UINT32 || u.key_type == EG_UINT64))
                eg::raise(EG_ERR_UNSUPPORTED, "features need 4- or 8-byte integer keys");
            if (u.rows < 0 || (u.rows > 0 && !u.keys)) eg::raise(EG_ERR_INVALID, "bad key column");
            h[(size_t)i] = eg::FQuery{u.keys, u.rows, u.group_bound, eg::type_size(u.key_type), u.nops, u.value_bytes, 0};
        }
        eg::DevBuf dq(h.size() * sizeof(eg::FQuery), s), feat((size_t)num * eg::kIn * 4, s), out((size_t)num * eg::kOut * 4, s);
        EG_HIP(hipMemcpyAsync(dq.get(), h.data(), h.size() * sizeof(eg::FQuery), hipMemcpyHostToDevice, s));
        eg::mlp_features(dq.get<eg::FQuery>(), num, feat.get<float>(), s);
        hipLaunchKernelGGL(eg::mlp_fwd_kernel, dim3((unsigned)((num + eg::kS - 1) / eg::kS)), dim3(eg::kMB), 0, s,
                           (const float*)m.params.get<float>(), (const float*)feat.get<float>(), (int)num, out.get<float>());
        EG_LAUNCH_CHECK();
        std::vector<float> y((size_t)num * eg::kOut);
        EG_HIP(hipMemcpyAsync(y.data(), out.get(), y.size() * 4, hipMemcpyDeviceToHost, s));
        EG_HIP(hipStreamSynchronize(s));
        for (int64_t i = 0; i < num; ++i) {
            const float* o = y.data() + (size_t)i * eg::kOut;
            const bool by_bound = eg::bound_picks_hash(queries[i].rows, queries[i].group_bound);
            algos[i] = by_bound ? EG_ALGO_HASH : eg::algo_of(o, queries[i].rows);
            if (card_est) {
                const double lg = std::min(62.0, std::max(0.0, (double)o[3] * 32.0));
                card_est[i] = queries[i].rows < eg::kSmallRows  // as mlp_select: hash sized by the row count
                                  ? queries[i].rows
                              : by_bound ? queries[i].group_bound
                                  : (int64_t)std::min<double>((double)std::max<int64_t>(queries[i].rows, 1), std::exp2(lg));
            }
        }
    });
}

int eg_mlp_loaded(void)
{
    eg::LoadedModel& m = eg::loaded_model();
    std::lock_guard<std::mutex> lk(m.mu);
    return m.loaded ? 1 : 0;
}

}  // extern "C"
