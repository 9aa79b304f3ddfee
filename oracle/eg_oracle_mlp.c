/*
 * eg_oracle_mlp.c -- CPU restatement (fp64) of the learned dispatcher's
 * sample features and of the 16 -> 64 -> 64 -> 4 MLP: forward pass, loss and
 * its analytic gradient.  TEST INFRASTRUCTURE ONLY (see eg_oracle.h).
 *
 * The reference has no learned dispatch (its dispatcher is the rule of
 * operators/groupby/groupby.cu:68-119); these functions follow the formulas
 * stated in the library's header for EG_MLP_* and are pinned in the tests by
 * finite differences (gradient) and by numpy (forward pass, features).
 */
#include "eg_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IN 16
#define H 64
#define OUT 4
#define OW1 0
#define OB1 (OW1 + H * IN)
#define OW2 (OB1 + H)
#define OB2 (OW2 + H * H)
#define OW3 (OB2 + H)
#define OB3 (OW3 + OUT * H)
#define NPARAM (OB3 + OUT)

static void forward1(const double* P, const float* x, double* h1, double* h2, double* o)
{
    for (int u = 0; u < H; ++u) {
        double z = P[OB1 + u];
        for (int f = 0; f < IN; ++f) z += P[OW1 + u * IN + f] * (double)x[f];
        h1[u] = z > 0 ? z : 0;
    }
    for (int u = 0; u < H; ++u) {
        double z = P[OB2 + u];
        for (int v = 0; v < H; ++v) z += P[OW2 + u * H + v] * h1[v];
        h2[u] = z > 0 ? z : 0;
    }
    for (int k = 0; k < OUT; ++k) {
        double z = P[OB3 + k];
        for (int v = 0; v < H; ++v) z += P[OW3 + k * H + v] * h2[v];
        o[k] = z;
    }
}

int orc_mlp_forward(const double* params, const float* x, int64_t n, double* y)
{
    double h1[H], h2[H];
    for (int64_t s = 0; s < n; ++s) forward1(params, x + s * IN, h1, h2, y + s * OUT);
    return 0;
}

/* Mean over samples of CE(softmax(o[0..2]), target) + lambda (o[3] - card)^2,
 * and its gradient w.r.t. every parameter. */
int orc_mlp_loss_grad(const double* P, const float* x, const float* target, const float* card, int64_t n,
                      double lambda, double* grad, double* loss)
{
    double h1[H], h2[H], o[OUT], dz[OUT], dh2[H], dh1[H];
    memset(grad, 0, sizeof(double) * NPARAM);
    double L = 0;
    const double inv = n > 0 ? 1.0 / (double)n : 0.0;
    for (int64_t s = 0; s < n; ++s) {
        const float* xs = x + s * IN;
        forward1(P, xs, h1, h2, o);
        double m = o[0];
        if (o[1] > m) m = o[1];
        if (o[2] > m) m = o[2];
        const double e0 = exp(o[0] - m), e1 = exp(o[1] - m), e2 = exp(o[2] - m);
        const double se = e0 + e1 + e2, lse = m + log(se);
        const double p0 = target[s * 3], p1 = target[s * 3 + 1], p2 = target[s * 3 + 2];
        const double d = o[3] - (double)card[s];
        L += -(p0 * (o[0] - lse) + p1 * (o[1] - lse) + p2 * (o[2] - lse)) + lambda * d * d;
        dz[0] = (e0 / se - p0) * inv;
        dz[1] = (e1 / se - p1) * inv;
        dz[2] = (e2 / se - p2) * inv;
        dz[3] = 2.0 * lambda * d * inv;
        for (int k = 0; k < OUT; ++k) {
            grad[OB3 + k] += dz[k];
            for (int v = 0; v < H; ++v) grad[OW3 + k * H + v] += dz[k] * h2[v];
        }
        for (int u = 0; u < H; ++u) {
            double g = 0;
            for (int k = 0; k < OUT; ++k) g += dz[k] * P[OW3 + k * H + u];
            dh2[u] = h2[u] > 0 ? g : 0;
        }
        for (int u = 0; u < H; ++u) {
            grad[OB2 + u] += dh2[u];
            for (int v = 0; v < H; ++v) grad[OW2 + u * H + v] += dh2[u] * h1[v];
        }
        for (int v = 0; v < H; ++v) {
            double g = 0;
            for (int u = 0; u < H; ++u) g += dh2[u] * P[OW2 + u * H + v];
            dh1[v] = h1[v] > 0 ? g : 0;
        }
        for (int v = 0; v < H; ++v) {
            grad[OB1 + v] += dh1[v];
            for (int f = 0; f < IN; ++f) grad[OW1 + v * IN + f] += dh1[v] * (double)xs[f];
        }
    }
    if (loss) *loss = L * inv;
    return 0;
}

/* ---- sample features -------------------------------------------------------
 * Sample: rows r = i * stride, i < ns = min(rows, 4096), stride = 1 when
 * rows <= 4096 else rows / 4096.  A 4-byte key is read as a sign-extended
 * int32 (for the span and the order test), an 8-byte key as int64. */
typedef struct { int64_t k; int64_t c; } kc_t;

static int cmp_i64(const void* a, const void* b)
{
    int64_t x = *(const int64_t*)a, y = *(const int64_t*)b;
    return (x > y) - (x < y);
}

static int cmp_c_desc(const void* a, const void* b)
{
    int64_t x = ((const kc_t*)a)->c, y = ((const kc_t*)b)->c;
    return (x < y) - (x > y);
}

int orc_sample_features(const void* keys, int32_t key_type, int64_t rows, int64_t group_bound, int32_t nops,
                        int32_t value_bytes, float* out)
{
    const int kb = (key_type == ORC_INT32 || key_type == ORC_UINT32) ? 4
                 : (key_type == ORC_INT64 || key_type == ORC_UINT64) ? 8 : 0;
    if (!kb) return -1;
    const int64_t ns = rows < 4096 ? rows : 4096;
    const int64_t stride = rows <= 4096 ? 1 : rows / 4096;
    int64_t* k = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns > 0 ? ns : 1));
    int64_t* srt = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ns > 0 ? ns : 1));
    kc_t* g = (kc_t*)malloc(sizeof(kc_t) * (size_t)(ns > 0 ? ns : 1));
    if (!k || !srt || !g) { free(k); free(srt); free(g); return -1; }
    for (int64_t i = 0; i < ns; ++i) {
        const int64_t r = i * stride;
        k[i] = kb == 4 ? (int64_t)((const int32_t*)keys)[r] : ((const int64_t*)keys)[r];
    }
    int64_t asc = 0;
    for (int64_t i = 0; i + 1 < ns; ++i) asc += k[i + 1] >= k[i];
    memcpy(srt, k, sizeof(int64_t) * (size_t)ns);
    qsort(srt, (size_t)ns, sizeof(int64_t), cmp_i64);
    int64_t nd = 0;
    for (int64_t i = 0; i < ns; ++i) {
        if (i == 0 || srt[i] != srt[i - 1]) { g[nd].k = srt[i]; g[nd].c = 0; ++nd; }
        g[nd - 1].c++;
    }
    double f1 = 0, f2 = 0, ent = 0, mx = 0;
    for (int64_t j = 0; j < nd; ++j) {
        const double c = (double)g[j].c, p = c / (double)ns;
        f1 += g[j].c == 1;
        f2 += g[j].c == 2;
        ent -= p * log2(p);
        if (c > mx) mx = c;
    }
    qsort(g, (size_t)nd, sizeof(kc_t), cmp_c_desc);
    double top8 = 0;
    for (int64_t j = 0; j < nd && j < 8; ++j) top8 += (double)g[j].c;
    const double ndv = (double)nd, R = (double)rows, nsf = (double)ns;
    const double chao = f2 > 0 ? ndv + f1 * f1 / (2 * f2) : ndv + f1 * (f1 - 1) * 0.5;
    const double gee = (ns > 0 ? sqrt(R / nsf) : 0) * f1 + (ndv - f1);
    const double span = ns > 0 ? (double)((uint64_t)srt[ns - 1] - (uint64_t)srt[0]) + 1.0 : 0.0;
    const int64_t bnd = (group_bound > 0 && group_bound < rows) ? group_bound : rows;
    const double inv = ns > 0 ? 1.0 / nsf : 0.0;
    out[0] = (float)(log2(R + 1) / 32);
    out[1] = (float)kb / 8.f;
    out[2] = (float)(log2(ndv + 1) / 12);
    out[3] = (float)(log2((chao < R ? chao : R) + 1) / 32);
    out[4] = (float)(log2((gee < R ? gee : R) + 1) / 32);
    out[5] = (float)(top8 * inv);
    out[6] = ns > 1 ? (float)(ent / log2(nsf)) : 0.f;
    out[7] = (float)(f1 * inv);
    out[8] = (float)(f2 * inv);
    out[9] = (float)(mx * inv);
    out[10] = (float)(log2(span + 1) / 64);
    out[11] = (float)(log2((double)bnd + 1) / 32);
    out[12] = (float)nops / 16.f;
    out[13] = (float)value_bytes / 32.f;
    out[14] = ns > 1 ? (float)((double)asc / (nsf - 1)) : 1.f;
    out[15] = (float)(nsf / 4096.0);
    free(k);
    free(srt);
    free(g);
    return 0;
}
