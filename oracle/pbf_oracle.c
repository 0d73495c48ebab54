/*
 * pbf_oracle.c -- CPU restatement of the reference PBF substep.
 *
 * TEST INFRASTRUCTURE ONLY (see pbf_oracle.h).  Built with
 * -ffp-contract=off -fno-fast-math so every expression below is evaluated as
 * written in IEEE binary32 (the fused multiply-adds are explicit fmaf
 * calls).  Parallel over particles with OpenMP; each particle's sums are
 * sequential in the reference's friend order, so results do not depend on
 * the thread count.
 */
#include "pbf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define F4(a, i) ((a) + 4 * (size_t)(i))

int pbfo_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Simulation.cpp:167,179-180: the factors are double expressions streamed
 * with std::showpoint at the default precision (6 significant digits) and
 * suffixed with 'f', so the program sees the float nearest to the 6-digit
 * decimal.  std::pow(float, int) promotes to double. */
void pbfo_kernel_factors(float h, float *poly6, float *spiky)
{
    char buf[64];
    const double hd = (double)h;
    const double p6 = 315.0f / (64.0f * M_PI * pow(hd, 9));
    const double sp = 45.0f / (M_PI * pow(hd, 6));
    snprintf(buf, sizeof buf, "%#g", p6);
    *poly6 = strtof(buf, NULL);
    snprintf(buf, sizeof buf, "%#g", sp);
    *spiky = strtof(buf, NULL);
}

/* utilities.cl:69-77, evaluated on the 32-bit two's-complement pattern. */
uint32_t pbfo_expand_bits(int32_t xi)
{
    uint32_t x = (uint32_t)xi;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

/* utilities.cl:79-87: Morton number (always < 2^30, so non-negative as int)
 * modulo GRID_BUF_SIZE. */
uint32_t pbfo_cell_hash(int32_t cx, int32_t cy, int32_t cz, uint32_t gbs)
{
    const uint32_t m = pbfo_expand_bits(cx) | (pbfo_expand_bits(cy) << 1) | (pbfo_expand_bits(cz) << 2);
    return (uint32_t)((int32_t)m % (int32_t)gbs);
}

/* convert_int of x/h (radixsort.cl:12, build_friends_list.cl:43): truncation
 * toward zero.  Out-of-range and NaN inputs are undefined in the reference;
 * here NaN maps to 0 and magnitudes are clamped to 2^24. */
int32_t pbfo_cell_coord(float q)
{
    if (!(q == q)) return 0;
    if (q >= 16777216.0f) return 16777216;
    if (q <= -16777216.0f) return -16777216;
    return (int32_t)q;
}

/* Sine for the wall edge-offset noise (compute_delta.cl:165-166).  Written
 * from the mathematics so that the CPU and GPU evaluate identical binary32
 * operations: Cody-Waite reduction by pi (three-part split of pi), then the
 * odd Taylor polynomial of degree 11 on [-pi/2, pi/2].  Inputs with
 * |x| >= 65536 or NaN return 0 (outside any box the configs use). */
float pbfo_sinf(float x)
{
    if (!(fabsf(x) < 65536.0f)) return 0.0f;
    const float k = rintf(x * 0x1.45f306p-2f);
    float r = x - k * 0x1.92p+1f;
    r = r - k * 0x1.fb6p-11f;
    r = r - k * -0x1.777a5cp-24f;
    const float r2 = r * r;
    float p = -2.50521083854417e-8f;          /* -1/11! */
    p = p * r2 + 2.75573192239859e-6f;        /*  1/9!  */
    p = p * r2 + -1.98412698412698e-4f;       /* -1/7!  */
    p = p * r2 + 8.33333333333333e-3f;        /*  1/5!  */
    p = p * r2 + -1.66666666666667e-1f;       /* -1/3!  */
    p = r + (r * r2) * p;
    return (((int32_t)k) & 1) ? -p : p;
}

/* compute_scaling.cl:50 (threshold 0.6) and compute_delta.cl:100,
 * apply_viscosity.cl:28, apply_vorticity.cl:26 (threshold 0.5). */
int pbfo_skip_circle(int processed, int total, float threshold)
{
    return ((float)processed) / (float)total > threshold;
}

/* predict_positions.cl:1-18 */
void pbfo_predict(const pbfo_cfg *c, uint32_t pause, const float *x, float *v, float *xs)
{
    const float dt = c->dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)c->n; ++i) {
        float *vi = F4(v, i);
        const float *xi = F4(x, i);
        if (pause == 0) {
            vi[0] = vi[0] + dt * c->gravity[0];
            vi[1] = vi[1] + dt * c->gravity[1];
            vi[2] = vi[2] + dt * c->gravity[2];
        }
        float *o = F4(xs, i);
        for (int k = 0; k < 4; ++k) o[k] = xi[k] + dt * vi[k];
    }
}

/* radixsort.cl:1-20 (the padding keys beyond n are not materialised). */
void pbfo_keys(const pbfo_cfg *c, const float *xs, uint32_t *keys)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)c->n; ++i) {
        const float *p = F4(xs, i);
        keys[i] = pbfo_cell_hash(pbfo_cell_coord(p[0] / c->h), pbfo_cell_coord(p[1] / c->h),
                                 pbfo_cell_coord(p[2] / c->h), c->grid_buf_size);
    }
}

/* The reference's radix sort (radixsort.cl:61-308 driven by
 * Simulation.cpp:466-520) is a stable LSD counting sort on 30 bits, with
 * perm initialised to the identity (radixsort.cl:19).  Its output is the
 * unique stable order; restated here as an 8-bit LSD counting sort. */
void pbfo_stable_sort(const uint32_t *keys, uint32_t n, uint32_t *keys_sorted, uint32_t *perm)
{
    uint32_t *kt = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    uint32_t *pt = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
    /* four 8-bit passes; even passes read (keys_sorted, perm), odd passes
     * read (kt, pt), so the result lands back in (keys_sorted, perm) */
    memcpy(keys_sorted, keys, sizeof(uint32_t) * n);
    for (uint32_t i = 0; i < n; ++i) perm[i] = i;
    for (int pass = 0; pass < 4; ++pass) {
        const uint32_t *ki = (pass & 1) ? kt : keys_sorted, *pi = (pass & 1) ? pt : perm;
        uint32_t *ko = (pass & 1) ? keys_sorted : kt, *po = (pass & 1) ? perm : pt;
        size_t cnt[257];
        memset(cnt, 0, sizeof cnt);
        const int sh = 8 * pass;
        for (uint32_t i = 0; i < n; ++i) cnt[((ki[i] >> sh) & 255u) + 1]++;
        for (int d = 0; d < 256; ++d) cnt[d + 1] += cnt[d];
        for (uint32_t i = 0; i < n; ++i) {
            const size_t dst = cnt[(ki[i] >> sh) & 255u]++;
            ko[dst] = ki[i];
            po[dst] = pi[i];
        }
    }
    free(kt);
    free(pt);
}

/* radixsort.cl:22-37 (one float4 array). */
void pbfo_gather(const uint32_t *perm, uint32_t n, const float *src, float *dst)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) memcpy(F4(dst, i), F4(src, perm[i]), 16);
}

/* update_cells.cl:1-26 on an all-empty table.  Deviation: with n == 1 the
 * reference never writes the end index (its buildFriendsList then scans to
 * 0xFFFFFFFF); here the last particle always closes its cell. */
void pbfo_cells(const uint32_t *ks, uint32_t n, uint32_t *cells)
{
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t k = ks[i];
        if (i == 0) {
            cells[2 * (size_t)k] = 0;
        } else if (ks[i - 1] != k) {
            cells[2 * (size_t)k] = i;
            cells[2 * (size_t)ks[i - 1] + 1] = i - 1;
        }
        if (i == n - 1) cells[2 * (size_t)k + 1] = i;
    }
}

static inline float dot3(const float *a, const float *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* |r| from r^2 for the sweeps: a pair with r^2 < 2^-96 (|r| < 2^-48) is
 * coincident, with length and direction 0 (h - |r| rounds to h there); the
 * same rule gives |omega_j| from its square.  The kernels evaluate the
 * square root and reciprocal as a hardware estimate plus one correction,
 * equal to sqrtf and 1.0f/x on exactly this domain. */
static inline float pair_rl(float r2) { return r2 >= 0x1p-96f ? sqrtf(r2) : 0.0f; }

static inline void cross3(const float *a, const float *b, float *o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Fused multiply-adds at fixed points, the same ones the GPU kernels fuse:
 * the squared distance of a pair, fmaf(rz, rz, fmaf(ry, ry, rx * rx)), and
 * each product that a sweep adds into a sum.  fmaf rounds once, as the
 * hardware FMA does; every other expression here stays unfused.  The
 * reference compiles its kernels with -cl-mad-enable (src/Simulation.cpp:160),
 * which lets its compiler fuse any of these. */
static inline float dist2(const float *r) { return fmaf(r[2], r[2], fmaf(r[1], r[1], r[0] * r[0])); }
/* a x b with each component a fused difference of products */
static inline void cross3_fused(const float *a, const float *b, float *o)
{
    o[0] = fmaf(a[1], b[2], -(a[2] * b[1]));
    o[1] = fmaf(a[2], b[0], -(a[0] * b[2]));
    o[2] = fmaf(a[0], b[1], -(a[1] * b[0]));
}


/* build_friends_list.cl:23-103 */
void pbfo_friends(const pbfo_cfg *c, const float *xs, const uint32_t *cells, uint32_t *fr)
{
    const uint32_t n = c->n;
    const float h = c->h;
    const float h2 = h * h;                 /* ParamUtils.cpp:97 */
    const float min_r = 0.3f * h;
    const float span = h - min_r;
    uint32_t *idx = fr + (size_t)PBFO_CIRCLES * n;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float *pi = F4(xs, i);
        int cnt[PBFO_CIRCLES] = {0, 0, 0, 0, 0};
        const int32_t cx = pbfo_cell_coord(pi[0] / h), cy = pbfo_cell_coord(pi[1] / h),
                      cz = pbfo_cell_coord(pi[2] / h);
        for (int dx = -1; dx <= 1; ++dx)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    const uint32_t k = pbfo_cell_hash(cx + dx, cy + dy, cz + dz, c->grid_buf_size);
                    const uint32_t s = cells[2 * (size_t)k], e = cells[2 * (size_t)k + 1];
                    if (s == PBFO_EMPTY) continue;
                    for (uint32_t j = s; j <= e; ++j) {
                        if (j == i) continue;
                        const float *pj = F4(xs, j);
                        float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                        const float r2 = dist2(r);
                        /* build_friends_list.cl:70 skips r^2 >= h^2; a NaN
                         * r^2 (relaxed math there: undefined) is skipped too */
                        if (!(r2 < h2)) continue;
                        const float a = fmaxf(0.0f, (sqrtf(r2) - min_r) / span);
                        int circ = (int)(a * a * a * (float)PBFO_CIRCLES);
                        if (circ > PBFO_CIRCLES - 1) circ = PBFO_CIRCLES - 1;
                        if (cnt[circ] >= PBFO_PER_CIRCLE) continue;
                        const int f = cnt[circ]++;
                        idx[((size_t)circ * PBFO_PER_CIRCLE + (size_t)f) * n + i] = j;
                    }
                }
        for (int q = 0; q < PBFO_CIRCLES; ++q) fr[(size_t)q * n + i] = (uint32_t)cnt[q];
    }
}

#define FRIEND(fr, n, circ, f, i) ((fr)[(size_t)PBFO_CIRCLES * (n) + ((size_t)(circ) * PBFO_PER_CIRCLE + (f)) * (n) + (i)])

/* compute_scaling.cl:1-113 */
void pbfo_scaling(const pbfo_cfg *c, const float *xs, const uint32_t *fr, float *rho, float *lambda)
{
    const uint32_t n = c->n;
    const float h = c->h, h2 = h * h, e = c->epsilon, rd = c->rest_density;
    float poly6, spiky;
    pbfo_kernel_factors(h, &poly6, &spiky);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float *pi = F4(xs, i);
        float dens = 0.0f, gsum = 0.0f, gi[3] = {0.0f, 0.0f, 0.0f};
        int cnt[PBFO_CIRCLES], total = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) total += cnt[q] = (int)fr[(size_t)q * n + i];
        int done = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) {
            if (pbfo_skip_circle(done, total, 0.6f)) continue;
            done += cnt[q];
            for (int f = 0; f < cnt[q]; ++f) {
                const float *pj = F4(xs, FRIEND(fr, n, q, f, i));
                const float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                const float r2 = dist2(r);
                if (r2 < h2) {
                    const float rl = pair_rl(r2);
                    const float hr = h - rl;
                    const float s = spiky * hr * hr;
                    const float inv = rl > 0.0f ? 1.0f / rl : 0.0f;   /* r/|r| = r * (1/|r|); 0 when coincident */
                    const float t = s * inv;
                    const float g[3] = {r[0] * t, r[1] * t, r[2] * t};
                    const float d = h2 - r2;
                    dens = fmaf(d * d, d, dens);
                    gsum = fmaf(g[2], g[2], fmaf(g[1], g[1], fmaf(g[0], g[0], gsum)));
                    gi[0] += g[0]; gi[1] += g[1]; gi[2] += g[2];
                }
            }
        }
        dens *= poly6;
        rho[i] = dens;
        gsum += dot3(gi, gi);
        const float C = (dens / rd) - 1.0f;
        lambda[i] = -1.0f * C / (gsum / (rd * rd) + e);
    }
}

/* pack_data.cl:1-11 */
void pbfo_pack(uint32_t n, const float *src4, const float *w, float *dst4)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        float *o = F4(dst4, i);
        const float *s = F4(src4, i);
        o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; o[3] = w[i];
    }
}

/* Box walls.  The reference's BouncePointQuad (compute_delta.cl:1-42) with
 * an infinite axis-aligned plane (so the s,t range test always passes), mask
 * = 1 and the plane point reconstructed exactly: a point behind the offset
 * plane whose previous position was within the force zone is moved onto the
 * offset plane.  Planes are applied in the order x-lo, x-hi, y-lo, y-hi,
 * z-lo, z-hi, as the reference chains its quads (compute_delta.cl:169-182). */
static inline int wall_lo(float *f, const float *prev, int a, float lo, float edge)
{
    const float pa = lo + edge;
    const float dp = f[a] - pa;
    if (dp > 0.0f) return 0;                     /* dot(n, deltaP) > 0   */
    const float nv = prev[a] - f[a];             /* dot(n, prev - next)  */
    if (fabsf(dp) > edge + nv) return 0;         /* outside force zone   */
    f[a] = pa;
    return 1;
}

static inline int wall_hi(float *f, const float *prev, int a, float hi, float edge)
{
    const float pa = hi - edge;
    const float dp = f[a] - pa;
    if (dp < 0.0f) return 0;
    const float nv = f[a] - prev[a];
    if (fabsf(dp) > edge + nv) return 0;
    f[a] = pa;
    return 1;
}

/* A quad with the constants BouncePointQuad derives from it
 * (compute_delta.cl:3-9), computed once per call. */
typedef struct {
    float b[3], e0[3], e1[3], n[3];
    float aa, bb, cc, invdet;
    int32_t mask_y, mask_h;
} quadc;

static void quad_consts(const pbfo_quad *q, quadc *o)
{
    for (int k = 0; k < 3; ++k) { o->b[k] = q->b[k]; o->e0[k] = q->e0[k]; o->e1[k] = q->e1[k]; }
    o->aa = dot3(q->e0, q->e0);
    o->bb = dot3(q->e0, q->e1);
    o->cc = dot3(q->e1, q->e1);
    /* the reference's 1.0 is a double literal; restated in binary32 */
    o->invdet = 1.0f / (o->aa * o->cc - o->bb * o->bb);
    float cr[3];
    cross3(q->e0, q->e1, cr);
    /* normalize() as v * (1/|v|) */
    const float il = 1.0f / sqrtf(dot3(cr, cr));
    for (int k = 0; k < 3; ++k) o->n[k] = cr[k] * il;
    o->mask_y = q->mask_y;
    o->mask_h = q->mask_h;
}

/* BouncePointQuad (compute_delta.cl:1-42): a point behind the quad's
 * offset plane, inside the quad and within the force zone, where the mask is
 * set, moves onto the offset plane.  Texel (int)(t W), mask_y + (int)(s h),
 * clamped to the image (CLK_ADDRESS_CLAMP_TO_EDGE, utilities.cl:89). */
static int bounce_quad(const quadc *q, const float *prev, float *f, float edge, const uint32_t *mask,
                       uint32_t mw, uint32_t mrows)
{
    const float D[3] = {q->b[0] - f[0], q->b[1] - f[1], q->b[2] - f[2]};
    const float e = dot3(q->e1, D);
    const float d = dot3(q->e0, D);
    const float s = q->invdet * (q->bb * e - q->cc * d);
    const float t = q->invdet * (q->bb * d - q->aa * e);
    if ((s < 0.0f) || (s > 1.0f) || (t < 0.0f) || (t > 1.0f)) return 0;
    float pp[3], dp[3], pv[3];
    for (int k = 0; k < 3; ++k) {
        pp[k] = q->b[k] + q->e0[k] * s + q->e1[k] * t + q->n[k] * edge;
        dp[k] = f[k] - pp[k];
        pv[k] = prev[k] - f[k];
    }
    if (dot3(q->n, dp) > 0.0f) return 0;
    const float nv = dot3(q->n, pv);
    if (sqrtf(dot3(dp, dp)) > edge + nv) return 0;
    int32_t mx = (int32_t)(t * (float)mw);
    int32_t my = (int32_t)((float)q->mask_y + s * (float)q->mask_h);
    mx = mx < 0 ? 0 : (mx >= (int32_t)mw ? (int32_t)mw - 1 : mx);
    my = my < 0 ? 0 : (my >= (int32_t)mrows ? (int32_t)mrows - 1 : my);
    if (mask[(size_t)my * mw + (size_t)mx] == 0u) return 0;
    for (int k = 0; k < 3; ++k) f[k] = pp[k];
    return 1;
}

void pbfo_delta(const pbfo_cfg *c, const float *x, const float *xs, const uint32_t *fr, float *delta)
{
    pbfo_delta_hits(c, x, xs, fr, delta, NULL);
}

/* compute_delta.cl:44-186 (delta.w is left untouched, as in the reference).
 * hits (optional): per particle, which planes moved it (bit 2a: axis a low,
 * bit 2a+1: axis a high, bit 6+k: quad k). */
void pbfo_delta_hits(const pbfo_cfg *c, const float *x, const float *xs, const uint32_t *fr, float *delta,
                     uint32_t *hits)
{
    quadc qc[26];
    const uint32_t nq = c->quads && c->mask ? (c->quad_count < 26 ? c->quad_count : 26) : 0;
    for (uint32_t k = 0; k < nq; ++k) quad_consts(&c->quads[k], &qc[k]);
    const uint32_t n = c->n;
    const float h = c->h, h2 = h * h, rd = c->rest_density, K = c->st_k;
    float poly6, spiky;
    pbfo_kernel_factors(h, &poly6, &spiky);
    /* compute_delta.cl:81-82: pow(dq*h, 2) and pow(h^2 - q^2, 3) as products */
    const float dqh = c->st_dist * h;
    const float q2 = dqh * dqh;
    const float tq = h2 - q2;
    const float poly6_q = tq * tq * tq;
    /* poly6_r / poly6_q (compute_delta.cl:151) as poly6_r * (1/poly6_q): the
     * divisor is a constant of the step, so one correctly rounded reciprocal
     * serves every pair (relaxed math permits the same in the reference) */
    const float inv_q = 1.0f / poly6_q;
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float *pi = F4(xs, i);
        float sum[3] = {0.0f, 0.0f, 0.0f};
        int cnt[PBFO_CIRCLES], total = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) total += cnt[q] = (int)fr[(size_t)q * n + i];
        int done = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) {
            if (pbfo_skip_circle(done, total, 0.5f)) continue;
            done += cnt[q];
            for (int f = 0; f < cnt[q]; ++f) {
                const float *pj = F4(xs, FRIEND(fr, n, q, f, i));
                const float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                const float r2 = dist2(r);
                if (r2 < h2) {
                    const float rl = pair_rl(r2);
                    const float hr = h - rl;
                    const float inv = rl > 0.0f ? 1.0f / rl : 0.0f;   /* r/|r| = r * (1/|r|); 0 when coincident */
                    const float t = inv * hr * hr;
                    const float g[3] = {r[0] * t, r[1] * t, r[2] * t};
                    const float d = h2 - r2;
                    const float p6 = d * d * d;
                    const float ratio = p6 * inv_q;
                    /* lambda_i + lambda_j + s_corr, s_corr = K ratio^4 as (ratio^2)^2 */
                    const float r4 = ratio * ratio;
                    const float coef = fmaf(K, r4 * r4, pi[3] + pj[3]);
                    for (int k = 0; k < 3; ++k) sum[k] = fmaf(coef, g[k], sum[k]);
                }
            }
        }
        float fut[3];
        for (int k = 0; k < 3; ++k) fut[k] = pi[k] + (-spiky * sum[k]) / rd;
        uint32_t hb = 0;
        if (c->walls || nq) {
            float ns[3];
            for (int k = 0; k < 3; ++k) ns[k] = pbfo_sinf(fut[k] * 5.0f / h);
            /* compute_delta.cl:166; the walls leave out its scene-scale "1.0 +" */
            const float edge = (3.0f + ns[0] + ns[1] + ns[2]) * h * 0.03f;
            const float *prev = F4(x, i);
            if (c->walls)
                for (int a = 0; a < 3; ++a) {
                    hb |= (uint32_t)wall_lo(fut, prev, a, c->box_lo[a], edge) << (2 * a);
                    hb |= (uint32_t)wall_hi(fut, prev, a, c->box_hi[a], edge) << (2 * a + 1);
                }
            const float eq = c->edge_base + edge;
            for (uint32_t k = 0; k < nq; ++k)
                hb |= (uint32_t)bounce_quad(&qc[k], prev, fut, eq, c->mask, c->mask_w, c->mask_rows) << (6 + k);
        }
        if (hits) hits[i] = hb;
        float *o = F4(delta, i);
        for (int k = 0; k < 3; ++k) o[k] = fut[k] - pi[k];
    }
}

/* update_predicted.cl:1-11 (w: the reference adds an unwritten delta.w; the
 * oracle keeps xs.w, which the next packData overwrites either way). */
void pbfo_update_predicted(uint32_t n, const float *xs, const float *delta, float *xs_out)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; ++i) {
        const float *s = F4(xs, i), *d = F4(delta, i);
        float *o = F4(xs_out, i);
        o[0] = s[0] + d[0]; o[1] = s[1] + d[1]; o[2] = s[2] + d[2]; o[3] = s[3];
    }
}

/* update_velocities.cl:1-18 (fast_length taken as sqrt(dot)). */
void pbfo_update_velocities(const pbfo_cfg *c, float *x, const float *xs, float *v)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)c->n; ++i) {
        float *xi = F4(x, i), *vi = F4(v, i);
        const float *p = F4(xs, i);
        for (int k = 0; k < 3; ++k) vi[k] = (p[k] - xi[k]) / c->dt;
        xi[0] = p[0]; xi[1] = p[1]; xi[2] = p[2];
        xi[3] = sqrtf(dot3(vi, vi));
    }
}


/* apply_viscosity.cl:1-87 with Jacobi semantics: every v_j is read from the
 * pre-kernel snapshot v_in (the reference reads and writes one buffer, which
 * is schedule dependent). */
void pbfo_viscosity(const pbfo_cfg *c, const float *xs, const float *v_in, const uint32_t *fr,
                    float *v_out, float *omega)
{
    const uint32_t n = c->n;
    const float h = c->h, h2 = h * h;
    float poly6, spiky;
    pbfo_kernel_factors(h, &poly6, &spiky);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float *pi = F4(xs, i), *vi = F4(v_in, i);
        float vs[3] = {0.0f, 0.0f, 0.0f}, om[3] = {0.0f, 0.0f, 0.0f};
        int cnt[PBFO_CIRCLES], total = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) total += cnt[q] = (int)fr[(size_t)q * n + i];
        int done = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) {
            if (pbfo_skip_circle(done, total, 0.5f)) continue;
            done += cnt[q];
            for (int f = 0; f < cnt[q]; ++f) {
                const uint32_t j = FRIEND(fr, n, q, f, i);
                const float *pj = F4(xs, j), *vj = F4(v_in, j);
                const float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                const float r2 = dist2(r);
                if (r2 < h2 && fabsf(pj[3]) > 1e-8f) {
                    const float vij[3] = {vj[0] - vi[0], vj[1] - vi[1], vj[2] - vi[2]};
                    const float d = h2 - r2;
                    const float rl = pair_rl(r2);
                    const float hr = h - rl;
                    const float inv = rl > 0.0f ? 1.0f / rl : 0.0f;   /* r/|r| = r * (1/|r|); 0 when coincident */
                    const float t = inv * hr * hr;
                    const float g[3] = {r[0] * t, r[1] * t, r[2] * t};
                    float cr[3];
                    cross3_fused(vij, g, cr);
                    om[0] += cr[0]; om[1] += cr[1]; om[2] += cr[2];
                    const float p6 = d * d * d;
                    for (int k = 0; k < 3; ++k) vs[k] = fmaf(vij[k], p6, vs[k]);
                }
            }
        }
        float *vo = F4(v_out, i), *oo = F4(omega, i);
        for (int k = 0; k < 3; ++k) vo[k] = vi[k] + c->viscosity * (vs[k] * poly6);
        vo[3] = vi[3] + c->viscosity * 0.0f;
        for (int k = 0; k < 3; ++k) oo[k] = om[k] * -spiky;
        /* w = |omega| where the reference stores 0 (apply_viscosity.cl:86):
         * the GPU's vorticity sweep reads it instead of a length per pair;
         * the vorticity step below still takes |omega_j| from x, y, z */
        oo[3] = pair_rl(dot3(oo, oo));
    }
}

/* apply_vorticity.cl:1-87 */
void pbfo_vorticity(const pbfo_cfg *c, const float *xs, float *v, const float *omega, const uint32_t *fr)
{
    const uint32_t n = c->n;
    const float h = c->h, h2 = h * h;
    float poly6, spiky;
    pbfo_kernel_factors(h, &poly6, &spiky);
#pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ++ii) {
        const uint32_t i = (uint32_t)ii;
        const float *pi = F4(xs, i);
        float eta[3] = {0.0f, 0.0f, 0.0f};
        int cnt[PBFO_CIRCLES], total = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) total += cnt[q] = (int)fr[(size_t)q * n + i];
        int done = 0;
        for (int q = 0; q < PBFO_CIRCLES; ++q) {
            if (pbfo_skip_circle(done, total, 0.5f)) continue;
            done += cnt[q];
            for (int f = 0; f < cnt[q]; ++f) {
                const uint32_t j = FRIEND(fr, n, q, f, i);
                const float *pj = F4(xs, j);
                const float r[3] = {pi[0] - pj[0], pi[1] - pj[1], pi[2] - pj[2]};
                const float r2 = dist2(r);
                if (r2 < h2 && fabsf(pj[3]) > 1e-8f) {
                    const float rl = pair_rl(r2);
                    const float hr = h - rl;
                    const float inv = rl > 0.0f ? 1.0f / rl : 0.0f;   /* r/|r| = r * (1/|r|); 0 when coincident */
                    const float t = inv * hr * hr;
                    const float g[3] = {r[0] * t, r[1] * t, r[2] * t};
                    const float *oj = F4(omega, j);
                    const float wl = pair_rl(dot3(oj, oj));
                    for (int k = 0; k < 3; ++k) eta[k] = fmaf(wl, g[k], eta[k]);
                }
            }
        }
        for (int k = 0; k < 3; ++k) eta[k] *= -spiky;
        const float l = sqrtf(dot3(eta, eta));
        if (l > 0.0f) { eta[0] /= l; eta[1] /= l; eta[2] /= l; }
        float cr[3];
        cross3(eta, F4(omega, i), cr);
        float *vi = F4(v, i);
        for (int k = 0; k < 3; ++k) vi[k] += (c->vorticity * cr[k]) * c->dt;
    }
}

/* ------------------------------------------------------------------------ */

struct pbfo_sim {
    pbfo_cfg c;
    float *x, *v, *xs, *xs2, *rho, *lambda, *omega, *delta, *tmp4;
    uint32_t *keys, *keys_sorted, *perm, *cells, *friends;
    uint32_t *wall_hits;  /* iters x n, written by each step when set */
    int sorted_once;
};

static void *xmalloc(size_t b)
{
    void *p = calloc(1, b ? b : 16);
    if (!p) { fprintf(stderr, "pbf_oracle: out of host memory (%zu B)\n", b); abort(); }
    return p;
}

pbfo_sim *pbfo_create(const pbfo_cfg *c)
{
    pbfo_sim *s = (pbfo_sim *)xmalloc(sizeof *s);
    s->c = *c;
    const size_t n = c->n;
    s->x = xmalloc(16 * n); s->v = xmalloc(16 * n); s->xs = xmalloc(16 * n); s->xs2 = xmalloc(16 * n);
    s->rho = xmalloc(4 * n); s->lambda = xmalloc(4 * n); s->omega = xmalloc(16 * n);
    s->delta = xmalloc(16 * n); s->tmp4 = xmalloc(16 * n);
    s->keys = xmalloc(4 * n); s->keys_sorted = xmalloc(4 * n); s->perm = xmalloc(4 * n);
    s->cells = xmalloc(8 * (size_t)c->grid_buf_size);
    memset(s->cells, 0xFF, 8 * (size_t)c->grid_buf_size);
    s->friends = xmalloc(4 * n * PBFO_CIRCLES * (1 + PBFO_PER_CIRCLE));
    return s;
}

void pbfo_destroy(pbfo_sim *s)
{
    if (!s) return;
    free(s->x); free(s->v); free(s->xs); free(s->xs2); free(s->rho); free(s->lambda);
    free(s->omega); free(s->delta); free(s->tmp4); free(s->keys); free(s->keys_sorted);
    free(s->perm); free(s->cells); free(s->friends); free(s);
}

void pbfo_set_state(pbfo_sim *s, const float *x, const float *v)
{
    memcpy(s->x, x, 16 * (size_t)s->c.n);
    memcpy(s->v, v, 16 * (size_t)s->c.n);
    s->sorted_once = 0;
}

/* Simulation::Step, Simulation.cpp:554-621. */
void pbfo_step(pbfo_sim *s, uint32_t pause)
{
    const pbfo_cfg *c = &s->c;
    const uint32_t n = c->n;
    pbfo_predict(c, pause, s->x, s->v, s->xs);
    /* radixsort() is skipped while paused (Simulation.cpp:566-567); before
     * the first sort there are no sorted keys, so the first step always sorts. */
    if (!pause || !s->sorted_once) {
        pbfo_keys(c, s->xs, s->keys);
        pbfo_stable_sort(s->keys, n, s->keys_sorted, s->perm);
        pbfo_gather(s->perm, n, s->x, s->tmp4);
        memcpy(s->x, s->tmp4, 16 * (size_t)n);
        pbfo_gather(s->perm, n, s->xs, s->tmp4);
        memcpy(s->xs, s->tmp4, 16 * (size_t)n);
        s->sorted_once = 1;
    }
    /* resetGrid (reset_grid.cl) leaves every cell empty before updateCells */
    memset(s->cells, 0xFF, 8 * (size_t)c->grid_buf_size);
    pbfo_cells(s->keys_sorted, n, s->cells);
    pbfo_friends(c, s->xs, s->cells, s->friends);
    for (uint32_t it = 0; it < c->iters; ++it) {
        pbfo_scaling(c, s->xs, s->friends, s->rho, s->lambda);
        pbfo_pack(n, s->xs, s->lambda, s->xs2);
        pbfo_delta_hits(c, s->x, s->xs2, s->friends, s->delta,
                        s->wall_hits ? s->wall_hits + (size_t)it * n : NULL);
        pbfo_update_predicted(n, s->xs2, s->delta, s->xs);
    }
    pbfo_pack(n, s->xs, s->rho, s->xs2);
    memcpy(s->xs, s->xs2, 16 * (size_t)n);
    pbfo_update_velocities(c, s->x, s->xs, s->v);
    pbfo_viscosity(c, s->xs, s->v, s->friends, s->tmp4, s->omega);
    memcpy(s->v, s->tmp4, 16 * (size_t)n);
    pbfo_vorticity(c, s->xs, s->v, s->omega, s->friends);
}

void pbfo_record_walls(pbfo_sim *s, uint32_t *hits) { s->wall_hits = hits; }

void *pbfo_buffer(pbfo_sim *s, int which)
{
    switch (which) {
    case 0: return s->x;
    case 1: return s->v;
    case 2: return s->xs;
    case 3: return s->rho;
    case 4: return s->lambda;
    case 5: return s->omega;
    case 6: return s->keys_sorted;
    case 7: return s->perm;
    case 8: return s->cells;
    case 9: return s->friends;
    case 10: return s->keys;
    default: return NULL;
    }
}
