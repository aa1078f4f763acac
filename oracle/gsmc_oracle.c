/*
 * CPU oracle for the Gamma-SMC cached forward-backward decoder.
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker: it is compiled into
 * liboracle.so and loaded by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, never by the product library or its Python
 * package.
 *
 * It restates, in plain C, the reference's recursion (A) -- the one
 * gamma_smc_cu.infer()/infer_blockwise() run -- in two precisions:
 *
 *  - FP32, operation for operation as the reference computes it:
 *      host table builder  src/flow_field.cpp:60-388
 *      device walk         src/kernels/gamma_smc_flow.cu:115-131,160-163,
 *                          408-514,567-609,632-669,678-839
 *    Explicit fmaf() appears exactly where the reference calls __fmaf_rn,
 *    and in the fast 10^x where nvcc's default contraction fuses
 *    `C2 * (x * LN10) + C1`; everything else is compiled with
 *    -ffp-contract=off so no other fusion happens.
 *  - FP64: the same steps in double (tables, walk, combine), used to report
 *    the FP32 path's error.
 *
 * Parity pins (tests/test_oracle_*.py): the FP32 tables are compared bit for
 * bit with the reference's own flow_field.cpp compiled into oracle/_ref; the
 * walk is pinned by an independent scipy interpolation of short walks, the
 * posterior identities (alpha/beta * 2Ne == mean, lower <= mean <= upper),
 * block/flank invariants and the FP64 restatement.
 */
#define _GNU_SOURCE /* exp10f */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MEAN_N 51
#define CV_N 50
#define GRID (MEAN_N * CV_N)

/* Table order inside the contiguous buffers exchanged with Python. */
enum { T_CACHE_M, T_CACHE_C, T_FHOM_M, T_FHOM_C, T_FHET_M, T_FHET_C,
       T_BHOM_M, T_BHOM_C, T_BHET_M, T_BHET_C, T_COUNT };

/* ------------------------------------------------------------------ */
/* Flow field file (src/flow_field.cpp:8-54)                            */
/* ------------------------------------------------------------------ */
int orc_load_flow_field(const char* path, float* u, float* v, float* bounds)
{
    FILE* f = fopen(path, "r");
    if (!f) return -1;
    float mlo, mhi, clo, chi;
    int mn, cn;
    if (fscanf(f, "%f %f %d", &mlo, &mhi, &mn) != 3 || mn != MEAN_N) { fclose(f); return -2; }
    if (fscanf(f, "%f %f %d", &clo, &chi, &cn) != 3 || cn != CV_N) { fclose(f); return -3; }
    bounds[0] = log10f(mlo);
    bounds[1] = log10f(mhi);
    bounds[2] = log10f(clo);
    bounds[3] = log10f(chi);
    for (int i = 0; i < GRID; i++)
        if (fscanf(f, "%f", &u[i]) != 1) { fclose(f); return -4; }
    for (int i = 0; i < GRID; i++)
        if (fscanf(f, "%f", &v[i]) != 1) { fclose(f); return -5; }
    fclose(f);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Entropy clip (src/flow_field.cpp:80-166)                             */
/* ------------------------------------------------------------------ */
#define ENT_MMAX 4.0
#define ENT_STEP 0.0008
#define ENT_N 5001

static double digamma_asym(double x)
{
    double r = 0.0;
    while (x < 6.0) { r -= 1.0 / x; x += 1.0; }
    double i1 = 1.0 / x, i2 = i1 * i1, i4 = i2 * i2, i6 = i4 * i2, i8 = i4 * i4, i10 = i8 * i2;
    r += log(x) - 0.5 * i1 - i2 / 12.0 + i4 / 120.0 - i6 / 252.0 + i8 / 240.0 - 5.0 * i10 / 660.0;
    return r;
}

static double gamma_entropy(double m, double c)
{
    double la = -2.0 * c, lb = la - m;
    double a = pow(10.0, la), b = pow(10.0, lb);
    return a - log(b) + lgamma(a) + (1.0 - a) * digamma_asym(a);
}

static double g_ent_d[ENT_N];
static float g_ent_f[ENT_N];
static int g_ent_ready = 0;

static void entropy_table(void)
{
    if (g_ent_ready) return;
    for (int i = 0; i < ENT_N; i++) {
        double m = i * ENT_STEP, lo = -2.0, hi = 0.0, r;
        if (gamma_entropy(m, lo) > 1.0) r = lo;
        else if (gamma_entropy(m, hi) <= 1.0) r = hi;
        else {
            for (int it = 0; it < 60; it++) {
                double mid = 0.5 * (lo + hi);
                if (gamma_entropy(m, mid) > 1.0) hi = mid; else lo = mid;
            }
            r = hi;
        }
        g_ent_d[i] = r;
        g_ent_f[i] = (float)r;
    }
    g_ent_ready = 1;
}

void orc_entropy_table(float* out_f, double* out_d)
{
    entropy_table();
    if (out_f) memcpy(out_f, g_ent_f, sizeof g_ent_f);
    if (out_d) memcpy(out_d, g_ent_d, sizeof g_ent_d);
}

/* std::min / std::max as the reference's host code uses them. */
static inline float mnf(float a, float b) { return (b < a) ? b : a; }
static inline float mxf(float a, float b) { return (a < b) ? b : a; }
static inline double mnd(double a, double b) { return (b < a) ? b : a; }
static inline double mxd(double a, double b) { return (a < b) ? b : a; }

typedef struct { float mlo, mhi, clo, chi, mstep, cstep; } box_f;
typedef struct { double mlo, mhi, clo, chi, mstep, cstep; } box_d;

static void clip_f(float* m, float* c, const box_f* b)
{
    *m = mxf(b->mlo, mnf(*m, b->mhi));
    *c = mxf(b->clo, mnf(*c, b->chi));
    if (*m >= 0.0f && *m <= ENT_MMAX) {
        float idx = *m / (float)ENT_STEP;
        int i0 = (int)idx;
        if (i0 >= ENT_N - 1) { *c = mnf(*c, g_ent_f[ENT_N - 1]); return; }
        float w = idx - i0;
        float lim = (1.0f - w) * g_ent_f[i0] + w * g_ent_f[i0 + 1];
        *c = mnf(*c, lim);
        *c = mxf(b->clo, mnf(*c, b->chi));
    }
}

static void clip_d(double* m, double* c, const box_d* b)
{
    *m = mxd(b->mlo, mnd(*m, b->mhi));
    *c = mxd(b->clo, mnd(*c, b->chi));
    if (*m >= 0.0 && *m <= ENT_MMAX) {
        double idx = *m / ENT_STEP;
        int i0 = (int)idx;
        if (i0 >= ENT_N - 1) { *c = mnd(*c, g_ent_d[ENT_N - 1]); return; }
        double w = idx - i0;
        double lim = (1.0 - w) * g_ent_d[i0] + w * g_ent_d[i0 + 1];
        *c = mnd(*c, lim);
        *c = mxd(b->clo, mnd(*c, b->chi));
    }
}

/* Host bilinear (src/flow_field.cpp:60-78): division by the step. */
static float bil_host_f(const float* t, float m, float c, const box_f* b)
{
    float fm = (m - b->mlo) / b->mstep;
    float fc = (c - b->clo) / b->cstep;
    fm = mxf(0.0f, mnf(fm, (float)(MEAN_N - 1)));
    fc = mxf(0.0f, mnf(fc, (float)(CV_N - 1)));
    int m0 = (int)fm, c0 = (int)fc;
    if (m0 == MEAN_N - 1) m0--;
    if (c0 == CV_N - 1) c0--;
    float wm = fm - m0, wc = fc - c0;
    int k = m0 * CV_N + c0;
    float v00 = t[k], v01 = t[k + 1], v10 = t[k + CV_N], v11 = t[k + CV_N + 1];
    return (1 - wm) * ((1 - wc) * v00 + wc * v01) + wm * ((1 - wc) * v10 + wc * v11);
}

static double bil_host_d(const double* t, double m, double c, const box_d* b)
{
    double fm = (m - b->mlo) / b->mstep;
    double fc = (c - b->clo) / b->cstep;
    fm = mxd(0.0, mnd(fm, (double)(MEAN_N - 1)));
    fc = mxd(0.0, mnd(fc, (double)(CV_N - 1)));
    int m0 = (int)fm, c0 = (int)fc;
    if (m0 == MEAN_N - 1) m0--;
    if (c0 == CV_N - 1) c0--;
    double wm = fm - m0, wc = fc - c0;
    int k = m0 * CV_N + c0;
    double v00 = t[k], v01 = t[k + 1], v10 = t[k + CV_N], v11 = t[k + CV_N + 1];
    return (1 - wm) * ((1 - wc) * v00 + wc * v01) + wm * ((1 - wc) * v10 + wc * v11);
}

/* One recombination step R (src/flow_field.cpp:241-246). */
static void recomb_f(float* m, float* c, const float* u, const float* v, float rho, const box_f* b)
{
    clip_f(m, c, b);
    float du = bil_host_f(u, *m, *c, b);
    float dv = bil_host_f(v, *m, *c, b);
    *m += du * rho;
    *c += dv * rho;
    clip_f(m, c, b);
}

static void recomb_d(double* m, double* c, const double* u, const double* v, double rho, const box_d* b)
{
    clip_d(m, c, b);
    double du = bil_host_d(u, *m, *c, b);
    double dv = bil_host_d(v, *m, *c, b);
    *m += du * rho;
    *c += dv * rho;
    clip_d(m, c, b);
}

/* Observed-site emission (src/flow_field.cpp:168-184). */
static void emit_f(int het, float mu, float* m, float* c)
{
    float la = -2.0f * *c;
    float lb = la - *m;
    lb = log10f(powf(10.0f, lb) + mu);
    if (het) la = log10f(powf(10.0f, la) + 1.0f);
    *m = la - lb;
    *c = -0.5f * la;
}

static void emit_d(int het, double mu, double* m, double* c)
{
    double la = -2.0 * *c;
    double lb = la - *m;
    lb = log10(pow(10.0, lb) + mu);
    if (het) la = log10(pow(10.0, la) + 1.0);
    *m = la - lb;
    *c = -0.5 * la;
}

/* ------------------------------------------------------------------ */
/* Table build (src/flow_field.cpp:205-388), FP32.                      */
/* out: T_COUNT arrays of n_steps*GRID floats, contiguous, enum order.  */
/* ------------------------------------------------------------------ */
int orc_build_tables_f32(const float* u, const float* v, const float* bounds,
                         int n_steps, float rho, float mu, float* out)
{
    if (n_steps < 1) return -1;
    entropy_table();
    box_f b;
    b.mlo = bounds[0]; b.mhi = bounds[1]; b.clo = bounds[2]; b.chi = bounds[3];
    b.mstep = (b.mhi - b.mlo) / (MEAN_N - 1);
    b.cstep = (b.chi - b.clo) / (CV_N - 1);
    size_t L = (size_t)n_steps * GRID;
    float* T[T_COUNT];
    for (int k = 0; k < T_COUNT; k++) T[k] = out + (size_t)k * L;

    #pragma omp parallel for schedule(static)
    for (int g = 0; g < GRID; g++) {
        int row = g / CV_N, col = g % CV_N;
        float m = b.mlo + row * b.mstep;
        float c = b.clo + col * b.cstep;
        recomb_f(&m, &c, u, v, rho, &b);
        T[T_CACHE_M][g] = m;
        T[T_CACHE_C][g] = c;
        for (int k = 1; k < n_steps; k++) {
            /* hom stretch: beta += mu, then R (flow_field.cpp:299-319) */
            float la = -2.0f * c;
            float lb = la - m;
            float bl = powf(10.0f, lb) + mu;
            lb = log10f(bl);
            m = la - lb;
            recomb_f(&m, &c, u, v, rho, &b);
            T[T_CACHE_M][(size_t)k * GRID + g] = m;
            T[T_CACHE_C][(size_t)k * GRID + g] = c;
        }
    }

    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n_steps; k++) {
        size_t off = (size_t)k * GRID;
        for (int g = 0; g < GRID; g++) {
            int row = g / CV_N, col = g % CV_N;
            for (int het = 0; het < 2; het++) {
                /* forward: stretch of k+1 bp then the observed site */
                float m = T[T_CACHE_M][off + g], c = T[T_CACHE_C][off + g];
                emit_f(het, mu, &m, &c);
                clip_f(&m, &c, &b);
                T[het ? T_FHET_M : T_FHOM_M][off + g] = m;
                T[het ? T_FHET_C : T_FHOM_C][off + g] = c;
                /* backward: observed site then a stretch of k+1 bp */
                float mb = b.mlo + row * b.mstep, cb = b.clo + col * b.cstep;
                emit_f(het, mu, &mb, &cb);
                clip_f(&mb, &cb, &b);
                float om = bil_host_f(T[T_CACHE_M] + off, mb, cb, &b);
                float oc = bil_host_f(T[T_CACHE_C] + off, mb, cb, &b);
                T[het ? T_BHET_M : T_BHOM_M][off + g] = om;
                T[het ? T_BHET_C : T_BHOM_C][off + g] = oc;
            }
        }
    }
    return 0;
}

/* FP64 tables: the same construction in double from the FP32-parsed
 * field; the entropy threshold stays double. */
int orc_build_tables_f64(const float* u32, const float* v32, const float* bounds,
                         int n_steps, double rho, double mu, double* out)
{
    if (n_steps < 1) return -1;
    entropy_table();
    double u[GRID], v[GRID];
    for (int i = 0; i < GRID; i++) { u[i] = u32[i]; v[i] = v32[i]; }
    box_d b;
    b.mlo = bounds[0]; b.mhi = bounds[1]; b.clo = bounds[2]; b.chi = bounds[3];
    b.mstep = (b.mhi - b.mlo) / (MEAN_N - 1);
    b.cstep = (b.chi - b.clo) / (CV_N - 1);
    size_t L = (size_t)n_steps * GRID;
    double* T[T_COUNT];
    for (int k = 0; k < T_COUNT; k++) T[k] = out + (size_t)k * L;

    #pragma omp parallel for schedule(static)
    for (int g = 0; g < GRID; g++) {
        int row = g / CV_N, col = g % CV_N;
        double m = b.mlo + row * b.mstep;
        double c = b.clo + col * b.cstep;
        recomb_d(&m, &c, u, v, rho, &b);
        T[T_CACHE_M][g] = m;
        T[T_CACHE_C][g] = c;
        for (int k = 1; k < n_steps; k++) {
            double la = -2.0 * c;
            double lb = la - m;
            lb = log10(pow(10.0, lb) + mu);
            m = la - lb;
            recomb_d(&m, &c, u, v, rho, &b);
            T[T_CACHE_M][(size_t)k * GRID + g] = m;
            T[T_CACHE_C][(size_t)k * GRID + g] = c;
        }
    }

    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n_steps; k++) {
        size_t off = (size_t)k * GRID;
        for (int g = 0; g < GRID; g++) {
            int row = g / CV_N, col = g % CV_N;
            for (int het = 0; het < 2; het++) {
                double m = T[T_CACHE_M][off + g], c = T[T_CACHE_C][off + g];
                emit_d(het, mu, &m, &c);
                clip_d(&m, &c, &b);
                T[het ? T_FHET_M : T_FHOM_M][off + g] = m;
                T[het ? T_FHET_C : T_FHOM_C][off + g] = c;
                double mb = b.mlo + row * b.mstep, cb = b.clo + col * b.cstep;
                emit_d(het, mu, &mb, &cb);
                clip_d(&mb, &cb, &b);
                double om = bil_host_d(T[T_CACHE_M] + off, mb, cb, &b);
                double oc = bil_host_d(T[T_CACHE_C] + off, mb, cb, &b);
                T[het ? T_BHET_M : T_BHOM_M][off + g] = om;
                T[het ? T_BHET_C : T_BHOM_C][off + g] = oc;
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Device walk, FP32 (src/kernels/gamma_smc_flow.cu)                    */
/* ------------------------------------------------------------------ */
#define DEV_MLO (-5.0f)
#define DEV_CLO (-2.0f)
static const float DEV_MSTEP = (2.0f - -5.0f) / (MEAN_N - 1);   /* gamma_smc_flow.cu:14 */
static const float DEV_CSTEP = (0.0f - -2.0f) / (CV_N - 1);     /* :15 */
static float dev_minv, dev_cinv;

static void dev_consts(void)
{
    dev_minv = 1.0f / DEV_MSTEP;   /* :16-17 */
    dev_cinv = 1.0f / DEV_CSTEP;
}

/* cache_bilinear_ldg (:408-430) on one layer of an (m or c) table. */
static inline void dev_bil2(const float* tm, const float* tc, float m, float c, float* om, float* oc)
{
    float fm = (m - DEV_MLO) * dev_minv;
    float fc = (c - DEV_CLO) * dev_cinv;
    fm = fmaxf(0.0f, fminf(fm, (float)(MEAN_N - 1)));
    fc = fmaxf(0.0f, fminf(fc, (float)(CV_N - 1)));
    int m0 = (int)fm, c0 = (int)fc;
    if (m0 == MEAN_N - 1) m0--;
    if (c0 == CV_N - 1) c0--;
    float wm = fm - (float)m0, wc = fc - (float)c0;
    int k = m0 * CV_N + c0;
    float omwm = 1.0f - wm, omwc = 1.0f - wc;
    *om = fmaf(wm, fmaf(wc, tm[k + CV_N + 1], omwc * tm[k + CV_N]),
               omwm * fmaf(wc, tm[k + 1], omwc * tm[k]));
    *oc = fmaf(wm, fmaf(wc, tc[k + CV_N + 1], omwc * tc[k + CV_N]),
               omwm * fmaf(wc, tc[k + 1], omwc * tc[k]));
}

/* __float2int_rz: truncate, saturating; NaN -> 0. */
static inline int f2i_rz(float x)
{
    if (x != x) return 0;
    if (x >= 2147483648.0f) return 2147483647;
    if (x <= -2147483648.0f) return (int)0x80000000u;
    return (int)x;
}

/* gamma_smc_fast_pow10 (:115-123): Schraudolph's exp bit construction. */
static inline float fast_pow10_f(float x)
{
    const float LN10 = 2.30258509299f, C1 = 1064872507.1541044f, C2 = 12102203.161561485f;
    int bits = f2i_rz(fmaf(C2, x * LN10, C1));
    float r;
    memcpy(&r, &bits, 4);
    return r;
}

float orc_fast_pow10(float x) { return fast_pow10_f(x); }

/* __double2int_rn: round half to even, saturating. */
static inline int d2i_rn(double x)
{
    if (x != x) return 0;
    double r = nearbyint(x);
    if (r >= 2147483647.0) return 2147483647;
    if (r <= -2147483648.0) return (int)0x80000000u;
    return (int)r;
}

static inline int seg_steps(double delta) { int r = d2i_rn(delta); return r > 1 ? r : 1; }

typedef struct {
    const float* T[T_COUNT];
    int N;
} tabs_f;

static void fwd_segment_f(float* m, float* c, const tabs_f* t, int seg, int het)
{
    const size_t G = GRID;
    while (seg > t->N) {
        dev_bil2(t->T[T_CACHE_M] + (size_t)(t->N - 1) * G, t->T[T_CACHE_C] + (size_t)(t->N - 1) * G, *m, *c, m, c);
        seg -= t->N;
    }
    const float* tm = t->T[het ? T_FHET_M : T_FHOM_M] + (size_t)(seg - 1) * G;
    const float* tc = t->T[het ? T_FHET_C : T_FHOM_C] + (size_t)(seg - 1) * G;
    dev_bil2(tm, tc, *m, *c, m, c);
}

static void bwd_segment_f(float* m, float* c, const tabs_f* t, int seg, int het)
{
    const size_t G = GRID;
    int rem = seg;
    while (rem > t->N) rem -= t->N;
    dev_bil2(t->T[het ? T_BHET_M : T_BHOM_M] + (size_t)(rem - 1) * G,
             t->T[het ? T_BHET_C : T_BHOM_C] + (size_t)(rem - 1) * G, *m, *c, m, c);
    seg -= rem;
    while (seg > 0) {
        dev_bil2(t->T[T_CACHE_M] + (size_t)(t->N - 1) * G, t->T[T_CACHE_C] + (size_t)(t->N - 1) * G, *m, *c, m, c);
        seg -= t->N;
    }
}

/*
 * Forward-backward of pairs over the site window [site_start, site_start+block_S).
 * site_start == 0 and block_S == S is infer()'s full-sequence walk
 * (:567-609, :678-754); otherwise the block walk (:632-669, :762-839),
 * whose first step is one bp when site_start > 0.
 * G: uint8 (n, S), het(s) = (G[i][s] ^ G[j][s]) & 1.
 * Outputs: site-major [block_S x P]; lower/upper/alpha/beta may be NULL.
 */
int orc_fb_window_f32(const uint8_t* G, int64_t n, int64_t S, const double* pos,
                      const int32_t* pi, const int32_t* pj, int64_t P,
                      const float* tables, int n_steps, float Ne,
                      int64_t site_start, int64_t block_S,
                      float* mean, float* lower, float* upper, float* alpha, float* beta)
{
    if (site_start < 0 || block_S < 0 || site_start + block_S > S || n_steps < 1) return -1;
    for (int64_t p = 0; p < P; p++)
        if (pi[p] < 0 || pi[p] >= n || pj[p] < 0 || pj[p] >= n) return -2;
    dev_consts();
    tabs_f t;
    size_t L = (size_t)n_steps * GRID;
    for (int k = 0; k < T_COUNT; k++) t.T[k] = tables + (size_t)k * L;
    t.N = n_steps;
    const float unscale = 2.0f * Ne;
    int ci = lower && upper;

    #pragma omp parallel
    {
        float* fm_buf = (float*)malloc(sizeof(float) * (size_t)(block_S > 0 ? block_S : 1));
        float* fc_buf = (float*)malloc(sizeof(float) * (size_t)(block_S > 0 ? block_S : 1));
        #pragma omp for schedule(dynamic, 1)
        for (int64_t p = 0; p < P; p++) {
            const uint8_t* gi = G + (size_t)pi[p] * S;
            const uint8_t* gj = G + (size_t)pj[p] * S;
            float m = 0.0f, c = 0.0f;
            double prev = (site_start == 0) ? -1.0 : (pos[site_start] - 1.0);
            for (int64_t s = 0; s < block_S; s++) {
                int64_t g = site_start + s;
                int seg = seg_steps(pos[g] - prev);
                prev = pos[g];
                fwd_segment_f(&m, &c, &t, seg, (gi[g] ^ gj[g]) & 1);
                fm_buf[s] = m;
                fc_buf[s] = c;
            }
            m = 0.0f; c = 0.0f;
            for (int64_t s = block_S - 1; s >= 0; s--) {
                int64_t g = site_start + s;
                float la = -2.0f * c;
                float ba = fast_pow10_f(la), bb = fast_pow10_f(la - m);
                float lf = -2.0f * fc_buf[s];
                float fa = fast_pow10_f(lf), fb = fast_pow10_f(lf - fm_buf[s]);
                float a_s = fa + ba - 1.0f;
                float b_s = fb + bb - 1.0f;
                float mg = (a_s / fmaxf(b_s, 1e-10f)) * unscale;
                size_t idx = (size_t)s * P + p;
                mean[idx] = mg;
                if (ci) {
                    float a_ci = fmaxf(a_s, 1.0f);
                    float inv9a = 1.0f / (9.0f * a_ci);
                    float sq = sqrtf(inv9a);
                    float base = 1.0f - inv9a;
                    float lo = fmaxf(fmaf(-1.96f, sq, base), 0.0f);
                    float hi = fmaf(1.96f, sq, base);
                    lower[idx] = fmaxf(mg * lo * lo * lo, 0.0f);
                    upper[idx] = mg * hi * hi * hi;
                }
                if (alpha) alpha[idx] = a_s;
                if (beta) beta[idx] = b_s;
                int seg;
                if (g == 0) seg = seg_steps(pos[0] + 1.0);
                else if (s == 0 && site_start > 0) seg = 1;
                else seg = seg_steps(pos[g] - pos[g - 1]);
                bwd_segment_f(&m, &c, &t, seg, (gi[g] ^ gj[g]) & 1);
            }
        }
        free(fm_buf);
        free(fc_buf);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Device walk, FP64 restatement                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    const double* T[T_COUNT];
    int N;
} tabs_d;

static inline void dev_bil2_d(const double* tm, const double* tc, double m, double c, double* om, double* oc)
{
    const double minv = 50.0 / 7.0, cinv = 49.0 / 2.0;
    double fm = (m + 5.0) * minv;
    double fc = (c + 2.0) * cinv;
    fm = fmax(0.0, fmin(fm, (double)(MEAN_N - 1)));
    fc = fmax(0.0, fmin(fc, (double)(CV_N - 1)));
    int m0 = (int)fm, c0 = (int)fc;
    if (m0 == MEAN_N - 1) m0--;
    if (c0 == CV_N - 1) c0--;
    double wm = fm - m0, wc = fc - c0;
    int k = m0 * CV_N + c0;
    double omwm = 1.0 - wm, omwc = 1.0 - wc;
    *om = fma(wm, fma(wc, tm[k + CV_N + 1], omwc * tm[k + CV_N]), omwm * fma(wc, tm[k + 1], omwc * tm[k]));
    *oc = fma(wm, fma(wc, tc[k + CV_N + 1], omwc * tc[k + CV_N]), omwm * fma(wc, tc[k + 1], omwc * tc[k]));
}

/* The same Schraudolph construction with FP64 intermediates, or the exact
 * power when exact != 0. */
static inline double pow10_d(double x, int exact)
{
    if (exact) return pow(10.0, x);
    double y = 12102203.161561485 * (x * 2.302585092994046) + 1064872507.1541044;
    int bits;
    if (y != y) bits = 0;
    else if (y >= 2147483647.0) bits = 2147483647;
    else if (y <= -2147483648.0) bits = (int)0x80000000u;
    else bits = (int)y;
    float r;
    memcpy(&r, &bits, 4);
    return (double)r;
}

static void fwd_segment_d(double* m, double* c, const tabs_d* t, int seg, int het)
{
    const size_t G = GRID;
    while (seg > t->N) {
        dev_bil2_d(t->T[T_CACHE_M] + (size_t)(t->N - 1) * G, t->T[T_CACHE_C] + (size_t)(t->N - 1) * G, *m, *c, m, c);
        seg -= t->N;
    }
    dev_bil2_d(t->T[het ? T_FHET_M : T_FHOM_M] + (size_t)(seg - 1) * G,
               t->T[het ? T_FHET_C : T_FHOM_C] + (size_t)(seg - 1) * G, *m, *c, m, c);
}

static void bwd_segment_d(double* m, double* c, const tabs_d* t, int seg, int het)
{
    const size_t G = GRID;
    int rem = seg;
    while (rem > t->N) rem -= t->N;
    dev_bil2_d(t->T[het ? T_BHET_M : T_BHOM_M] + (size_t)(rem - 1) * G,
               t->T[het ? T_BHET_C : T_BHOM_C] + (size_t)(rem - 1) * G, *m, *c, m, c);
    seg -= rem;
    while (seg > 0) {
        dev_bil2_d(t->T[T_CACHE_M] + (size_t)(t->N - 1) * G, t->T[T_CACHE_C] + (size_t)(t->N - 1) * G, *m, *c, m, c);
        seg -= t->N;
    }
}

int orc_fb_window_f64(const uint8_t* G, int64_t n, int64_t S, const double* pos,
                      const int32_t* pi, const int32_t* pj, int64_t P,
                      const double* tables, int n_steps, double Ne, int exact_pow10,
                      int64_t site_start, int64_t block_S,
                      double* mean, double* alpha, double* beta)
{
    if (site_start < 0 || block_S < 0 || site_start + block_S > S || n_steps < 1) return -1;
    for (int64_t p = 0; p < P; p++)
        if (pi[p] < 0 || pi[p] >= n || pj[p] < 0 || pj[p] >= n) return -2;
    tabs_d t;
    size_t L = (size_t)n_steps * GRID;
    for (int k = 0; k < T_COUNT; k++) t.T[k] = tables + (size_t)k * L;
    t.N = n_steps;
    const double unscale = 2.0 * Ne;

    #pragma omp parallel
    {
        double* fm_buf = (double*)malloc(sizeof(double) * (size_t)(block_S > 0 ? block_S : 1));
        double* fc_buf = (double*)malloc(sizeof(double) * (size_t)(block_S > 0 ? block_S : 1));
        #pragma omp for schedule(dynamic, 1)
        for (int64_t p = 0; p < P; p++) {
            const uint8_t* gi = G + (size_t)pi[p] * S;
            const uint8_t* gj = G + (size_t)pj[p] * S;
            double m = 0.0, c = 0.0;
            double prev = (site_start == 0) ? -1.0 : (pos[site_start] - 1.0);
            for (int64_t s = 0; s < block_S; s++) {
                int64_t g = site_start + s;
                int seg = seg_steps(pos[g] - prev);
                prev = pos[g];
                fwd_segment_d(&m, &c, &t, seg, (gi[g] ^ gj[g]) & 1);
                fm_buf[s] = m;
                fc_buf[s] = c;
            }
            m = 0.0; c = 0.0;
            for (int64_t s = block_S - 1; s >= 0; s--) {
                int64_t g = site_start + s;
                double la = -2.0 * c, lf = -2.0 * fc_buf[s];
                double a_s = pow10_d(lf, exact_pow10) + pow10_d(la, exact_pow10) - 1.0;
                double b_s = pow10_d(lf - fm_buf[s], exact_pow10) + pow10_d(la - m, exact_pow10) - 1.0;
                size_t idx = (size_t)s * P + p;
                mean[idx] = (a_s / fmax(b_s, 1e-10)) * unscale;
                if (alpha) alpha[idx] = a_s;
                if (beta) beta[idx] = b_s;
                int seg;
                if (g == 0) seg = seg_steps(pos[0] + 1.0);
                else if (s == 0 && site_start > 0) seg = 1;
                else seg = seg_steps(pos[g] - pos[g - 1]);
                bwd_segment_d(&m, &c, &t, seg, (gi[g] ^ gj[g]) & 1);
            }
        }
        free(fm_buf);
        free(fc_buf);
    }
    return 0;
}

/* Forward states only, FP32 (the reference's run_fwd_states debug path,
 * :567-609), for pinning the forward half on its own. */
int orc_forward_states_f32(const uint8_t* G, int64_t n, int64_t S, const double* pos,
                           int32_t i, int32_t j, const float* tables, int n_steps,
                           float* out_m, float* out_c)
{
    if (i < 0 || i >= n || j < 0 || j >= n || n_steps < 1) return -2;
    dev_consts();
    tabs_f t;
    size_t L = (size_t)n_steps * GRID;
    for (int k = 0; k < T_COUNT; k++) t.T[k] = tables + (size_t)k * L;
    t.N = n_steps;
    float m = 0.0f, c = 0.0f;
    double prev = -1.0;
    for (int64_t s = 0; s < S; s++) {
        int seg = seg_steps(pos[s] - prev);
        prev = pos[s];
        fwd_segment_f(&m, &c, &t, seg, (G[(size_t)i * S + s] ^ G[(size_t)j * S + s]) & 1);
        out_m[s] = m;
        out_c[s] = c;
    }
    return 0;
}

/* Device-side bilinear exposed for the indexing pins. */
void orc_dev_bilinear(const float* tm, const float* tc, float m, float c, float* om, float* oc)
{
    dev_consts();
    dev_bil2(tm, tc, m, c, om, oc);
}

int orc_seg_steps(double delta) { return seg_steps(delta); }

/* ------------------------------------------------------------------ */
/* Forward-only filter (src/kernels/gamma_smc_flow.cu:1139-1232,        */
/* src/bindings.cpp:2794-2849)                                          */
/* ------------------------------------------------------------------ */
/* (int)(x) of a double: truncate, saturating; NaN -> 0. */
static inline int d2i_rz(double x)
{
    if (x != x) return 0;
    if (x >= 2147483647.0) return 2147483647;
    if (x <= -2147483648.0) return (int)0x80000000u;
    return (int)x;
}

/* The h2 path, FP32: tm / tc are the composed transition tables cache[k]
 * (N layers of 2550 nodes) as the reference stores them (FP16, widened back
 * here by the caller). Per pair and site s: for s > 0, advance by
 * gap = (int)(pos[s] - pos[s-1] + 0.5) bp when gap > 0 (full N-bp layers
 * first, then layer gap-1); at a het site add 1 to alpha (exact libm 10^x
 * and log10 stand in for the reference's __exp10f / __log10f, and
 * b_log = -2c - m is fused as nvcc fuses it); then
 * mean = 2^(m * log2(10)) * 2Ne and, with CI, the Wilson-Hilferty bounds
 * (the two 1.96 terms fused). Outputs site-major [S][P]. */
int orc_fwd_only_f32(const uint8_t* G, int64_t n, int64_t S, const double* pos,
                     const int32_t* pi, const int32_t* pj, int64_t P,
                     const float* tm, const float* tc, int N, float Ne,
                     float* mean, float* lower, float* upper)
{
    dev_consts();
    const float two_ne = 2.0f * Ne;
    for (int64_t p = 0; p < P; ++p)
        if (pi[p] < 0 || pi[p] >= n || pj[p] < 0 || pj[p] >= n) return -1;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < P; ++p) {
        const uint8_t* gi = G + (size_t)pi[p] * (size_t)S;
        const uint8_t* gj = G + (size_t)pj[p] * (size_t)S;
        float m = 0.0f, c = 0.0f;
        for (int64_t s = 0; s < S; ++s) {
            if (s > 0) {
                int gap = d2i_rz(pos[s] - pos[s - 1] + 0.5);
                if (gap > 0) {
                    while (gap > N) {
                        dev_bil2(tm + (size_t)(N - 1) * GRID, tc + (size_t)(N - 1) * GRID, m, c, &m, &c);
                        gap -= N;
                    }
                    dev_bil2(tm + (size_t)(gap - 1) * GRID, tc + (size_t)(gap - 1) * GRID, m, c, &m, &c);
                }
            }
            if ((gi[s] ^ gj[s]) & 1) {
                const float alpha = exp10f(-2.0f * c) + 1.0f;
                const float a_log = log10f(alpha);
                const float b_log = fmaf(-2.0f, c, -m);
                m = a_log - b_log;
                c = -0.5f * a_log;
            }
            const float mg = exp2f(m * 3.321928094887362f) * two_ne;
            const size_t idx = (size_t)s * (size_t)P + (size_t)p;
            mean[idx] = mg;
            if (lower && upper) {
                const float alpha = exp10f(-2.0f * c);
                const float inv9a = 1.0f / (9.0f * alpha);
                const float sq = sqrtf(inv9a);
                const float base = 1.0f - inv9a;
                const float lo = fmaxf(fmaf(-1.96f, sq, base), 0.0f);
                const float hi = fmaf(1.96f, sq, base);
                lower[idx] = fmaxf(mg * lo * lo * lo, 0.0f);
                upper[idx] = mg * hi * hi * hi;
            }
        }
    }
    return 0;
}

/* The same recursion in FP64 on FP64 tables (no FP16 rounding), exact
 * powers and logarithms: the reference point of the FP32 path's error. */
int orc_fwd_only_f64(const uint8_t* G, int64_t n, int64_t S, const double* pos,
                     const int32_t* pi, const int32_t* pj, int64_t P,
                     const double* tm, const double* tc, int N, double Ne, double* mean)
{
    for (int64_t p = 0; p < P; ++p)
        if (pi[p] < 0 || pi[p] >= n || pj[p] < 0 || pj[p] >= n) return -1;
    #pragma omp parallel for schedule(dynamic, 1)
    for (int64_t p = 0; p < P; ++p) {
        const uint8_t* gi = G + (size_t)pi[p] * (size_t)S;
        const uint8_t* gj = G + (size_t)pj[p] * (size_t)S;
        double m = 0.0, c = 0.0;
        for (int64_t s = 0; s < S; ++s) {
            if (s > 0) {
                int gap = d2i_rz(pos[s] - pos[s - 1] + 0.5);
                if (gap > 0) {
                    while (gap > N) {
                        dev_bil2_d(tm + (size_t)(N - 1) * GRID, tc + (size_t)(N - 1) * GRID, m, c, &m, &c);
                        gap -= N;
                    }
                    dev_bil2_d(tm + (size_t)(gap - 1) * GRID, tc + (size_t)(gap - 1) * GRID, m, c, &m, &c);
                }
            }
            if ((gi[s] ^ gj[s]) & 1) {
                const double a_log = log10(pow(10.0, -2.0 * c) + 1.0);
                const double b_log = -2.0 * c - m;
                m = a_log - b_log;
                c = -0.5 * a_log;
            }
            mean[(size_t)s * (size_t)P + (size_t)p] = pow(10.0, m) * 2.0 * Ne;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Log-quantized output codes (src/kernels/gamma_smc.cu:23-39)          */
/* ------------------------------------------------------------------ */
/* code = (unsigned)clamp(fma(logf(max(mean, 1)) * inv_range, levels, 0.5),
 * 0, levels) for n means; levels 255 or 15. */
void orc_log_codes(const float* mean, int64_t n, float inv_range, float levels, uint8_t* out)
{
    for (int64_t i = 0; i < n; ++i) {
        const float x = logf(fmaxf(mean[i], 1.0f)) * inv_range;
        const float q = fmaf(x, levels, 0.5f);
        out[i] = (uint8_t)fminf(fmaxf(q, 0.0f), levels);
    }
}
