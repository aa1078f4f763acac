// Host construction of the composed per-bp transition tables read by the
// decoder kernels.
//
// The tables follow the reference's build_flow_field_cache
// (src/flow_field.cpp:205-388): for every node of the 51 x 50 (m, c) grid,
//   stretch[k] = R(E_hom(stretch[k-1])), stretch[0] = R(node)
//   fwd_h[k]   = clip(E_site(h)(stretch[k](node)))
//   bwd_h[k]   = B_host[stretch[k]](clip(E_site(h)(node)))
// where R is one recombination step along the (u, v) flow field with the
// entropy clip, E_hom adds mu to beta, and E_site additionally adds 1 to
// alpha at a het site. Two precisions:
//   kTablesF32 - FP32 arithmetic with the same operation order and libm
//                calls as the reference, so the tables are bit-identical to
//                the reference's (compile with g++ -ffp-contract=off, no
//                -ffast-math);
//   kTablesF64 - the same construction in double, rounded to float only
//                when stored.
//
// Output layout (what the kernels read): per step k, a [2][2550] float2
// slab {(m, c) of hom, (m, c) of het} for the forward and backward tables,
// plus the stretch layer N-1 (the only stretch layer the walk reads).
#include "tables.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace gsmc {
namespace {

constexpr double kEntropyMaxMean = 4.0;   // flow_field.cpp:80-83
constexpr double kEntropyStep = 0.0008;
constexpr int kEntropyN = 5001;

// Asymptotic digamma with upward recurrence to x >= 6.
double digamma_series(double x)
{
    double acc = 0.0;
    for (; x < 6.0; x += 1.0) acc -= 1.0 / x;
    const double r = 1.0 / x, r2 = r * r, r4 = r2 * r2, r6 = r4 * r2, r8 = r4 * r4, r10 = r8 * r2;
    acc += std::log(x) - 0.5 * r - r2 / 12.0 + r4 / 120.0 - r6 / 252.0 + r8 / 240.0 - 5.0 * r10 / 660.0;
    return acc;
}

double entropy_of(double m, double c)
{
    const double la = -2.0 * c, lb = la - m;
    const double a = std::pow(10.0, la), b = std::pow(10.0, lb);
    return a - std::log(b) + std::lgamma(a) + (1.0 - a) * digamma_series(a);
}

// CV (log10) at which the Gamma's differential entropy reaches 1, per mean
// step of 0.0008 on [0, 4]; 60 bisection steps on [-2, 0].
struct EntropyLimit {
    double d[kEntropyN];
    float f[kEntropyN];
    EntropyLimit()
    {
        for (int i = 0; i < kEntropyN; ++i) {
            const double m = i * kEntropyStep;
            double lo = -2.0, hi = 0.0, r;
            if (entropy_of(m, lo) > 1.0) {
                r = lo;
            } else if (entropy_of(m, hi) <= 1.0) {
                r = hi;
            } else {
                for (int it = 0; it < 60; ++it) {
                    const double mid = 0.5 * (lo + hi);
                    (entropy_of(m, mid) > 1.0 ? hi : lo) = mid;
                }
                r = hi;
            }
            d[i] = r;
            f[i] = static_cast<float>(r);
        }
    }
};

const EntropyLimit& entropy_limit()
{
    static const EntropyLimit t;
    return t;
}

// Precision-generic pieces. `Real` is float or double; the float branch
// must evaluate exactly like the reference's host code.
template <class Real> struct Ops;
template <> struct Ops<float> {
    static float p10(float x) { return powf(10.0f, x); }
    static float lg(float x) { return log10f(x); }
    static float limit(int i) { return entropy_limit().f[i]; }
};
template <> struct Ops<double> {
    static double p10(double x) { return std::pow(10.0, x); }
    static double lg(double x) { return std::log10(x); }
    static double limit(int i) { return entropy_limit().d[i]; }
};

template <class Real> struct Box {
    Real mlo, mhi, clo, chi, mstep, cstep;
};

template <class Real> inline Real lo_of(Real a, Real b) { return (b < a) ? b : a; }  // std::min
template <class Real> inline Real hi_of(Real a, Real b) { return (a < b) ? b : a; }  // std::max

template <class Real> void clip(Real& m, Real& c, const Box<Real>& b)
{
    m = hi_of(b.mlo, lo_of(m, b.mhi));
    c = hi_of(b.clo, lo_of(c, b.chi));
    if (m >= Real(0) && m <= kEntropyMaxMean) {
        const Real idx = m / static_cast<Real>(kEntropyStep);
        const int i0 = static_cast<int>(idx);
        if (i0 >= kEntropyN - 1) {
            c = lo_of(c, Ops<Real>::limit(kEntropyN - 1));
            return;
        }
        const Real w = idx - i0;
        const Real lim = (Real(1) - w) * Ops<Real>::limit(i0) + w * Ops<Real>::limit(i0 + 1);
        c = lo_of(c, lim);
        c = hi_of(b.clo, lo_of(c, b.chi));
    }
}

// Host bilinear: fractional index by division by the grid step.
template <class Real> Real interp(const Real* t, Real m, Real c, const Box<Real>& b)
{
    Real fm = (m - b.mlo) / b.mstep;
    Real fc = (c - b.clo) / b.cstep;
    fm = hi_of(Real(0), lo_of(fm, static_cast<Real>(kMeanN - 1)));
    fc = hi_of(Real(0), lo_of(fc, static_cast<Real>(kCvN - 1)));
    int m0 = static_cast<int>(fm), c0 = static_cast<int>(fc);
    if (m0 == kMeanN - 1) --m0;
    if (c0 == kCvN - 1) --c0;
    const Real wm = fm - m0, wc = fc - c0;
    const int k = m0 * kCvN + c0;
    return (1 - wm) * ((1 - wc) * t[k] + wc * t[k + 1]) + wm * ((1 - wc) * t[k + kCvN] + wc * t[k + kCvN + 1]);
}

template <class Real>
void recombine(Real& m, Real& c, const Real* u, const Real* v, Real rho, const Box<Real>& b)
{
    clip(m, c, b);
    const Real du = interp(u, m, c, b);
    const Real dv = interp(v, m, c, b);
    m += du * rho;
    c += dv * rho;
    clip(m, c, b);
}

template <class Real> void emit_site(bool het, Real mu, Real& m, Real& c)
{
    Real la = Real(-2) * c;
    Real lb = la - m;
    lb = Ops<Real>::lg(Ops<Real>::p10(lb) + mu);
    if (het) la = Ops<Real>::lg(Ops<Real>::p10(la) + Real(1));
    m = la - lb;
    c = Real(-0.5) * la;
}

template <class Real>
void build(const FlowField& ff, int N, Real rho, Real mu, HostTables& out)
{
    Box<Real> b;
    b.mlo = static_cast<Real>(ff.bounds[0]);
    b.mhi = static_cast<Real>(ff.bounds[1]);
    b.clo = static_cast<Real>(ff.bounds[2]);
    b.chi = static_cast<Real>(ff.bounds[3]);
    b.mstep = (b.mhi - b.mlo) / (kMeanN - 1);
    b.cstep = (b.chi - b.clo) / (kCvN - 1);
    std::vector<Real> u(kGrid), v(kGrid);
    for (int i = 0; i < kGrid; ++i) { u[i] = ff.u[i]; v[i] = ff.v[i]; }

    // Stretch layers, kept in full precision while composing.
    std::vector<Real> sm((size_t)N * kGrid), sc((size_t)N * kGrid);
    #pragma omp parallel for schedule(static)
    for (int g = 0; g < kGrid; ++g) {
        Real m = b.mlo + (g / kCvN) * b.mstep;
        Real c = b.clo + (g % kCvN) * b.cstep;
        recombine(m, c, u.data(), v.data(), rho, b);
        sm[g] = m;
        sc[g] = c;
        for (int k = 1; k < N; ++k) {
            const Real la = Real(-2) * c;
            Real lb = la - m;
            const Real bl = Ops<Real>::p10(lb) + mu;
            lb = Ops<Real>::lg(bl);
            m = la - lb;
            recombine(m, c, u.data(), v.data(), rho, b);
            sm[(size_t)k * kGrid + g] = m;
            sc[(size_t)k * kGrid + g] = c;
        }
    }

    out.n_steps = N;
    out.fwd.assign((size_t)N * 2 * kGrid * 2, 0.0f);
    out.bwd.assign((size_t)N * 2 * kGrid * 2, 0.0f);
    out.stretch_last.assign((size_t)kGrid * 2, 0.0f);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < N; ++k) {
        const Real* lm = sm.data() + (size_t)k * kGrid;
        const Real* lc = sc.data() + (size_t)k * kGrid;
        for (int h = 0; h < 2; ++h) {
            float* fw = out.fwd.data() + ((size_t)k * 2 + h) * kGrid * 2;
            float* bw = out.bwd.data() + ((size_t)k * 2 + h) * kGrid * 2;
            for (int g = 0; g < kGrid; ++g) {
                Real m = lm[g], c = lc[g];
                emit_site<Real>(h != 0, mu, m, c);
                clip(m, c, b);
                fw[2 * g] = static_cast<float>(m);
                fw[2 * g + 1] = static_cast<float>(c);
                Real mb = b.mlo + (g / kCvN) * b.mstep;
                Real cb = b.clo + (g % kCvN) * b.cstep;
                emit_site<Real>(h != 0, mu, mb, cb);
                clip(mb, cb, b);
                bw[2 * g] = static_cast<float>(interp(lm, mb, cb, b));
                bw[2 * g + 1] = static_cast<float>(interp(lc, mb, cb, b));
            }
        }
    }
    for (int g = 0; g < kGrid; ++g) {
        out.stretch_last[2 * g] = static_cast<float>(sm[(size_t)(N - 1) * kGrid + g]);
        out.stretch_last[2 * g + 1] = static_cast<float>(sc[(size_t)(N - 1) * kGrid + g]);
    }
    out.cache.resize((size_t)N * kGrid * 2);
    for (size_t i = 0; i < (size_t)N * kGrid; ++i) {
        out.cache[2 * i] = static_cast<float>(sm[i]);
        out.cache[2 * i + 1] = static_cast<float>(sc[i]);
    }
}

}  // namespace

bool load_flow_field(const char* path, FlowField& ff, std::string& err)
{
    FILE* f = std::fopen(path, "r");
    if (!f) {
        err = std::string("cannot open flow field file: ") + path;
        return false;
    }
    float lo, hi;
    int n;
    bool ok = std::fscanf(f, "%f %f %d", &lo, &hi, &n) == 3 && n == kMeanN;
    if (ok) {
        ff.bounds[0] = log10f(lo);
        ff.bounds[1] = log10f(hi);
        ok = std::fscanf(f, "%f %f %d", &lo, &hi, &n) == 3 && n == kCvN;
    }
    if (ok) {
        ff.bounds[2] = log10f(lo);
        ff.bounds[3] = log10f(hi);
        for (int i = 0; ok && i < kGrid; ++i) ok = std::fscanf(f, "%f", &ff.u[i]) == 1;
        for (int i = 0; ok && i < kGrid; ++i) ok = std::fscanf(f, "%f", &ff.v[i]) == 1;
    }
    std::fclose(f);
    if (!ok) err = std::string("malformed flow field file (expected a 51 x 50 grid): ") + path;
    return ok;
}

void build_tables(const FlowField& ff, int n_steps, float scaled_rho, float scaled_mu,
                  double scaled_rho64, double scaled_mu64, int precision, HostTables& out)
{
    if (precision == kTablesF64)
        build<double>(ff, n_steps, scaled_rho64, scaled_mu64, out);
    else
        build<float>(ff, n_steps, scaled_rho, scaled_mu, out);
}

}  // namespace gsmc
