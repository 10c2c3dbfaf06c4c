// nait.hip -- DensityExpression 5 (nAiT) on gfx950.
//
// n(r) = beta0 a_r tau beta1 int CurlyA(x) [Li_{1/2}(arg) + b Li_{-3/2}(arg)] dx,
// arg = -(U - x a_r)/T, by the reference's adaptive Boole bisection per grid
// point (getDenEnAt, mpDPFT.cpp:3363-3453; AdaptiveBoole_nAiT :3553-3578;
// nAiTintegrand :3580-3623).  One thread walks one point's bisection tree
// depth first with an explicit stack; accepted panels are summed left to
// right.  The Airy kinetic energy EkinAiT03D (:5610-5709) is the second kernel.
//
// Special functions: CurlyA = int_y^inf Ai as the reference evaluates it (a
// cubic spline with parabolic ends through the knots -150 + 0.01 i, its
// asymptotic forms outside), Li_s(-e^x) and Ai by quintic Hermite
// interpolation of the chain tables in nait_tables.bin (built by
// mpdpft_amd/tables.py), with the reference's branches for |x| > 699.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "solver.h"

namespace {

constexpr double kPI = 3.141592653589793;
constexpr double kMP = 1.0e-12;
constexpr int kStack = 64;
// Nominal FP64 operations of one integrand evaluation on the table path:
// arg 3, Li pair (index 8, Hermite basis 36, two combinations 30) 74,
// CurlyA cubic 11, product 4, Boole weight and accumulate 2.
constexpr double kNaitFlopsPerEval = 94.;
// The panel-reuse kernel's evaluation on its table path: abscissa 2, argument
// 3, Li index and offset 6, two degree-5 Horner chains 20, CurlyA index,
// offset and cubic 11, product 4, Boole weights into both half sums 4.
constexpr double kNaitReuseFlopsPerEval = 50.;

struct TabView {
    const double4 *cc;  // CurlyA cubic per knot interval: c0 + u(c1 + u(c2 + u c3))
    const double *p;    // [pn][6] Li_s(-e^x), s = 3/2 .. -7/2
    const double2 *a;   // (Ai, Ai')
    int cn, pn, an;
    double cx0, ch, cinv, px0, ph, pinv, pxmax, ax0, ah, ainv;
    double ginv[4];      // 1/Gamma(s+1), s = 3/2, 1/2, -1/2, -3/2 (as the reference spells them)
    double pw2[4], pw3[4];  // 2^-s, 3^-s
    const double *asym;     // [4][12] Sommerfeld coefficients (device copy of kAsym)
    const double2 *pp;      // [pn - 1 + nb][6]: per interval, the monomial coefficients in u of the
                            // quintic Hermite interpolants of Li_{1/2} (0..5) and Li_{-3/2} (6..11);
                            // intervals pn - 1 ... cover (pxmax, pxmax + nb] with unit step
    int nb;
};

// Sommerfeld coefficients 2 eta(2k) (s)_{2k}, k = 1..12, for s = 3/2 - j
constexpr double kTwoEta[12] = {1.6449340668482264, 1.8940656589944918, 1.9711021825948705, 1.992466003705296,
                                1.9980790151965429, 1.9995153702877164, 1.9998783406919596, 1.9999695284298122,
                                1.9999923757392204, 1.999998093223163,  1.9999995232264616, 1.9999998807977848};
struct AsymTable {
    double v[4][12];
};
constexpr AsymTable make_asym() {
    AsymTable r{};
    for (int j = 0; j < 4; j++) {
        double sj = 1.5 - (double)j, ff = 1.;
        for (int k = 1; k <= 12; k++) {
            ff *= (sj - 2 * k + 2) * (sj - 2 * k + 1);
            r.v[j][k - 1] = kTwoEta[k - 1] * ff;
        }
    }
    return r;
}
constexpr AsymTable kAsym = make_asym();

// x^s for the four half-integer orders
__device__ __forceinline__ double xpow(int j, double x) {
    double r = sqrt(x);
    return j == 0 ? x * r : (j == 1 ? r : (j == 2 ? 1. / r : 1. / (x * r)));
}

// CurlyA's oscillatory asymptotic form (y < -149.99), kept out of line: only
// panels below the knot range reach it
__device__ __noinline__ double curly_a_asym(double y) {
    return 1. - cos(0.25 * 3.141592653589793 + 0.6666666666666666666 * (-y) * sqrt(-y)) /
                    sqrt(3.141592653589793 * (-y) * sqrt(-y));
}

__device__ __forceinline__ double curly_a(const TabView &t, double y) {
    if (y < -1.0e+50) return 1.;
    if (y < -149.99) return curly_a_asym(y);
    if (y >= 99.99) return 0.;
    int j = (int)floor((y - t.cx0) * t.cinv);
    j = j < 0 ? 0 : (j > t.cn - 2 ? t.cn - 2 : j);
    double u = y - (t.cx0 + t.ch * (double)j);
    double4 c = t.cc[j];
    return c.x + u * (c.y + u * (c.z + u * c.w));
}

__device__ __forceinline__ double qherm(double t, double h, double f0, double g0, double c0, double f1, double g1,
                                        double c1) {
    double t2 = t * t, t3 = t2 * t, t4 = t3 * t, t5 = t4 * t;
    double H0 = 1. - 10. * t3 + 15. * t4 - 6. * t5, H1 = t - 6. * t3 + 8. * t4 - 3. * t5;
    double H2 = 0.5 * t2 - 1.5 * t3 + 1.5 * t4 - 0.5 * t5, H3 = 10. * t3 - 15. * t4 + 6. * t5;
    double H4 = -4. * t3 + 7. * t4 - 3. * t5, H5 = 0.5 * t3 - t4 + 0.5 * t5;
    return f0 * H0 + h * g0 * H1 + h * h * c0 * H2 + f1 * H3 + h * g1 * H4 + h * h * c1 * H5;
}

// Li_s(-e^x) for chain index j (s = 3/2 - j), j in 0..3
__device__ double polylog_j(const TabView &t, int j, double x) {
    if (x < -699.) return 0.;
    if (x > 699.) return -xpow(j, x) * t.ginv[j];
    if (x < t.px0) {
        double e = exp(x);
        return -e + e * e * t.pw2[j] - e * e * e * t.pw3[j];
    }
    if (x > t.pxmax) {
        double x2 = 1. / (x * x), acc = 0.;
        for (int k = 11; k >= 0; k--) acc = x2 * (t.asym[12 * j + k] + acc);
        return -xpow(j, x) * t.ginv[j] * (1. + acc);
    }
    int k = (int)floor((x - t.px0) * t.pinv);
    if (k > t.pn - 2) k = t.pn - 2;
    double u = (x - (t.px0 + t.ph * (double)k)) * t.pinv;
    const double *a = t.p + 6 * k, *b = a + 6;
    return qherm(u, t.ph, a[j], a[j + 1], a[j + 2], b[j], b[j + 1], b[j + 2]);
}

// Li_{1/2}(-e^x) and Li_{-3/2}(-e^x) together (the integrand's pair)
__device__ __forceinline__ void polylog_pair(const TabView &t, double x, double &l1, double &l3) {
    if (x < -699.) {
        l1 = 0.;
        l3 = 0.;
        return;
    }
    if (x < t.px0) {  // the series branch of polylog_j for both orders, one exp
        double e = exp(x), e2 = e * e, e3 = e2 * e;
        l1 = -e + e2 * t.pw2[1] - e3 * t.pw3[1];
        l3 = -e + e2 * t.pw2[3] - e3 * t.pw3[3];
        return;
    }
    if (x > t.pxmax) {  // the asymptotic branches of polylog_j for both orders
        double r = sqrt(x), p1 = r, p3 = 1. / (x * r);
        if (x > 699.) {
            l1 = -p1 * t.ginv[1];
            l3 = -p3 * t.ginv[3];
            return;
        }
        double x2 = 1. / (x * x), c1 = 0., c3 = 0.;
#pragma unroll 1
        for (int k = 11; k >= 0; k--) {
            c1 = x2 * (t.asym[12 + k] + c1);
            c3 = x2 * (t.asym[36 + k] + c3);
        }
        l1 = -p1 * t.ginv[1] * (1. + c1);
        l3 = -p3 * t.ginv[3] * (1. + c3);
        return;
    }
    int k = (int)floor((x - t.px0) * t.pinv);
    if (k > t.pn - 2) k = t.pn - 2;
    double u = (x - (t.px0 + t.ph * (double)k)) * t.pinv;
    const double *a = t.p + 6 * k, *b = a + 6;
    double a1 = a[1], a2 = a[2], a3 = a[3], a4 = a[4], a5 = a[5];
    double b1 = b[1], b2 = b[2], b3 = b[3], b4 = b[4], b5 = b[5];
    double u2 = u * u, u3 = u2 * u, u4 = u3 * u, u5 = u4 * u, h = t.ph;
    double H0 = 1. - 10. * u3 + 15. * u4 - 6. * u5, H1 = u - 6. * u3 + 8. * u4 - 3. * u5;
    double H2 = 0.5 * u2 - 1.5 * u3 + 1.5 * u4 - 0.5 * u5, H3 = 10. * u3 - 15. * u4 + 6. * u5;
    double H4 = -4. * u3 + 7. * u4 - 3. * u5, H5 = 0.5 * u3 - u4 + 0.5 * u5;
    l1 = a1 * H0 + h * a2 * H1 + h * h * a3 * H2 + b1 * H3 + h * b2 * H4 + h * h * b3 * H5;
    l3 = a3 * H0 + h * a4 * H1 + h * h * a5 * H2 + b3 * H3 + h * b4 * H4 + h * h * b5 * H5;
}

// Table geometry of nait_tables.bin as written by mpdpft_amd/tables.py; the
// host checks the loaded file against these before using the panel-reuse
// kernel, whose lookups then need no runtime parameters.
constexpr double kCX0 = -150., kCH = 0.01, kCINV = 100.;
constexpr int kCN = 25001;
constexpr double kPX0 = -40., kPH = 0.01, kPINV = 100., kPXMAX = 200.;
constexpr int kPN = 24001, kNB = 500;

// What the panel-reuse kernel reads in its inner loop.
struct ReuseTab {
    const double2 *pp;  // [kPN - 1 + kNB][6] monomial coefficients (see TabView::pp)
    const double4 *cc;  // CurlyA cubic per knot interval
    double ginv1, ginv3;
};

__device__ __forceinline__ double curly_a_fast(const ReuseTab &t, double y) {
    if (y < -1.0e+50) return 1.;
    if (y < -149.99) return curly_a_asym(y);
    if (y >= 99.99) return 0.;
    int j = (int)floor((y - kCX0) * kCINV);  // 0 < j < kCN - 2 for -149.99 <= y < 99.99
    double u = y - (kCX0 + kCH * (double)j);
    double4 c = t.cc[j];
    return c.x + u * (c.y + u * (c.z + u * c.w));
}

// e^x for the far branch, -699 <= x < -40, inline: Cody-Waite reduction x = n
// ln2 + r with the two-part ln2 of fdlibm (n ln2_hi exact), |r| <= ln2/2, the
// degree-13 Taylor polynomial of e^r (truncation < 5e-18 relative), then the
// exponent n.  Results stay normal in this range (e^-699 ~ 2.6e-304); within
// 2 ulp of exp.
__device__ __forceinline__ double exp_far(double x) {
    const double n = rint(x * 1.4426950408889634074);
    double r = fma(-n, 6.93147180369123816490e-01, x);
    r = fma(-n, 1.90821492927058770002e-10, r);
    double p = 1. / 6227020800.;  // 1/13!
    p = fma(p, r, 1. / 479001600.);
    p = fma(p, r, 1. / 39916800.);
    p = fma(p, r, 1. / 3628800.);
    p = fma(p, r, 1. / 362880.);
    p = fma(p, r, 1. / 40320.);
    p = fma(p, r, 1. / 5040.);
    p = fma(p, r, 1. / 720.);
    p = fma(p, r, 1. / 120.);
    p = fma(p, r, 1. / 24.);
    p = fma(p, r, 1. / 6.);
    p = fma(p, r, 0.5);
    p = fma(p, r, 1.);
    p = fma(p, r, 1.);
    return ldexp(p, (int)n);
}

// The integrand's Li pair for the panel-reuse kernel.  Below px0 and above
// 699 the branches of polylog_j; on [px0, pxmax] the quintic Hermite
// interpolants of the chain table in monomial form; on (pxmax, 699] a second
// quintic Hermite table with unit step built from the Sommerfeld expansion
// that polylog_j sums there (agreement ~1e-17 relative).
__device__ __forceinline__ void polylog_pair_fast(const ReuseTab &t, double x, double &l1, double &l3) {
    if (x < kPX0) {
        if (x < -699.) {
            l1 = l3 = 0.;
        } else {
            // -e + e^2 2^-s - e^3 3^-s rounds to -e: e <= e^-40, so the two
            // corrections lie below half an ulp of e
            l1 = l3 = -exp_far(x);
        }
        return;
    }
    if (x > 699.) {
        double r = sqrt(x);
        l1 = -r * t.ginv1;
        l3 = -(1. / (x * r)) * t.ginv3;
        return;
    }
    int k;
    double u;
    if (x <= kPXMAX) {
        k = (int)floor((x - kPX0) * kPINV);
        if (k > kPN - 2) k = kPN - 2;
        u = (x - (kPX0 + kPH * (double)k)) * kPINV;
    } else {
        int m = (int)floor(x - kPXMAX);
        if (m > kNB - 1) m = kNB - 1;
        u = x - (kPXMAX + (double)m);
        k = kPN - 1 + m;
    }
    const double2 *c = t.pp + 6 * k;
    double2 c01 = c[0], c23 = c[1], c45 = c[2], d01 = c[3], d23 = c[4], d45 = c[5];
    l1 = ((((c45.y * u + c45.x) * u + c23.y) * u + c23.x) * u + c01.y) * u + c01.x;
    l3 = ((((d45.y * u + d45.x) * u + d23.y) * u + d23.x) * u + d01.y) * u + d01.x;
}

__device__ double airy_ai(const TabView &t, double y) {
    if (y > 100.) return 0.;
    if (y <= -300.) return cos(2. * (-y) * sqrt(-y) / 3. - 0.25 * kPI) / (sqrt(kPI) * pow(-y, 0.25));
    int k = (int)floor((y - t.ax0) * t.ainv);
    k = k < 0 ? 0 : (k > t.an - 2 ? t.an - 2 : k);
    double x0 = t.ax0 + t.ah * (double)k, x1 = x0 + t.ah;
    double u = (y - x0) * t.ainv;
    double2 A = t.a[k], B = t.a[k + 1];
    return qherm(u, t.ah, A.x, A.y, x0 * A.x, B.x, B.y, x1 * B.x);
}

__device__ __forceinline__ bool negligible(double big, double small, double maxexp) {
    if (small == 0.) return true;
    if (big == 0. || big < small) return false;
    return log10(fabs(big)) - log10(fabs(small)) > maxexp;
}

// wavefront-uniform sum: every lane ends with the same bits (xor butterfly,
// each exchange adds the same two operands)
__device__ __forceinline__ double wave_sum(double v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

struct PointArgs {
    double pref, pref2, U, ar, invT;
};

__device__ __forceinline__ double integrand(const TabView &t, const PointArgs &q, double x) {
    double arg = -(q.U - x * q.ar) * q.invT;
    double l1, l3;
    polylog_pair(t, arg, l1, l3);
    return q.pref * curly_a(t, x) * (l1 + q.pref2 * l3);
}

__device__ __forceinline__ double integrand_fast(const ReuseTab &t, const PointArgs &q, double x) {
    double arg = -(q.U - x * q.ar) * q.invT;
    double l1, l3;
    polylog_pair_fast(t, arg, l1, l3);
    return q.pref * curly_a_fast(t, x) * (l1 + q.pref2 * l3);
}

// Boole weight of sample yi of n = 4 BP + 1 (BooleRule1D, mpDPFT.cpp:8050-8058)
__device__ __forceinline__ double boole_w(int yi, int n) {
    if (yi == 0 || yi == n - 1) return 7.;
    int m = (yi - 1) & 3;
    return m == 1 ? 12. : (m == 3 ? 14. : 32.);
}

struct Counts {
    double itermax, relacc, evals;
};

struct WaveStack {
    double a[64], b[64], prev[64];
    int it[64];
};

// AdaptiveBoole_nAiT for one grid point, all 64 lanes together: the 2(4BP+1)
// samples of a node are spread over the lanes, the panel sums reduced across
// the wavefront; the tree walk is uniform (depth first, left panel first).
__device__ double adaptive_boole_wave(const TabView &t, const PointArgs &q, int BP, int maxIter, double rtpna,
                                      double compare, double maxexp, double a0, double b0, Counts &cnt,
                                      WaveStack &S, int lane) {
    int sp = 0;
    S.a[0] = a0;
    S.b[0] = b0;
    S.prev[0] = 0.;
    S.it[0] = 0;
    sp = 1;
    __builtin_amdgcn_wave_barrier();
    const int n = 4 * BP + 1, tot = 2 * n;
    double total = 0.;
    while (sp > 0) {
        sp--;
        double a = S.a[sp], b = S.b[sp], prev = S.prev[sp];
        int iter = S.it[sp];
        __builtin_amdgcn_wave_barrier();
        double half = 0.5 * (b - a), Delta = 0.25 * half / ((double)BP);
        double sa = 0., sb = 0.;
        for (int base = 0; base < tot; base += 64) {
            int idx = base + lane;
            if (idx < tot) {
                bool right = idx >= n;
                int yi = right ? idx - n : idx;
                double x = a + (double)yi * Delta;
                if (right) x = x + half;
                double f = boole_w(yi, n) * integrand(t, q, x);
                if (right) sb += f;
                else sa += f;
            }
        }
        sa = wave_sum(sa);
        sb = wave_sum(sb);
        cnt.evals += (double)tot;
        double ha = (a + half - a) / (((double)BP) * 4.), hb = (b - (a + half)) / (((double)BP) * 4.);
        double Inta = sa * (2. * ha / 45.), Intb = sb * (2. * hb / 45.);
        double Int = Inta + Intb;
        if (Int != Int) {  // a NaN panel makes the point's density NaN, mapped to 0
            total += Int;
        } else if (iter > 1 && fabs(Int) > 0. && fabs((Int - prev) / Int) < rtpna) {
            cnt.relacc += 1.;
            total += Int;
        } else if (iter > maxIter) {
            cnt.itermax += 1.;
            total += Int;
        } else if (negligible(compare, pow(2., (double)iter) * Int, maxexp)) {
            total += Int;
        } else if (Delta < kMP) {
            total += Int;
        } else {
            if (sp + 2 > 64) return __builtin_nan("");
            S.a[sp] = a + half;
            S.b[sp] = b;
            S.prev[sp] = Intb;
            S.it[sp] = iter + 1;
            S.a[sp + 1] = a;
            S.b[sp + 1] = a + half;
            S.prev[sp + 1] = Inta;
            S.it[sp + 1] = iter + 1;
            sp += 2;
            __builtin_amdgcn_wave_barrier();
        }
    }
    return total;
}

struct DensityArgs {
    const double *V, *grad2, *lap;
    double mu, T, degeneracy, alpha, thr, rtpna, compare;
    int BP, maxIter;
};

__device__ __forceinline__ double trap_w(const mk::Grid &g, int idx) {
    double w = 1.;
    int rem = idx;
    for (int d = 0; d < g.D; d++) {
        int c = rem % g.EL;
        rem /= g.EL;
        if (c == 0 || c == g.s) w *= 0.5;
    }
    return w;
}

__device__ __forceinline__ double bsum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

// One wavefront per grid point (grid-stride over points).  partials: [0]
// weighted density sum, [1] iterMax returns, [2] relAcc returns.
__global__ __launch_bounds__(256) void k_nait_density(
    mk::Grid g, TabView t, DensityArgs da, double *n, double *partials) {
    __shared__ double sh[4];
    __shared__ WaveStack stacks[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    WaveStack &S = stacks[wv];
    double acc = 0., cim = 0., cra = 0., cev = 0.;
    const int G = (int)g.G;
    const double tau = 1. / da.T, beta0 = -da.degeneracy / pow(2. * kPI, 1.5), beta1 = pow(da.alpha * tau, -1.5);
    const double maxexp = -log10(da.rtpna);
    for (int i = blockIdx.x * 4 + wv; i < G; i += gridDim.x * 4) {
        double U = da.V[i] - da.mu;
        double ar = 0.5 * pow(da.alpha * 1. * da.grad2[i], 1. / 3.);
        double beta2 = -da.lap[i] / (12. * sqrt(da.alpha * tau));
        double res;
        if (ar < fabs(da.thr)) {
            double arg = -tau * U;
            res = beta0 * (beta1 * polylog_j(t, 0, arg) + beta2 * polylog_j(t, 2, arg));
        } else {
            PointArgs q;
            q.invT = 1. / da.T;
            q.U = U;
            q.ar = ar;
            q.pref = beta0 * ar * tau * beta1;
            q.pref2 = beta2 / beta1;
            // lower limit: the first x = -2^k (k = 1, 2, ...) where the
            // integrand is below 1e-12, stopping once x < -1e8 (k = 27)
            bool small = true;
            if (lane < 27) {
                double x = -ldexp(1., lane + 1), arg = -tau * (U - x * ar);
                double l1 = polylog_j(t, 1, arg), l3 = polylog_j(t, 3, arg);
                double test = beta0 * curly_a(t, x) * ar * tau * (beta1 * l1 + beta2 * l3);
                small = !(fabs(test) > kMP);
            }
            unsigned long long bal = __ballot(small ? 1 : 0);
            int k = __ffsll((long long)bal) - 1;  // lowest lane with a small integrand
            if (k < 0 || k > 26) k = 26;
            double x = -ldexp(1., k + 1);
            Counts cnt{0., 0., 27.};
            double rem = adaptive_boole_wave(t, q, da.BP, da.maxIter, da.rtpna, da.compare, maxexp, 2. * x, x, cnt, S,
                                             lane);
            res = rem;
            res += adaptive_boole_wave(t, q, da.BP, da.maxIter, da.rtpna, da.compare, maxexp, x, 100., cnt, S, lane);
            cim += cnt.itermax;
            cra += cnt.relacc;
            cev += cnt.evals;
        }
        if (res != res) res = 0.;
        if (lane == 0) {
            n[i] = res;
            acc += trap_w(g, i) * res;
        }
    }
    if (lane != 0) {
        cim = 0.;
        cra = 0.;
        cev = 0.;
    }
    double r0 = bsum(acc, sh);
    double r1 = bsum(cim, sh);
    double r2 = bsum(cra, sh);
    double r3 = bsum(cev, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = r0;
        partials[mk::kRedBlocks + blockIdx.x] = r1;
        partials[2 * mk::kRedBlocks + blockIdx.x] = r2;
        partials[3 * mk::kRedBlocks + blockIdx.x] = r3;
    }
}

// ---------------------------------------------------------------------------
// Derivatives of the computed density for the adjoint: the same quadrature
// tree (lower limit, panels and acceptance tests of the value integral), with
// the integrand's partial derivatives summed over the accepted panels, i.e.
// the exact derivative of the computed n(U, a_r, beta2) wherever the tree is
// locally constant.
// ---------------------------------------------------------------------------

// quintic Hermite basis derivatives d/dt (see qherm)
__device__ __forceinline__ double qherm_d(double t, double h, double f0, double g0, double c0, double f1, double g1,
                                          double c1) {
    double t2 = t * t, t3 = t2 * t, t4 = t3 * t;
    double H0 = -30. * t2 + 60. * t3 - 30. * t4, H1 = 1. - 18. * t2 + 32. * t3 - 15. * t4;
    double H2 = t - 4.5 * t2 + 6. * t3 - 2.5 * t4, H3 = 30. * t2 - 60. * t3 + 30. * t4;
    double H4 = -12. * t2 + 28. * t3 - 15. * t4, H5 = 1.5 * t2 - 4. * t3 + 2.5 * t4;
    return f0 * H0 + h * g0 * H1 + h * h * c0 * H2 + f1 * H3 + h * g1 * H4 + h * h * c1 * H5;
}

// d/dx of polylog_j(t, j, x) as computed, branch by branch
__device__ double polylog_j_d(const TabView &t, int j, double x) {
    const double sj = 1.5 - (double)j;
    if (x < -699.) return 0.;
    if (x > 699.) return -sj * xpow(j, x) / x * t.ginv[j];
    if (x < t.px0) {
        double e = exp(x);
        return -e + 2. * e * e * t.pw2[j] - 3. * e * e * e * t.pw3[j];
    }
    if (x > t.pxmax) {
        double x2 = 1. / (x * x), acc = 0., dacc = 0.;
        for (int k = 11; k >= 0; k--) {
            acc = x2 * (t.asym[12 * j + k] + acc);
            dacc = x2 * ((double)(k + 1) * t.asym[12 * j + k] + dacc);
        }
        // S = sum_k a_k x^-2(k+1), S' = -2 sum_k (k+1) a_k x^-2(k+1) / x
        return -t.ginv[j] * (sj * xpow(j, x) / x * (1. + acc) + xpow(j, x) * (-2. * dacc / x));
    }
    int k = (int)floor((x - t.px0) * t.pinv);
    if (k > t.pn - 2) k = t.pn - 2;
    double u = (x - (t.px0 + t.ph * (double)k)) * t.pinv;
    const double *a = t.p + 6 * k, *b = a + 6;
    return qherm_d(u, t.ph, a[j], a[j + 1], a[j + 2], b[j], b[j + 1], b[j + 2]) * t.pinv;
}

// polylog_pair and the x-derivatives of its two values, branch by branch
__device__ void polylog_pair_d(const TabView &t, double x, double &l1, double &l3, double &d1, double &d3) {
    if (x < -699.) {
        l1 = l3 = d1 = d3 = 0.;
        return;
    }
    if (x < t.px0 || x > t.pxmax) {
        polylog_pair(t, x, l1, l3);
        d1 = polylog_j_d(t, 1, x);
        d3 = polylog_j_d(t, 3, x);
        return;
    }
    polylog_pair(t, x, l1, l3);
    int k = (int)floor((x - t.px0) * t.pinv);
    if (k > t.pn - 2) k = t.pn - 2;
    double u = (x - (t.px0 + t.ph * (double)k)) * t.pinv;
    const double *a = t.p + 6 * k, *b = a + 6;
    d1 = qherm_d(u, t.ph, a[1], a[2], a[3], b[1], b[2], b[3]) * t.pinv;
    d3 = qherm_d(u, t.ph, a[3], a[4], a[5], b[3], b[4], b[5]) * t.pinv;
}

struct Deriv4 {
    double v, u, a, b;  // value and d/dU, d/da_r, d/dbeta2
};

__device__ __forceinline__ Deriv4 integrand_d(const TabView &t, const PointArgs &q, double beta1, double x) {
    double arg = -(q.U - x * q.ar) * q.invT;
    double l1, l3, d1, d3;
    polylog_pair_d(t, arg, l1, l3, d1, d3);
    double base = q.pref * curly_a(t, x);
    double ds = d1 + q.pref2 * d3;
    Deriv4 r;
    r.v = base * (l1 + q.pref2 * l3);
    r.u = -base * ds * q.invT;
    r.a = r.v / q.ar + base * ds * x * q.invT;  // pref is linear in a_r
    r.b = base * l3 / beta1;                    // pref2 = beta2 / beta1
    return r;
}

__device__ Deriv4 adaptive_boole_wave_d(const TabView &t, const PointArgs &q, double beta1, int BP, int maxIter,
                                        double rtpna, double compare, double maxexp, double a0, double b0,
                                        WaveStack &S, int lane) {
    int sp = 0;
    S.a[0] = a0;
    S.b[0] = b0;
    S.prev[0] = 0.;
    S.it[0] = 0;
    sp = 1;
    __builtin_amdgcn_wave_barrier();
    const int n = 4 * BP + 1, tot = 2 * n;
    Deriv4 total{0., 0., 0., 0.};
    while (sp > 0) {
        sp--;
        double a = S.a[sp], b = S.b[sp], prev = S.prev[sp];
        int iter = S.it[sp];
        __builtin_amdgcn_wave_barrier();
        double half = 0.5 * (b - a), Delta = 0.25 * half / ((double)BP);
        Deriv4 sa{0., 0., 0., 0.}, sb{0., 0., 0., 0.};
        for (int base = 0; base < tot; base += 64) {
            int idx = base + lane;
            if (idx < tot) {
                bool right = idx >= n;
                int yi = right ? idx - n : idx;
                double x = a + (double)yi * Delta;
                if (right) x = x + half;
                double w = boole_w(yi, n);
                Deriv4 f = integrand_d(t, q, beta1, x);
                Deriv4 &acc = right ? sb : sa;
                acc.v += w * f.v;
                acc.u += w * f.u;
                acc.a += w * f.a;
                acc.b += w * f.b;
            }
        }
        sa.v = wave_sum(sa.v);
        sb.v = wave_sum(sb.v);
        sa.u = wave_sum(sa.u);
        sb.u = wave_sum(sb.u);
        sa.a = wave_sum(sa.a);
        sb.a = wave_sum(sb.a);
        sa.b = wave_sum(sa.b);
        sb.b = wave_sum(sb.b);
        double ha = (a + half - a) / (((double)BP) * 4.), hb = (b - (a + half)) / (((double)BP) * 4.);
        double ca = 2. * ha / 45., cb = 2. * hb / 45.;
        double Inta = sa.v * ca, Intb = sb.v * cb;
        double Int = Inta + Intb;
        bool accept;
        if (Int != Int) accept = true;
        else if (iter > 1 && fabs(Int) > 0. && fabs((Int - prev) / Int) < rtpna) accept = true;
        else if (iter > maxIter) accept = true;
        else if (negligible(compare, pow(2., (double)iter) * Int, maxexp)) accept = true;
        else if (Delta < kMP) accept = true;
        else accept = false;
        if (accept) {
            total.v += Int;
            total.u += sa.u * ca + sb.u * cb;
            total.a += sa.a * ca + sb.a * cb;
            total.b += sa.b * ca + sb.b * cb;
        } else {
            if (sp + 2 > 64) return Deriv4{__builtin_nan(""), 0., 0., 0.};
            S.a[sp] = a + half;
            S.b[sp] = b;
            S.prev[sp] = Intb;
            S.it[sp] = iter + 1;
            S.a[sp + 1] = a;
            S.b[sp + 1] = a + half;
            S.prev[sp + 1] = Inta;
            S.it[sp + 1] = iter + 1;
            sp += 2;
            __builtin_amdgcn_wave_barrier();
        }
    }
    return total;
}

// One wavefront per grid point: n and its partial derivatives with respect to
// U = V - mu, |grad V|^2 and lap V (through a_r and beta2).
__global__ __launch_bounds__(256) void k_nait_partials(mk::Grid g, TabView t, DensityArgs da, double *n,
                                                       double *dU, double *dg2, double *dl) {
    __shared__ WaveStack stacks[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    WaveStack &S = stacks[wv];
    const int G = (int)g.G;
    const double tau = 1. / da.T, beta0 = -da.degeneracy / pow(2. * kPI, 1.5), beta1 = pow(da.alpha * tau, -1.5);
    const double sqrt_at = sqrt(da.alpha * tau);
    const double maxexp = -log10(da.rtpna);
    for (int i = blockIdx.x * 4 + wv; i < G; i += gridDim.x * 4) {
        double U = da.V[i] - da.mu;
        double g2 = da.grad2[i];
        double ar = 0.5 * pow(da.alpha * 1. * g2, 1. / 3.);
        double beta2 = -da.lap[i] / (12. * sqrt_at);
        Deriv4 r;
        if (ar < fabs(da.thr)) {
            double arg = -tau * U;
            double P0 = polylog_j(t, 0, arg), P2 = polylog_j(t, 2, arg);
            r.v = beta0 * (beta1 * P0 + beta2 * P2);
            r.u = beta0 * (beta1 * polylog_j_d(t, 0, arg) + beta2 * polylog_j_d(t, 2, arg)) * (-tau);
            r.a = 0.;
            r.b = beta0 * P2;
        } else {
            PointArgs q;
            q.invT = 1. / da.T;
            q.U = U;
            q.ar = ar;
            q.pref = beta0 * ar * tau * beta1;
            q.pref2 = beta2 / beta1;
            bool small = true;
            if (lane < 27) {
                double x = -ldexp(1., lane + 1), arg = -tau * (U - x * ar);
                double l1 = polylog_j(t, 1, arg), l3 = polylog_j(t, 3, arg);
                double test = beta0 * curly_a(t, x) * ar * tau * (beta1 * l1 + beta2 * l3);
                small = !(fabs(test) > kMP);
            }
            unsigned long long bal = __ballot(small ? 1 : 0);
            int k = __ffsll((long long)bal) - 1;
            if (k < 0 || k > 26) k = 26;
            double x = -ldexp(1., k + 1);
            Deriv4 r1 = adaptive_boole_wave_d(t, q, beta1, da.BP, da.maxIter, da.rtpna, da.compare, maxexp, 2. * x,
                                              x, S, lane);
            Deriv4 r2 = adaptive_boole_wave_d(t, q, beta1, da.BP, da.maxIter, da.rtpna, da.compare, maxexp, x, 100.,
                                              S, lane);
            r.v = r1.v + r2.v;
            r.u = r1.u + r2.u;
            r.a = r1.a + r2.a;
            r.b = r1.b + r2.b;
        }
        if (r.v != r.v) r = Deriv4{0., 0., 0., 0.};
        if (lane == 0) {
            n[i] = r.v;
            dU[i] = r.u;
            // a_r = (alpha g2)^{1/3} / 2: da_r/dg2 = a_r / (3 g2); beta2 = -lap / (12 sqrt(alpha tau))
            dg2[i] = (g2 > 0. && r.a != 0.) ? r.a * ar / (3. * g2) : 0.;
            dl[i] = -r.b / (12. * sqrt_at);
        }
    }
}

// ---------------------------------------------------------------------------
// Panel-reuse variant of the same quadrature.  A node [a, b] is sampled at
// k = 0..8BP (left half a + k Delta, right half (a + k' Delta) + half with
// k = 4BP + k'); the shared midpoint is evaluated once.  When a node is
// bisected, each child's even samples are its parent's half, so a child
// evaluates only its 4BP odd samples.  The accepted panels, the acceptance
// tests and their order are those of AdaptiveBoole_nAiT (:3553-3578); a reused
// abscissa can differ from the reference's recomputed one in the last bit.
// One wavefront per grid point; LDS holds the node's samples and, per tree
// level, the sample half of the pending right child.
// ---------------------------------------------------------------------------
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double readlane_d(double v, int l) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double uniform_d(double v) {
    int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes: quad, half-row and row butterflies (DPP), then the
// four row sums combined in a fixed order; the result is wave-uniform.
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_mov<0x141>(v);  // row_half_mirror
    v += dpp_mov<0x140>(v);  // row_mirror
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

struct PointScalars {
    const double *V, *grad2, *lap;
    double mu, invT, tau, beta0, beta1, sqrt_at, alpha, thr, rtpna, compare, log10cmp, maxexp;
    double negl_lo, negl_hi;  // |2^iter Int| below / above: NegligibleMagnitudeQ decided without log10
    int BP, maxIter, depth;
    long long G;
};

constexpr int kReuseMaxPer = 4;  // 4BP + 1 <= 256 samples per half

// One workgroup per grid point, dispatched round-robin over the 8 XCDs
// (workgroup w runs on XCD w mod 8).  The 16 points of one 128-byte line of a
// field go to the same XCD, lines alternate between XCDs: the 8-byte
// per-point reads and writes then fill whole L2 lines once instead of pulling
// each line into every L2, and the work stays interleaved across the chip.
constexpr int kXcds = 8, kLinePts = 16;
__device__ __forceinline__ long long xcd_point(long long) {
    const long long w = blockIdx.x, j = w / kXcds;
    return ((j / kLinePts) * kXcds + w % kXcds) * kLinePts + j % kLinePts;
}
inline unsigned xcd_blocks(long long G) {
    long long lines = (G + kLinePts - 1) / kLinePts, per = (lines + kXcds - 1) / kXcds;
    return (unsigned)(per * kXcds * kLinePts);
}

// Boole weight of sample k of a node with 8BP + 1 samples, in its left (k <=
// 4BP) and right (k >= 4BP) half; 0 outside the half.
__device__ __forceinline__ double wleft(int k, int h4) { return k <= h4 ? boole_w(k, h4 + 1) : 0.; }
__device__ __forceinline__ double wright(int k, int h4) { return k >= h4 ? boole_w(k - h4, h4 + 1) : 0.; }

// The exact forms of the two acceptance tests near their thresholds, out of
// line: their constants then stay out of the walk's registers.
__device__ __noinline__ bool negl_exact(const double *consts, double sm) {
    return consts[0] - log10(fabs(sm)) > consts[1];  // log10(compare) - log10|sm| > -log10(rtpna)
}
__device__ __noinline__ bool relacc_exact(double Int, double prev, double rtpna) {
    return fabs((Int - prev) / Int) < rtpna;
}

// fabs((Int - prev) / Int) < rtpna, decided by a product unless it lies
// within 1e-12 of the threshold (then by the division itself)
__device__ __forceinline__ bool relacc_ok(double Int, double prev, double rtpna) {
    double d = fabs(Int - prev), m = fabs(Int) * rtpna;
    if (d < m * (1. - 1e-12)) return true;
    if (d > m * (1. + 1e-12)) return false;
    return relacc_exact(Int, prev, rtpna);
}

// Four waves per SIMD (at most 128 VGPRs): the walk is latency bound, and the
// few spills this costs are cheaper than the lost occupancy.
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nait_reuse(
    ReuseTab rt, const TabView *__restrict__ tv, PointScalars ps, double *__restrict__ n, int *__restrict__ counts) {
    extern __shared__ double lds[];
    const TabView &t = *tv;  // chain tables and constants for the lower limit and the a_r = 0 form
    const int lane = threadIdx.x;
    const long long i = xcd_point(ps.G);
    if (i >= ps.G) return;
    const int BP = ps.BP, h4 = 4 * BP, full = 8 * BP;
    const int R = (h4 + 64) / 64;                 // rounds over the 4BP + 1 samples of a half
    double *cur = lds;                            // [8BP + 1] samples of the current node
    double *stk = lds + full + 1;                 // [depth][4BP + 1] pending right halves
    double *stn = stk + (size_t)ps.depth * (h4 + 1);  // [depth][4]: a, b, prev, iter
    double *cst = stn + 4 * (size_t)ps.depth;         // [2]: log10(compare), -log10(rtpna)
    if (lane == 0) {
        cst[0] = ps.log10cmp;
        cst[1] = ps.maxexp;
    }

    const double U = ps.V[i] - ps.mu;
    const double ar = 0.5 * pow(ps.alpha * 1. * ps.grad2[i], 1. / 3.);
    const double beta2 = -ps.lap[i] / (12. * ps.sqrt_at);
    double res = 0.;
    int evals = 0, nim = 0, nra = 0;
    if (ar < fabs(ps.thr)) {
        double arg = -ps.tau * U;
        res = ps.beta0 * (ps.beta1 * polylog_j(t, 0, arg) + beta2 * polylog_j(t, 2, arg));
    } else {
        PointArgs q;
        q.invT = ps.invT;
        q.U = U;
        q.ar = ar;
        q.pref = ps.beta0 * ar * ps.tau * ps.beta1;
        q.pref2 = beta2 / ps.beta1;
        // lower limit: the first x = -2^k (k = 1, 2, ...) where the integrand
        // is below 1e-12, stopping once x < -1e8 (k = 27)
        bool small = true;
        if (lane < 27) {
            double x = -ldexp(1., lane + 1), arg = -ps.tau * (U - x * ar);
            double l1, l3;
            polylog_pair(t, arg, l1, l3);  // the values of polylog_j(1, .) and polylog_j(3, .)
            double test = ps.beta0 * curly_a(t, x) * ar * ps.tau * (ps.beta1 * l1 + beta2 * l3);
            small = !(fabs(test) > kMP);
        }
        unsigned long long bal = __ballot(small ? 1 : 0);
        int kx = __ffsll((long long)bal) - 1;
        if (kx < 0 || kx > 26) kx = 26;
        const double xlow = -ldexp(1., kx + 1);
        evals = 27;
#pragma unroll 1
        for (int root = 0; root < 2; root++) {
            double a = root ? xlow : 2. * xlow, b = root ? 100. : xlow, prev = 0.;
            // per-root constants: at level iter an exactly bisected node has
            // half = half0 2^-iter, Delta = Delta0 2^-iter, 2h/45 = c0 2^-iter
            const double half0 = 0.5 * (b - a), Delta0 = 0.25 * half0 / ((double)BP);
            const double c0 = 2. * (((a + half0) - a) / (((double)BP) * 4.)) / 45.;
            int iter = 0, sp = 0;
            bool fresh = true;  // the root evaluates all 8BP + 1 samples
            double sa = 0., sb = 0.;
            for (;;) {
                // the walk is wave-uniform; say so to the compiler (scalar registers, scalar branches)
                a = uniform_d(a);
                b = uniform_d(b);
                prev = uniform_d(prev);
                iter = __builtin_amdgcn_readfirstlane(iter);
                sp = __builtin_amdgcn_readfirstlane(sp);
                const double half = 0.5 * (b - a);
                double Delta, ca, cb;
                if (half == ldexp(half0, -iter) && (a + half) - a == half && b - (a + half) == half) {
                    Delta = ldexp(Delta0, -iter);
                    ca = cb = ldexp(c0, -iter);
                } else {
                    Delta = 0.25 * half / ((double)BP);
                    ca = 2. * (((a + half) - a) / (((double)BP) * 4.)) / 45.;
                    cb = 2. * ((b - (a + half)) / (((double)BP) * 4.)) / 45.;
                }
                // samples to evaluate: all (root) or the 4BP odd ones (the even
                // ones and their Boole sums were placed by the parent step)
                const int nslots = fresh ? full + 1 : h4;
                for (int sl = lane; sl < nslots; sl += 64) {
                    int k = fresh ? sl : 2 * sl + 1;
                    double x = k <= h4 ? a + (double)k * Delta : (a + (double)(k - h4) * Delta) + half;
                    double f = integrand_fast(rt, q, x);
                    cur[k] = f;
                    double wl, wr;
                    if (fresh) {
                        wl = wleft(k, h4);
                        wr = wright(k, h4);
                    } else {  // odd samples: weight 32 in the half that holds them
                        wl = k < h4 ? 32. : 0.;
                        wr = k < h4 ? 0. : 32.;
                    }
                    sa += wl * f;
                    sb += wr * f;
                }
                evals += nslots;
                fresh = false;
                const double Sa = wave_total(sa), Sb = wave_total(sb);
                const double Inta = Sa * ca, Intb = Sb * cb, Int = Inta + Intb;
                bool accept = true;
                if (Int != Int) {  // (Int, Inta, Intb are wave-uniform: wave_total)
                } else if (iter > 1 && fabs(Int) > 0. && relacc_ok(Int, prev, ps.rtpna)) {
                    nra++;
                } else if (iter > ps.maxIter) {
                    nim++;
                } else {
                    double sm = ldexp(Int, iter);  // 2^iter Int (NegligibleMagnitudeQ)
                    bool negl;
                    if (sm == 0.) negl = true;
                    else if (ps.compare == 0. || ps.compare < sm) negl = false;
                    else if (fabs(sm) < ps.negl_lo) negl = true;
                    else if (fabs(sm) > ps.negl_hi) negl = false;
                    else negl = negl_exact(cst, sm);
                    accept = negl || Delta < kMP;
                }
                accept = __builtin_amdgcn_readfirstlane((int)accept) != 0;
                __syncthreads();  // cur complete
                const double *src;
                if (accept) {
                    res += Int;
                    if (sp == 0) break;
                    sp--;
                    a = uniform_d(stn[4 * sp]);
                    b = uniform_d(stn[4 * sp + 1]);
                    prev = uniform_d(stn[4 * sp + 2]);
                    iter = (int)uniform_d(stn[4 * sp + 3]);
                    src = stk + (size_t)sp * (h4 + 1);
                } else {
                    if (sp >= ps.depth) {  // deeper than maxIter allows; cannot happen
                        res = __builtin_nan("");
                        break;
                    }
                    // the right child waits on the stack with the right half of cur
                    double *dst = stk + (size_t)sp * (h4 + 1);
                    for (int j = lane; j <= h4; j += 64) dst[j] = cur[h4 + j];
                    if (lane == 0) {
                        stn[4 * sp] = a + half;
                        stn[4 * sp + 1] = b;
                        stn[4 * sp + 2] = Intb;
                        stn[4 * sp + 3] = (double)(iter + 1);
                    }
                    sp++;
                    b = a + half;
                    prev = Inta;
                    iter++;
                    src = cur;  // the left child: the left half of cur, in place
                }
                // the next node's even samples k = 2j are src[j]; their Boole
                // sums are formed here, the odd samples are added above
                double v[kReuseMaxPer];
                sa = 0.;
                sb = 0.;
#pragma unroll
                for (int r = 0; r < kReuseMaxPer; r++) {
                    int j = lane + 64 * r;
                    v[r] = (r < R && j <= h4) ? src[j] : 0.;
                }
                __syncthreads();
#pragma unroll
                for (int r = 0; r < kReuseMaxPer; r++) {
                    int j = lane + 64 * r;
                    if (r < R && j <= h4) {
                        cur[2 * j] = v[r];
                        sa += wleft(2 * j, h4) * v[r];
                        sb += wright(2 * j, h4) * v[r];
                    }
                }
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        const long long G = ps.G;
        n[i] = res != res ? 0. : res;
        counts[i] = evals;
        counts[G + i] = nim;
        counts[2 * G + i] = nra;
    }
}

// The same walk with a fixed lane layout, for BP <= 15 (2BP + 1 <= 32): lanes
// 0..31 hold the left half of a node, lanes 32..63 the right half.  Lane m of a
// half holds that half's even sample 2m (m <= 2BP) and odd sample 2m + 1 (m <
// 2BP); the midpoint is the last even sample of the left half and the first of
// the right one.  A lane's Boole weights are then constants of the lane, and
// one row-sum reduction of w_e e + 32 o yields both half sums (rows 0-1: left,
// rows 2-3: right).  Abscissae, panels and acceptance tests are those of
// k_nait_reuse; only the order of the panel-sum additions differs.
constexpr int kLanesMaxBP = 15;

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_nait_lanes(
    ReuseTab rt, const TabView *__restrict__ tv, PointScalars ps, double *__restrict__ n, int *__restrict__ counts) {
    extern __shared__ double lds[];
    const TabView &t = *tv;
    const int lane = threadIdx.x;
    const long long i = xcd_point(ps.G);
    if (i >= ps.G) return;
    const int BP = ps.BP, h4 = 4 * BP, full = 8 * BP, H = 2 * BP;
    double *cur = lds;                                // [8BP + 1] samples of the current node
    double *stk = lds + full + 1;                     // [depth][4BP + 1] pending right halves
    double *stn = stk + (size_t)ps.depth * (h4 + 1);  // [depth][4]: a, b, prev, iter
    double *cst = stn + 4 * (size_t)ps.depth;         // [2]: log10(compare), -log10(rtpna)
    if (lane == 0) {
        cst[0] = ps.log10cmp;
        cst[1] = ps.maxexp;
    }
    // this lane's slots: half hs, index m within the half
    const int hs = lane >> 5, m = lane & 31;
    const bool has_even = m <= H, has_odd = m < H;
    const int ke = hs * h4 + 2 * m, ko = ke + 1;  // sample indices in the node (0..8BP)
    const int je = hs * H + m;                    // where the next node's even sample 2 je comes from
    const double we = !has_even ? 0. : ((m == 0 || m == H) ? 7. : (((2 * m - 1) & 3) == 1 ? 12. : 14.));
    const double wo = has_odd ? 32. : 0.;

    const double U = ps.V[i] - ps.mu;
    const double ar = 0.5 * pow(ps.alpha * 1. * ps.grad2[i], 1. / 3.);
    const double beta2 = -ps.lap[i] / (12. * ps.sqrt_at);
    double res = 0.;
    int evals = 0, nim = 0, nra = 0;
    if (ar < fabs(ps.thr)) {
        double arg = -ps.tau * U;
        res = ps.beta0 * (ps.beta1 * polylog_j(t, 0, arg) + beta2 * polylog_j(t, 2, arg));
    } else {
        PointArgs q;
        q.invT = ps.invT;
        q.U = U;
        q.ar = ar;
        q.pref = ps.beta0 * ar * ps.tau * ps.beta1;
        q.pref2 = beta2 / ps.beta1;
        // lower limit: the first x = -2^k (k = 1, 2, ...) where the integrand
        // is below 1e-12, stopping once x < -1e8 (k = 27)
        // (the table forms of the walk decide; within 1e-12 of the threshold, where
        // they and the interpolants of the lower-limit test could disagree, the latter)
        bool small = true;
        if (lane < 27) {
            double x = -ldexp(1., lane + 1), arg = -ps.tau * (U - x * ar);
            double l1, l3;
            polylog_pair_fast(rt, arg, l1, l3);
            double test = ps.beta0 * curly_a_fast(rt, x) * ar * ps.tau * (ps.beta1 * l1 + beta2 * l3);
            const double at = fabs(test);
            if (at > kMP * (1. - 1e-12) && at < kMP * (1. + 1e-12)) {
                polylog_pair(t, arg, l1, l3);
                test = ps.beta0 * curly_a(t, x) * ar * ps.tau * (ps.beta1 * l1 + beta2 * l3);
            }
            small = !(fabs(test) > kMP);
        }
        unsigned long long bal = __ballot(small ? 1 : 0);
        int kx = __ffsll((long long)bal) - 1;
        if (kx < 0 || kx > 26) kx = 26;
        const double xlow = -ldexp(1., kx + 1);
        evals = 27;
#pragma unroll 1
        for (int root = 0; root < 2; root++) {
            double a = root ? xlow : 2. * xlow, b = root ? 100. : xlow, prev = 0.;
            const double half0 = 0.5 * (b - a), Delta0 = 0.25 * half0 / ((double)BP);
            const double c0 = 2. * (((a + half0) - a) / (((double)BP) * 4.)) / 45.;
            int iter = 0, sp = 0;
            bool fresh = true;  // the root evaluates its even samples too
            double e = 0.;      // this lane's even sample (placed by the parent step)
            for (;;) {
                // a, b, prev are wave-uniform by construction and only feed VALU arithmetic (FP64 has no
                // scalar ALU), so they stay in vector registers; iter and sp drive scalar branches
                iter = __builtin_amdgcn_readfirstlane(iter);
                sp = __builtin_amdgcn_readfirstlane(sp);
                const double half = 0.5 * (b - a);
                double Delta, ca, cb;
                if (half == ldexp(half0, -iter) && (a + half) - a == half && b - (a + half) == half) {
                    Delta = ldexp(Delta0, -iter);
                    ca = cb = ldexp(c0, -iter);
                } else {
                    Delta = 0.25 * half / ((double)BP);
                    ca = 2. * (((a + half) - a) / (((double)BP) * 4.)) / 45.;
                    cb = 2. * ((b - (a + half)) / (((double)BP) * 4.)) / 45.;
                }
                if (fresh) {
                    if (has_even) {
                        double x = ke <= h4 ? a + (double)ke * Delta : (a + (double)(ke - h4) * Delta) + half;
                        e = integrand_fast(rt, q, x);
                        cur[ke] = e;
                    }
                    evals += full + 1;
                } else {
                    evals += h4;
                }
                fresh = false;
                double o = 0.;
                if (has_odd) {
                    double x = ko <= h4 ? a + (double)ko * Delta : (a + (double)(ko - h4) * Delta) + half;
                    o = integrand_fast(rt, q, x);
                    cur[ko] = o;
                }
                double v = we * e + wo * o;
                v += dpp_mov<0xB1>(v);   // quad_perm [1,0,3,2]
                v += dpp_mov<0x4E>(v);   // quad_perm [2,3,0,1]
                v += dpp_mov<0x141>(v);  // row_half_mirror
                v += dpp_mov<0x140>(v);  // row_mirror: every lane holds its row's sum
                const double Sa = readlane_d(v, 0) + readlane_d(v, 16), Sb = readlane_d(v, 32) + readlane_d(v, 48);
                const double Inta = Sa * ca, Intb = Sb * cb, Int = Inta + Intb;
                bool accept = true;
                if (Int != Int) {
                } else if (iter > 1 && fabs(Int) > 0. && relacc_ok(Int, prev, ps.rtpna)) {
                    nra++;
                } else if (iter > ps.maxIter) {
                    nim++;
                } else {
                    double sm = ldexp(Int, iter);  // 2^iter Int (NegligibleMagnitudeQ)
                    bool negl;
                    if (sm == 0.) negl = true;
                    else if (ps.compare == 0. || ps.compare < sm) negl = false;
                    else if (fabs(sm) < ps.negl_lo) negl = true;
                    else if (fabs(sm) > ps.negl_hi) negl = false;
                    else negl = negl_exact(cst, sm);
                    accept = negl || Delta < kMP;
                }
                accept = __builtin_amdgcn_readfirstlane((int)accept) != 0;
                __syncthreads();  // cur complete
                const double *src;
                if (accept) {
                    res += Int;
                    if (sp == 0) break;
                    sp--;
                    a = stn[4 * sp];
                    b = stn[4 * sp + 1];
                    prev = stn[4 * sp + 2];
                    iter = (int)stn[4 * sp + 3];
                    src = stk + (size_t)sp * (h4 + 1);
                } else {
                    if (sp >= ps.depth) {  // deeper than maxIter allows; cannot happen
                        res = __builtin_nan("");
                        break;
                    }
                    double *dst = stk + (size_t)sp * (h4 + 1);
                    if (lane <= h4) dst[lane] = cur[h4 + lane];  // 4BP + 1 <= 61 samples
                    if (lane == 0) {
                        stn[4 * sp] = a + half;
                        stn[4 * sp + 1] = b;
                        stn[4 * sp + 2] = Intb;
                        stn[4 * sp + 3] = (double)(iter + 1);
                    }
                    sp++;
                    b = a + half;
                    prev = Inta;
                    iter++;
                    src = cur;  // the left child: the left half of cur, in place
                }
                // the next node's even sample 2 je is src[je]
                e = has_even ? src[je] : 0.;
                __syncthreads();
                if (has_even) cur[ke] = e;
            }
            __syncthreads();
        }
    }
    if (lane == 0) {
        const long long G = ps.G;
        n[i] = res != res ? 0. : res;
        counts[i] = evals;
        counts[G + i] = nim;
        counts[2 * G + i] = nra;
    }
}

// Trapezoid-weighted density sum and the three per-point counters, as
// per-block partials (fixed order).
__global__ __launch_bounds__(256) void k_nait_sums(mk::Grid g, const double *__restrict__ n,
                                                  const int *__restrict__ counts, double *partials) {
    __shared__ double sh[4];
    double acc = 0., ev = 0., im = 0., ra = 0.;
    const long long G = g.G;
    for (long long i = blockIdx.x * 256 + threadIdx.x; i < G; i += (long long)gridDim.x * 256) {
        acc += trap_w(g, (int)i) * n[i];
        ev += (double)counts[i];
        im += (double)counts[G + i];
        ra += (double)counts[2 * G + i];
    }
    double r0 = bsum(acc, sh), r1 = bsum(im, sh), r2 = bsum(ra, sh), r3 = bsum(ev, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = r0;
        partials[mk::kRedBlocks + blockIdx.x] = r1;
        partials[2 * mk::kRedBlocks + blockIdx.x] = r2;
        partials[3 * mk::kRedBlocks + blockIdx.x] = r3;
    }
}

struct EkinArgs {
    const double *V, *grad2, *lap;
    double mu, thr, c0, c1, c2, internal_acc, cmp, area_pref, neg_exp;
    double tail;  // nodes with y > max(U/a, 0) + tail are left out (infinity: every node)
    int MaxIt;
    int resume;  // continue each point from the refinement the previous launch stopped at
    long long G;
};

// Ai on [-300, 100]: the quintic Hermite interpolants of the (Ai, Ai') table
// (Ai'' = y Ai) in monomial form per interval; the asymptotic form below.
constexpr double kAX0 = -300., kAH = 0.005, kAINV = 200.;
constexpr int kAN = 80001;

__device__ __noinline__ double airy_asym(double y) {
    return cos(2. * (-y) * sqrt(-y) / 3. - 0.25 * kPI) / (sqrt(kPI) * pow(-y, 0.25));
}

// the table part of airy_fast (y > -300; the index is clamped, so any y reads in bounds)
__device__ __forceinline__ double airy_poly(const double2 *__restrict__ ap, double y) {
    int k = (int)floor((y - kAX0) * kAINV);
    k = k < 0 ? 0 : (k > kAN - 2 ? kAN - 2 : k);
    double u = (y - (kAX0 + kAH * (double)k)) * kAINV;
    const double2 *c = ap + 3 * k;
    double2 c01 = c[0], c23 = c[1], c45 = c[2];
    return ((((c45.y * u + c45.x) * u + c23.y) * u + c23.x) * u + c01.y) * u + c01.x;
}

__device__ __forceinline__ double airy_fast(const double2 *__restrict__ ap, double y) {
    if (y <= -300.) return airy_asym(y);
    int k = (int)floor((y - kAX0) * kAINV);
    k = k < 0 ? 0 : (k > kAN - 2 ? kAN - 2 : k);
    double u = (y - (kAX0 + kAH * (double)k)) * kAINV;
    const double2 *c = ap + 3 * k;
    double2 c01 = c[0], c23 = c[1], c45 = c[2];
    return ((((c45.y * u + c45.x) * u + c23.y) * u + c23.x) * u + c01.y) * u + c01.x;
}

// Trapezoid node j of EkinAiT03D's refinement with gs = dg + 1 nodes on [0, X]:
// x = j X / dg, correctly rounded (product by the reciprocal plus one residual
// correction; the residual is exact by FMA).
__device__ __forceinline__ double ekin_node(int j, double X, double dg, double inv) {
    double num = (double)j * X, q = num * inv;
    return fma(fma(-q, dg, num), inv, q);
}

// c1 (2ax)^{5/2}/5 - c2 lap (2ax)^{1/2}/6 times Ai(y); tt = 2ax
__device__ __forceinline__ double ekin_term(double AI, double tt, double k1, double k2) {
    double sq = sqrt(tt);
    return AI * (k1 * tt * tt * sq - k2 * sq);
}

// Node ranges of one refinement.  [0, asym): nodes that may have y = x + U/a <=
// -300, where Ai takes its asymptotic form (two nodes of margin; y is monotone in
// j).  (last, gs): nodes beyond y = max(U/a, 0) + tail.  There |Ai(y)| <
// exp(-(2/3) tail^{3/2}), 8e-38 for tail = 25, of Ai's size where the integrand
// lives, so with the (2ax)^{5/2} growth over x <= 400 a left-out term is below
// 1e-30 of the terms before it and below half an ulp of the lane's partial sum:
// leaving the tail out does not change the sum.  A NaN U/a keeps every node.
struct EkinNodes {
    int asym, last;
};
__device__ __forceinline__ EkinNodes ekin_nodes(int gs, double dg, double X, double Ua, double tail) {
    EkinNodes e;
    const double xa = -300. - Ua;  // y <= -300 for x <= xa (xa < X)
    e.asym = xa >= 0. ? min(gs, (int)(xa / X * dg) + 2) : 0;
    const double xc = fmax(Ua, 0.) + tail - Ua;
    e.last = xc < X ? min(gs - 1, (int)(xc / X * dg) + 1) : gs - 1;
    return e;
}

// Airy kinetic-energy density (EkinAiT03D inner loop, mpDPFT.cpp:5624-5709),
// one wavefront per grid point: the trapezoid samples of each refinement are
// spread over the lanes.  The host raises MaxIt by one per launch within one
// energy evaluation; a point's refinements up to the previous cap are the same
// in every launch, so each launch resumes from the stored last refinement
// (st_m: its index, or -(index + 2) once the convergence test stopped the point at
// that refinement, -1 for the a < thr form; st_v: its
// value) and evaluates only the new one.
__global__ __launch_bounds__(64) void k_ekin_ait(const double2 *__restrict__ ap, EkinArgs ea, double *out,
                                                double *st_v, int *st_m) {
    const int lane = threadIdx.x;
    const long long i = xcd_point(ea.G);
    if (i >= ea.G) return;
    double U = ea.V[i] - ea.mu, a = 0.5 * ea.c0 * pow(fabs(ea.grad2[i]), 1. / 3.);
    double lap = ea.lap[i];
    double val;
    if (a < ea.thr) {
        val = (-U > 0.) ? ea.c1 * pow(2. * (-U), 2.5) / 5. - ea.c2 * lap * sqrt(2. * (-U)) / 6. : 0.;
        if (lane == 0) {
            out[i] = val;
            st_v[i] = val;
            st_m[i] = -1;
        }
        return;
    }
    double prev = 1.;
    int iteration = 0;
    if (ea.resume) {
        int m0 = st_m[i];
        if (m0 < 0) {
            if (lane == 0) out[i] = st_v[i];
            return;
        }
        iteration = m0 + 1;
        prev = st_v[i];
    }
    const double xmax = fmax(1., 100. - U / a);
    const double Ua = U / a, xr = fmax(1., 100. - Ua);
    const double k1 = ea.c1 / 5., k2 = ea.c2 * lap / 6., twoa = 2. * a;
    bool done = false;
    for (;;) {
        const int gs = 64 << iteration;
        const double dg = (double)(gs - 1), inv = 1. / dg;
        const EkinNodes nd = ekin_nodes(gs, dg, xr, Ua, ea.tail);
        double r = 0.;
        // each lane sums the nodes j = lane (mod 64) in increasing order
        int j = lane;
        for (; j < nd.asym; j += 64) {
            double x = ekin_node(j, xr, dg, inv), y = x + Ua;
            double AI = y > 100. ? 0. : airy_fast(ap, y);
            r += ((j == 0 || j == gs - 1) ? 0.5 : 1.) * ekin_term(AI, twoa * x, k1, k2);
        }
        for (; j + 64 <= nd.last; j += 128) {  // two nodes per pass: both table reads in flight
            double x0 = ekin_node(j, xr, dg, inv), y0 = x0 + Ua;
            double x1 = ekin_node(j + 64, xr, dg, inv), y1 = x1 + Ua;
            double A0 = airy_poly(ap, y0), A1 = airy_poly(ap, y1);
            double f0 = ekin_term(y0 > 100. ? 0. : A0, twoa * x0, k1, k2);
            double f1 = ekin_term(y1 > 100. ? 0. : A1, twoa * x1, k1, k2);
            r += (j == 0 ? 0.5 : 1.) * f0;
            r += (j + 64 == gs - 1 ? 0.5 : 1.) * f1;
        }
        if (j <= nd.last) {
            double x = ekin_node(j, xr, dg, inv), y = x + Ua;
            double AI = airy_poly(ap, y);
            r += ((j == 0 || j == gs - 1) ? 0.5 : 1.) * ekin_term(y > 100. ? 0. : AI, twoa * x, k1, k2);
        }
        r = wave_total(r);
        r *= (xmax - 0.) / ((double)gs - 1.);
        val = r;
        double ac = (fabs(val) < kMP && fabs(prev) < kMP) ? 0. : fabs((val - prev) / (val + prev));
        double as = (val == 0. && prev == 0.) ? 0. : (val - prev) / (val + prev);
        if (iteration > 0 &&
            (ac < kMP || (fabs(as) < ea.internal_acc && negligible(ea.cmp, ea.area_pref * val, ea.neg_exp)))) {
            done = true;
            break;
        }
        if (iteration >= ea.MaxIt) break;
        iteration++;
        prev = val;
    }
    if (lane == 0) {
        out[i] = val;
        st_v[i] = val;
        st_m[i] = done ? -(iteration + 2) : iteration;  // < 0: converged at refinement -(st_m + 2)
    }
}

// d Ai / dy of airy_fast as computed (polynomial derivative; asymptotic form below -300)
// Ai and Ai' together (one table lookup): the quintic and its derivative on the
// interval's monomial coefficients; the asymptotic forms below y = -300.
__device__ __forceinline__ void airy_pair(const double2 *__restrict__ ap, double y, double &A, double &Ad) {
    if (y > 100.) {
        A = Ad = 0.;
        return;
    }
    if (y <= -300.) {
        const double z = -y, q = pow(z, 0.25), phi = 2. * z * sqrt(z) / 3. - 0.25 * kPI;
        A = cos(phi) / (sqrt(kPI) * q);
        Ad = (sin(phi) * sqrt(z) / q + cos(phi) * 0.25 / (q * z)) / sqrt(kPI);
        return;
    }
    int k = (int)floor((y - kAX0) * kAINV);
    k = k < 0 ? 0 : (k > kAN - 2 ? kAN - 2 : k);
    const double u = (y - (kAX0 + kAH * (double)k)) * kAINV;
    const double2 *c = ap + 3 * k;
    const double2 c01 = c[0], c23 = c[1], c45 = c[2];
    A = ((((c45.y * u + c45.x) * u + c23.y) * u + c23.x) * u + c01.y) * u + c01.x;
    Ad = ((((5. * c45.y * u + 4. * c45.x) * u + 3. * c23.y) * u + 2. * c23.x) * u + c01.y) * kAINV;
}

// Partial derivatives of the EkinAiT03D energy density (k_ekin_ait with the
// final MaxIt, no resume) with respect to U = V - mu, |grad V|^2 and lap V:
// the refinement that ends the point's loop, differentiated with its node
// count held fixed (the nodes x_j = j X / (gs - 1) move with X = max(1, 100 -
// U / a)).  With st_m (the refinements k_ekin_ait recorded for the same V, mu
// and MaxIt) only that refinement is summed; without it the loop is rerun with
// the value sums to find it.
struct EkinPoint {
    double U, a, Ua, X, xmax, dXdU, dXda, k1, k2, twoa, inva, Ua2;
};
// value and derivative sums of refinement `iteration` of one point (wave totals)
__device__ __forceinline__ void ekin_level(const double2 *__restrict__ ap, const EkinArgs &ea, const EkinPoint &e,
                                           int iteration, bool derivs, double &r, double &rU, double &ra,
                                           double &rl) {
    const int lane = threadIdx.x;
    const int gs = 64 << iteration;
    const double dg = (double)(gs - 1), inv = 1. / dg;
    const int last = ekin_nodes(gs, dg, e.X, e.Ua, ea.tail).last;  // the energy's node range
    r = rU = ra = rl = 0.;
    const double c26 = -(ea.c2 / 6.);
    if (!derivs) {
        for (int j = lane; j <= last; j += 64) {
            const double x = ekin_node(j, e.X, dg, inv), y = x + e.Ua;
            const double wj = (j == 0 || j == gs - 1) ? 0.5 : 1.;
            const double tt = e.twoa * x, sq = sqrt(tt);
            const double h = e.k1 * tt * tt * sq - e.k2 * sq;
            r += wj * (y > 100. ? 0. : airy_fast(ap, y)) * h;
        }
        r = wave_total(r);
        return;
    }
    // two nodes per pass (both table reads in flight); sqrt(tt) and 1/sqrt(tt) from one reciprocal square
    // root with a Newton step (the derivatives need no correctly rounded root)
#pragma unroll 2
    for (int j = lane; j <= last; j += 64) {
        const double x = ekin_node(j, e.X, dg, inv), y = x + e.Ua;
        const double wj = (j == 0 || j == gs - 1) ? 0.5 : 1.;
        const double tt = e.twoa * x;
        double isq = 0., sq = 0.;
        if (tt > 0.) {
            isq = rsqrt(tt);
            isq = isq * (1.5 - 0.5 * tt * isq * isq);
            sq = tt * isq;
        }
        const double h = e.k1 * tt * tt * sq - e.k2 * sq;
        double AI, AId;
        airy_pair(ap, y, AI, AId);
        const double dxU = (double)j * inv * e.dXdU, dxa = (double)j * inv * e.dXda;
        const double dyU = dxU + e.inva, dya = dxa - e.Ua2;
        // dh/dx and dh/da (both vanish with x; the x = 0 node contributes no x-derivative)
        const double t5 = 5. * e.k1 * tt * sq - e.k2 * isq;
        const double hx = e.a * t5, ha = x * t5;
        r += wj * AI * h;
        rU += wj * (AId * dyU * h + AI * hx * dxU);
        ra += wj * (AId * dya * h + AI * (hx * dxa + ha));
        rl += wj * AI * c26 * sq;
    }
    r = wave_total(r);
    rU = wave_total(rU);
    ra = wave_total(ra);
    rl = wave_total(rl);
}

__global__ __launch_bounds__(64) void k_ekin_partials(const double2 *__restrict__ ap, EkinArgs ea,
                                                      const int *__restrict__ st_m, double *eU, double *eg2,
                                                      double *el) {
    const int lane = threadIdx.x;
    const long long i = xcd_point(ea.G);
    if (i >= ea.G) return;
    const double U = ea.V[i] - ea.mu, g2 = fabs(ea.grad2[i]), a = 0.5 * ea.c0 * pow(g2, 1. / 3.);
    const double lap = ea.lap[i];
    if (a < ea.thr) {
        double dU = 0., dl = 0.;
        if (-U > 0.) {
            double t = 2. * (-U), st = sqrt(t);
            dU = -ea.c1 * t * st + ea.c2 * lap / (6. * st);
            dl = -ea.c2 * st / 6.;
        }
        if (lane == 0) {
            eU[i] = dU;
            eg2[i] = 0.;
            el[i] = dl;
        }
        return;
    }
    EkinPoint e;
    e.U = U;
    e.a = a;
    e.Ua = U / a;
    e.X = fmax(1., 100. - e.Ua);
    e.xmax = fmax(1., 100. - U / a);
    const bool Xvar = 100. - e.Ua > 1.;
    e.dXdU = Xvar ? -1. / a : 0.;
    e.dXda = Xvar ? U / (a * a) : 0.;
    e.k1 = ea.c1 / 5.;
    e.k2 = ea.c2 * lap / 6.;
    e.twoa = 2. * a;
    e.inva = 1. / a;
    e.Ua2 = U / (a * a);
    int level = -1;
    if (st_m) {  // the refinement k_ekin_ait stopped this point at
        const int m0 = st_m[i];
        level = m0 >= 0 ? m0 : (m0 <= -2 ? -m0 - 2 : -1);
    }
    if (level < 0) {  // rerun the value sums of the refinements to find it
        double prev = 1.;
        for (int iteration = 0;; iteration++) {
            double r, rU, ra, rl;
            ekin_level(ap, ea, e, iteration, false, r, rU, ra, rl);
            const double val = r * (e.xmax - 0.) / ((double)(64 << iteration) - 1.);
            const double ac = (fabs(val) < kMP && fabs(prev) < kMP) ? 0. : fabs((val - prev) / (val + prev));
            const double as = (val == 0. && prev == 0.) ? 0. : (val - prev) / (val + prev);
            if ((iteration > 0 &&
                 (ac < kMP || (fabs(as) < ea.internal_acc && negligible(ea.cmp, ea.area_pref * val, ea.neg_exp)))) ||
                iteration >= ea.MaxIt) {
                level = iteration;
                break;
            }
            prev = val;
        }
    }
    double r, rU, ra, rl;
    ekin_level(ap, ea, e, level, true, r, rU, ra, rl);
    const double inv = 1. / (double)((64 << level) - 1);
    const double dU = (e.dXdU * r + e.X * rU) * inv;
    const double da = (e.dXda * r + e.X * ra) * inv;
    const double dl = e.X * rl * inv;
    if (lane == 0) {
        eU[i] = dU;
        eg2[i] = g2 > 0. ? da * a / (3. * g2) : 0.;  // a = c0 |g2|^{1/3} / 2
        el[i] = dl;
    }
}

}  // namespace

struct NaitTables {
    TabView view;
    double4 *d_c = nullptr;
    double2 *d_a = nullptr;
    double *d_p = nullptr;
    double *d_asym = nullptr;   // [4][12]
    double2 *d_pp = nullptr;    // monomial form of the Li pair interpolants
    TabView *d_view = nullptr;  // device copy of view
    double2 *d_ap = nullptr;    // Ai in monomial form, [kAN - 1][3]
    double *d_ekin_v = nullptr; // [G] last refinement value per point (EkinAiT03D)
    int *d_ekin_m = nullptr;    // [G] its index, -1 when converged
    bool reuse_ok = false;      // the file's geometry is the one the panel-reuse kernel assumes
    // the last EkinAiT03D evaluation (its d_ekin_m refinements): species, fingerprint of V
    // (<V,V>, <V,Env>), mu and MaxIteration
    int ekin_s = -1, ekin_maxit = -1;
    double ekin_fp[2] = {0., 0.};
    double ekin_mu = 0.;
    int *d_counts = nullptr;    // [3][G] per-point evaluations, iterMax and relAcc returns
    double Mpp[MPDPFT_MAX_SPECIES][3];
    double last_itermax[MPDPFT_MAX_SPECIES], last_relacc[MPDPFT_MAX_SPECIES];
    int maxiter_capped[MPDPFT_MAX_SPECIES];  // the maxIter cap was reported in this solve
};

static std::string default_tables_path() {
    if (const char *e = getenv("MPDPFT_TABLES")) return e;
    Dl_info info;
    if (dladdr((void *)&default_tables_path, &info) && info.dli_fname) {
        std::string p(info.dli_fname);
        size_t k = p.rfind('/');
        return (k == std::string::npos ? std::string(".") : p.substr(0, k)) + "/nait_tables.bin";
    }
    return "nait_tables.bin";
}

// First derivatives of the cubic spline with parabolically terminated ends
// through (x_i, y_i) (tridiagonal system, Thomas algorithm).
static void spline_derivatives(const std::vector<double> &x, const std::vector<double> &y, std::vector<double> &d) {
    size_t n = x.size();
    std::vector<double> lo(n), di(n), up(n), rhs(n);
    di[0] = 1.;
    up[0] = 1.;
    rhs[0] = 2. * (y[1] - y[0]) / (x[1] - x[0]);
    for (size_t i = 1; i + 1 < n; i++) {
        lo[i] = x[i + 1] - x[i];
        di[i] = 2. * (x[i + 1] - x[i - 1]);
        up[i] = x[i] - x[i - 1];
        rhs[i] = 3. * (y[i] - y[i - 1]) / (x[i] - x[i - 1]) * (x[i + 1] - x[i]) +
                 3. * (y[i + 1] - y[i]) / (x[i + 1] - x[i]) * (x[i] - x[i - 1]);
    }
    lo[n - 1] = 1.;
    di[n - 1] = 1.;
    rhs[n - 1] = 2. * (y[n - 1] - y[n - 2]) / (x[n - 1] - x[n - 2]);
    std::vector<double> cp(n), dp(n);
    cp[0] = up[0] / di[0];
    dp[0] = rhs[0] / di[0];
    for (size_t i = 1; i < n; i++) {
        double m = di[i] - lo[i] * cp[i - 1];
        cp[i] = up[i] / m;
        dp[i] = (rhs[i] - lo[i] * dp[i - 1]) / m;
    }
    d.assign(n, 0.);
    d[n - 1] = dp[n - 1];
    for (size_t i = n - 1; i-- > 0;) d[i] = dp[i] - cp[i] * d[i + 1];
}

int nait_init(mpdpft_ctx *c) {
    std::string path = default_tables_path();
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return ctx_fail(c, MPDPFT_EIO, "cannot open nAiT tables " + path);
    char head[16];
    long long n3[3];
    double r[6];
    bool ok = std::fread(head, 1, 16, f) == 16 && std::memcmp(head, "NAITTAB1", 8) == 0;
    for (int k = 0; k < 3 && ok; k++) ok = std::fread(&n3[k], 8, 1, f) == 1 && std::fread(&r[2 * k], 8, 2, f) == 2;
    if (!ok) {
        std::fclose(f);
        return ctx_fail(c, MPDPFT_EIO, "bad nAiT table header " + path);
    }
    int cn = (int)n3[0], an = (int)n3[1], pn = (int)n3[2];
    std::vector<double> cv(cn), av(2 * (size_t)an), pv(6 * (size_t)pn);
    ok = std::fread(cv.data(), 8, cn, f) == (size_t)cn && std::fread(av.data(), 8, av.size(), f) == av.size() &&
         std::fread(pv.data(), 8, pv.size(), f) == pv.size();
    std::fclose(f);
    if (!ok) return ctx_fail(c, MPDPFT_EIO, "truncated nAiT tables " + path);
    std::vector<double> x(cn), d;
    for (int i = 0; i < cn; i++) x[i] = r[0] + r[1] * (double)i;
    spline_derivatives(x, cv, d);
    // cubic coefficients per knot interval (the Hermite form of the spline)
    std::vector<double4> cc(cn);
    double h = r[1];
    for (int i = 0; i + 1 < cn; i++) {
        double y0 = cv[i], y1 = cv[i + 1], d0 = d[i], d1 = d[i + 1];
        cc[i] = double4{y0, d0, (3. * (y1 - y0) - 2. * d0 * h - d1 * h) / (h * h),
                        (2. * (y0 - y1) + d0 * h + d1 * h) / (h * h * h)};
    }
    cc[cn - 1] = double4{cv[cn - 1], d[cn - 1], 0., 0.};
    NaitTables *t = new NaitTables();
    c->nait = t;
    if (hipMalloc((void **)&t->d_c, sizeof(double4) * cn) != hipSuccess ||
        hipMalloc((void **)&t->d_a, sizeof(double) * av.size()) != hipSuccess ||
        hipMalloc((void **)&t->d_p, sizeof(double) * pv.size()) != hipSuccess)
        return ctx_fail(c, MPDPFT_ENOMEM, "nAiT tables");
    if (hipMemcpy(t->d_c, cc.data(), sizeof(double4) * cn, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->d_a, av.data(), sizeof(double) * av.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->d_p, pv.data(), sizeof(double) * pv.size(), hipMemcpyHostToDevice) != hipSuccess)
        return ctx_fail(c, MPDPFT_EHIP, "nAiT table upload");
    TabView &v = t->view;
    v.cc = t->d_c;
    v.a = t->d_a;
    v.p = t->d_p;
    v.cn = cn;
    v.an = an;
    v.pn = pn;
    v.cx0 = r[0];
    v.ch = r[1];
    v.cinv = 1. / r[1];
    v.ax0 = r[2];
    v.ah = r[3];
    v.ainv = 1. / r[3];
    v.px0 = r[4];
    v.ph = r[5];
    v.pinv = 1. / r[5];
    v.pxmax = r[4] + r[5] * (double)(pn - 1);
    const double sq = std::sqrt(kPI);
    const double gam[4] = {0.75 * sq, 0.5 * sq, sq, -2. * sq};  // Gamma(s+1) as in mpDPFT.cpp:112
    for (int j = 0; j < 4; j++) {
        double sj = 1.5 - (double)j;
        v.ginv[j] = 1. / gam[j];
        v.pw2[j] = std::pow(2., -sj);
        v.pw3[j] = std::pow(3., -sj);
    }
    for (int s = 0; s < c->S; s++)
        for (int k = 0; k < 3; k++) t->Mpp[s][k] = c->p.Mpp[k];
    if (hipMalloc((void **)&t->d_asym, sizeof(kAsym.v)) != hipSuccess ||
        hipMalloc((void **)&t->d_counts, sizeof(int) * 3 * c->G) != hipSuccess)
        return ctx_fail(c, MPDPFT_ENOMEM, "nAiT coefficients/counters");
    if (hipMemcpy(t->d_asym, &kAsym.v[0][0], sizeof(kAsym.v), hipMemcpyHostToDevice) != hipSuccess)
        return ctx_fail(c, MPDPFT_EHIP, "nAiT coefficient upload");
    v.asym = t->d_asym;
    {  // monomial coefficients of the quintic Hermite interpolants of L_{1/2}, L_{-3/2}
        const int nb = 500;  // (pxmax, pxmax + 500], unit step
        std::vector<double2> pp(6 * (size_t)(pn - 1 + nb));
        const long double hh = r[5];
        for (int k = 0; k + 1 < pn; k++) {
            for (int o = 0; o < 2; o++) {
                int j = o == 0 ? 1 : 3;
                const double *A = pv.data() + 6 * (size_t)k, *B = A + 6;
                long double F0 = A[j], G0 = hh * A[j + 1], C0 = hh * hh * A[j + 2];
                long double F1 = B[j], G1 = hh * B[j + 1], C1 = hh * hh * B[j + 2];
                long double m[6] = {F0, G0, 0.5L * C0,
                                    -10 * F0 - 6 * G0 - 1.5L * C0 + 10 * F1 - 4 * G1 + 0.5L * C1,
                                    15 * F0 + 8 * G0 + 1.5L * C0 - 15 * F1 + 7 * G1 - C1,
                                    -6 * F0 - 3 * G0 - 0.5L * C0 + 6 * F1 - 3 * G1 + 0.5L * C1};
                for (int q2 = 0; q2 < 3; q2++)
                    pp[6 * (size_t)k + 3 * o + q2] = double2{(double)m[2 * q2], (double)m[2 * q2 + 1]};
            }
        }
        // second segment: L_s(x) = -x^s/Gamma(s+1) (1 + sum_k 2 eta(2k) (s)_{2k} x^-2k),
        // s = 1/2 - m (m = 0..4 covers both orders and two derivatives each)
        auto Lasym = [](int m, long double x) {
            const long double s = 0.5L - m;
            const long double sqpi = 1.772453850905516027298167483341145183L;
            const long double gam[5] = {0.5L * sqpi, sqpi, -2.L * sqpi, 4.L * sqpi / 3.L, -8.L * sqpi / 15.L};
            long double x2 = 1.L / (x * x), ff = 1.L, acc = 1.L, xp = 1.L;
            for (int k = 1; k <= 12; k++) {
                ff *= (s - 2 * k + 2) * (s - 2 * k + 1);
                xp *= x2;
                acc += (long double)kTwoEta[k - 1] * ff * xp;
            }
            return -powl(x, s) / gam[m] * acc;
        };
        for (int mb = 0; mb < nb; mb++) {
            long double x0 = (long double)v.pxmax + mb, x1 = x0 + 1.L;
            for (int o = 0; o < 2; o++) {
                int m0 = o == 0 ? 0 : 2;  // L_{1/2} chain from m = 0, L_{-3/2} chain from m = 2
                long double F0 = Lasym(m0, x0), G0 = Lasym(m0 + 1, x0), C0 = Lasym(m0 + 2, x0);
                long double F1 = Lasym(m0, x1), G1 = Lasym(m0 + 1, x1), C1 = Lasym(m0 + 2, x1);
                long double m[6] = {F0, G0, 0.5L * C0,
                                    -10 * F0 - 6 * G0 - 1.5L * C0 + 10 * F1 - 4 * G1 + 0.5L * C1,
                                    15 * F0 + 8 * G0 + 1.5L * C0 - 15 * F1 + 7 * G1 - C1,
                                    -6 * F0 - 3 * G0 - 0.5L * C0 + 6 * F1 - 3 * G1 + 0.5L * C1};
                size_t row = (size_t)(pn - 1 + mb);
                for (int q2 = 0; q2 < 3; q2++)
                    pp[6 * row + 3 * o + q2] = double2{(double)m[2 * q2], (double)m[2 * q2 + 1]};
            }
        }
        v.nb = nb;
        if (hipMalloc((void **)&t->d_pp, sizeof(double2) * pp.size()) != hipSuccess)
            return ctx_fail(c, MPDPFT_ENOMEM, "nAiT polynomial table");
        if (hipMemcpy(t->d_pp, pp.data(), sizeof(double2) * pp.size(), hipMemcpyHostToDevice) != hipSuccess)
            return ctx_fail(c, MPDPFT_EHIP, "nAiT polynomial table upload");
        v.pp = t->d_pp;
    }
    if (v.ax0 != kAX0 || v.ah != kAH || v.ainv != kAINV || v.an != kAN)
        return ctx_fail(c, MPDPFT_EIO, "nAiT Airy table geometry differs from the one compiled in");
    {  // Ai in monomial form per interval (Ai'' = y Ai)
        std::vector<double2> ap(3 * (size_t)(an - 1));
        const long double hh = v.ah;
        for (int k = 0; k + 1 < an; k++) {
            long double x0 = (long double)(v.ax0 + v.ah * (double)k), x1 = x0 + hh;
            long double F0 = av[2 * (size_t)k], G0 = hh * av[2 * (size_t)k + 1], C0 = hh * hh * x0 * F0;
            long double F1 = av[2 * (size_t)k + 2], G1 = hh * av[2 * (size_t)k + 3], C1 = hh * hh * x1 * F1;
            long double m[6] = {F0, G0, 0.5L * C0,
                                -10 * F0 - 6 * G0 - 1.5L * C0 + 10 * F1 - 4 * G1 + 0.5L * C1,
                                15 * F0 + 8 * G0 + 1.5L * C0 - 15 * F1 + 7 * G1 - C1,
                                -6 * F0 - 3 * G0 - 0.5L * C0 + 6 * F1 - 3 * G1 + 0.5L * C1};
            for (int q2 = 0; q2 < 3; q2++) ap[3 * (size_t)k + q2] = double2{(double)m[2 * q2], (double)m[2 * q2 + 1]};
        }
        if (hipMalloc((void **)&t->d_ap, sizeof(double2) * ap.size()) != hipSuccess ||
            hipMalloc((void **)&t->d_ekin_v, sizeof(double) * c->G) != hipSuccess ||
            hipMalloc((void **)&t->d_ekin_m, sizeof(int) * c->G) != hipSuccess)
            return ctx_fail(c, MPDPFT_ENOMEM, "nAiT Airy table/state");
        if (hipMemcpy(t->d_ap, ap.data(), sizeof(double2) * ap.size(), hipMemcpyHostToDevice) != hipSuccess)
            return ctx_fail(c, MPDPFT_EHIP, "nAiT Airy table upload");
    }
    if (hipMalloc((void **)&t->d_view, sizeof(TabView)) != hipSuccess)
        return ctx_fail(c, MPDPFT_ENOMEM, "nAiT view");
    if (hipMemcpy(t->d_view, &v, sizeof(TabView), hipMemcpyHostToDevice) != hipSuccess)
        return ctx_fail(c, MPDPFT_EHIP, "nAiT view upload");
    t->reuse_ok = v.cx0 == kCX0 && v.ch == kCH && v.cinv == kCINV && v.cn == kCN && v.px0 == kPX0 &&
                  v.ph == kPH && v.pinv == kPINV && v.pxmax == kPXMAX && v.pn == kPN && v.nb == kNB;
    return MPDPFT_OK;
}

void nait_free(mpdpft_ctx *c) {
    if (!c->nait) return;
    (void)hipFree(c->nait->d_c);
    (void)hipFree(c->nait->d_a);
    (void)hipFree(c->nait->d_p);
    (void)hipFree(c->nait->d_asym);
    (void)hipFree(c->nait->d_pp);
    (void)hipFree(c->nait->d_view);
    (void)hipFree(c->nait->d_ap);
    (void)hipFree(c->nait->d_ekin_v);
    (void)hipFree(c->nait->d_ekin_m);
    (void)hipFree(c->nait->d_counts);
    delete c->nait;
    c->nait = nullptr;
}

void nait_reset_state(mpdpft_ctx *c) {
    if (!c->nait) return;
    for (int s = 0; s < c->S; s++) {
        for (int k = 0; k < 3; k++) c->nait->Mpp[s][k] = c->p.Mpp[k];
        c->nait->maxiter_capped[s] = 0;
    }
}

int nait_partials(mpdpft_ctx *c, int s, double mu, double *n, double *dU, double *dg2, double *dl) {
    NaitTables *t = c->nait;
    if (!t) return ctx_fail(c, MPDPFT_EINVAL, "nAiT tables not loaded");
    size_t G = c->G;
    if (c->delfield_valid != s) {
        int rc = ctx_delfield(c, c->d_V + s * G, s, c->p.DelFieldMethod);
        if (rc) return rc;
        c->delfield_valid = s;
    }
    DensityArgs da;
    da.V = c->d_V + s * G;
    da.grad2 = c->d_Grad2 + s * G;
    da.lap = c->d_Lap + s * G;
    da.mu = mu;
    da.T = c->TVec[s];
    da.degeneracy = c->p.degeneracy;
    da.alpha = c->p.Units == 2 ? 1. / 0.1312342130759153 : 1.;
    da.thr = t->Mpp[s][2];
    da.rtpna = c->p.InternalAcc;
    da.compare = c->p.Abundances[s] / std::pow(c->p.edge, (double)c->D);
    da.BP = (int)t->Mpp[s][0];
    da.maxIter = (int)(t->Mpp[s][1] + 0.5);
    if (da.BP < 1 || 4 * da.BP + 1 > 4096 || da.maxIter > kStack - 4)
        return ctx_fail(c, MPDPFT_EINVAL, "nAiT Mpp out of range");
    unsigned blocks = (unsigned)std::min<size_t>(4096, (G + 3) / 4);
    hipLaunchKernelGGL(k_nait_partials, dim3(blocks), dim3(256), 0, c->st, c->g, t->view, da, n, dU, dg2, dl);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(c, e, "k_nait_partials");
    return MPDPFT_OK;
}

int nait_density(mpdpft_ctx *c, int s, double mu, double *den_out, double *N) {
    NaitTables *t = c->nait;
    if (!t) return ctx_fail(c, MPDPFT_EINVAL, "nAiT tables not loaded");
    size_t G = c->G;
    if (c->delfield_valid != s) {  // DelField(V) per GetnAiT call (:3456); V is fixed in a bisection
        int rc = ctx_delfield(c, c->d_V + s * G, s, c->p.DelFieldMethod);
        if (rc) return rc;
        c->delfield_valid = s;
    }
    DensityArgs da;
    da.V = c->d_V + s * G;
    da.grad2 = c->d_Grad2 + s * G;
    da.lap = c->d_Lap + s * G;
    da.mu = mu;
    da.T = c->TVec[s];
    da.degeneracy = c->p.degeneracy;
    da.alpha = c->p.Units == 2 ? 1. / 0.1312342130759153 : 1.;
    da.thr = t->Mpp[s][2];
    da.rtpna = c->p.InternalAcc;
    da.compare = c->p.Abundances[s] / std::pow(c->p.edge, (double)c->D);
    da.BP = (int)t->Mpp[s][0];
    da.maxIter = (int)(t->Mpp[s][1] + 0.5);
    if (da.BP < 1 || 4 * da.BP + 1 > 4096 || da.maxIter > kStack - 4)
        return ctx_fail(c, MPDPFT_EINVAL, "nAiT Mpp out of range");
    unsigned blocks = (unsigned)std::min<size_t>(mk::kRedBlocks, (G + 255) / 256);
    hipError_t e = hipMemsetAsync(c->d_partials, 0, sizeof(double) * 4 * mk::kRedBlocks, c->st);
    if (e != hipSuccess) return hip_fail(c, e, "memset");
    // panel-reuse kernel when its LDS (node samples + one pending half per
    // tree level) fits; otherwise the plain per-node kernel
    PointScalars ps;
    ps.V = da.V;
    ps.grad2 = da.grad2;
    ps.lap = da.lap;
    ps.mu = mu;
    ps.invT = 1. / da.T;
    ps.tau = 1. / da.T;
    ps.beta0 = -da.degeneracy / std::pow(2. * kPI, 1.5);
    ps.beta1 = std::pow(da.alpha * ps.tau, -1.5);
    ps.sqrt_at = std::sqrt(da.alpha * ps.tau);
    ps.alpha = da.alpha;
    ps.thr = da.thr;
    ps.rtpna = da.rtpna;
    ps.compare = da.compare;
    ps.log10cmp = da.compare != 0. ? std::log10(std::fabs(da.compare)) : 0.;
    ps.maxexp = -std::log10(da.rtpna);
    {  // NegligibleMagnitudeQ: log10(compare) - log10|sm| > maxexp <=> |sm| < compare 10^-maxexp
        double thr10 = std::pow(10., ps.log10cmp - ps.maxexp);
        ps.negl_lo = thr10 * (1. - 1e-9);
        ps.negl_hi = thr10 * (1. + 1e-9);
    }
    ps.BP = da.BP;
    ps.maxIter = da.maxIter;
    ps.depth = da.maxIter + 2;
    ps.G = (long long)G;
    size_t lds = sizeof(double) * ((size_t)(8 * da.BP + 1) + (size_t)ps.depth * (4 * da.BP + 1 + 4) + 2);
    const bool reuse = !c->nait_plain && t->reuse_ok && 4 * da.BP + 1 <= 64 * kReuseMaxPer && lds <= 65536;
    ReuseTab rt;
    rt.pp = t->view.pp;
    rt.cc = t->view.cc;
    rt.ginv1 = t->view.ginv[1];
    rt.ginv3 = t->view.ginv[3];
    stage_begin(c);
    if (reuse && da.BP <= kLanesMaxBP && !c->nait_generic) {
        hipLaunchKernelGGL(k_nait_lanes, dim3(xcd_blocks((long long)G)), dim3(64), lds, c->st, rt, t->d_view, ps, den_out,
                           t->d_counts);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(c, e, "k_nait_lanes");
    } else if (reuse) {
        hipLaunchKernelGGL(k_nait_reuse, dim3(xcd_blocks((long long)G)), dim3(64), lds, c->st, rt, t->d_view, ps, den_out,
                           t->d_counts);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(c, e, "k_nait_reuse");
    } else {
        hipLaunchKernelGGL(k_nait_density, dim3(blocks), dim3(256), 0, c->st, c->g, t->view, da, den_out,
                           c->d_partials);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(c, e, "k_nait_density");
    }
    stage_end(c, "nait_density", 0., 0., 0.);  // the kernel alone
    if (reuse) {
        hipLaunchKernelGGL(k_nait_sums, dim3(mk::kRedBlocks), dim3(256), 0, c->st, c->g, den_out, t->d_counts,
                           c->d_partials);
        e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(c, e, "k_nait_sums");
    }
    double r[4];
    int rc = ctx_reduce(c, 4, 1., r);
    if (rc) return rc;
    // roofline numerators: integrand evaluations x nominal FP64 operations each;
    // HBM traffic V, |grad V|^2, lap V in and n out
    // HBM: V, |grad V|^2, lap V in, n out (+ three int counters per point for the reuse kernel)
    stage_add(c, "nait_density", r[3], (reuse ? 44. : 32.) * (double)G,
              r[3] * (reuse ? kNaitReuseFlopsPerEval : kNaitFlopsPerEval));
    double a = (c->p.edge / 2. - (-c->p.edge / 2.)) / ((double)c->EL - 1.);
    for (int d = 0; d < c->D; d++) r[0] *= a;
    *N = r[0];
    t->last_itermax[s] = r[1];
    t->last_relacc[s] = r[2];
    return MPDPFT_OK;
}

// AdaptnAiT (mpDPFT.cpp:3475-3551).
void nait_adapt(mpdpft_ctx *c) {
    NaitTables *t = c->nait;
    if (!t) return;
    for (int s = 0; s < c->S; s++) {
        if (t->last_itermax[s] > 0.01 * t->last_relacc[s]) {
            t->Mpp[s][0] += 1.;
            t->Mpp[s][1] += 1.;
        }
        t->Mpp[s][0] = std::min(t->Mpp[s][0], 50.);
        // the reference raises maxIter without bound; the device walk's stack holds kStack - 4
        // levels, so maxIter stops there (noted once per solve in the control stream)
        if (t->Mpp[s][1] > (double)(kStack - 4)) {
            t->Mpp[s][1] = (double)(kStack - 4);
            if (!t->maxiter_capped[s]) {
                t->maxiter_capped[s] = 1;
                c->control << " AdaptnAiT[s=" << s << "]: maxIter held at " << (kStack - 4)
                           << " (depth of the quadrature stack)\n";
            }
        }
    }
}

void nait_state(const mpdpft_ctx *c, int s, double *BP, double *maxIter) {
    *BP = c->nait ? c->nait->Mpp[s][0] : 0.;
    *maxIter = c->nait ? c->nait->Mpp[s][1] : 0.;
}

// Where EkinAiT03D's node sums stop (see ekin_nodes); MPDPFT_EKIN_FULL=1 sums
// every node up to y = 100 as the reference's loop does (same results).
static double ekin_tail() {
    const char *e = getenv("MPDPFT_EKIN_FULL");
    return (e && *e && *e != '0') ? HUGE_VAL : 25.;
}

// (<V_s, V_s>, <V_s, Env_s>): tells whether V changed since a recorded evaluation
static int v_fingerprint(mpdpft_ctx *c, int s, double *fp) {
    const size_t G = c->G;
    hipError_t e = mk::conv_dots_S(c->g, 1, c->d_V + s * G, c->d_Env + s * G, c->d_partials, c->st);
    if (e != hipSuccess) return hip_fail(c, e, "fingerprint");
    double r[3];
    int rc = ctx_reduce(c, 3, 1., r);
    if (rc) return rc;
    fp[0] = r[0];
    fp[1] = r[2];
    return MPDPFT_OK;
}

// EkinAiT03D (mpDPFT.cpp:5610-5709); c->MaxIteration persists across calls.
int nait_energy_partials(mpdpft_ctx *c, int s, double *eU, double *eg2, double *el, double *prefactor) {
    NaitTables *t = c->nait;
    if (!t) return ctx_fail(c, MPDPFT_EINVAL, "nAiT tables not loaded");
    size_t G = c->G;
    if (c->delfield_valid != s) {
        int rc = ctx_delfield(c, c->d_V + s * G, s, c->p.DelFieldMethod);
        if (rc) return rc;
        c->delfield_valid = s;
    }
    double alpha = c->p.Units == 2 ? 1. / 0.1312342130759153 : 1.;
    EkinArgs ea;
    ea.V = c->d_V + s * G;
    ea.grad2 = c->d_Grad2 + s * G;
    ea.lap = c->d_Lap + s * G;
    ea.mu = c->muVec[s];
    ea.thr = t->Mpp[s][2];
    ea.c0 = std::pow(alpha, 1. / 3.);
    ea.c1 = std::pow(alpha, -1.5);
    ea.c2 = std::pow(alpha, -0.5);
    ea.internal_acc = c->p.InternalAcc;
    ea.cmp = c->EnvE[s];
    *prefactor = c->p.degeneracy / (4. * kPI * kPI);
    ea.area_pref = std::pow(c->p.edge, (double)c->D) * (*prefactor);
    ea.neg_exp = -std::log(c->p.InternalAcc);
    ea.tail = ekin_tail();
    ea.G = (long long)G;
    ea.MaxIt = c->MaxIteration;
    ea.resume = 0;
    if (c->MaxIteration > 24) return ctx_fail(c, MPDPFT_EINVAL, "EkinAiT03D: MaxIteration above 24");
    // the refinements of the last energy evaluation, if it saw this V (same fingerprint, bit for bit),
    // mu and MaxIteration (DelField of the same V is the same)
    double fp[2];
    if (int rc = v_fingerprint(c, s, fp)) return rc;
    const bool state = t->ekin_s == s && t->ekin_fp[0] == fp[0] && t->ekin_fp[1] == fp[1] &&
                       t->ekin_mu == c->muVec[s] && t->ekin_maxit == c->MaxIteration;
    const char *vb = std::getenv("MPDPFT_VERBOSE");
    if (state && vb && *vb && *vb != '0') {  // refinement levels of the points (stderr)
        std::vector<int> m(G);
        hipError_t ce = hipMemcpyAsync(m.data(), t->d_ekin_m, sizeof(int) * G, hipMemcpyDeviceToHost, c->st);
        if (ce == hipSuccess) ce = hipStreamSynchronize(c->st);
        if (ce != hipSuccess) return hip_fail(c, ce, "ekin levels");
        long long hist[26] = {}, open = 0;
        double nodes = 0.;
        for (size_t i = 0; i < G; i++) {
            const int lv = m[i] >= 0 ? m[i] : (m[i] <= -2 ? -m[i] - 2 : -1);
            if (lv < 0 || lv > 25) continue;
            hist[lv]++;
            open += m[i] >= 0;
            nodes += 64. * std::ldexp(1., lv);
        }
        std::fprintf(stderr, "ekin partials: MaxIteration %d, points per level:", c->MaxIteration);
        for (int l = 0; l < 26; l++)
            if (hist[l]) std::fprintf(stderr, " %d:%lld", l, hist[l]);
        std::fprintf(stderr, "; unconverged %lld; nodes %.3e\n", open, nodes);
    }
    hipLaunchKernelGGL(k_ekin_partials, dim3(xcd_blocks((long long)G)), dim3(64), 0, c->st, t->d_ap, ea,
                       state ? t->d_ekin_m : nullptr, eU, eg2, el);
    hipError_t he = hipGetLastError();
    if (he != hipSuccess) return hip_fail(c, he, "k_ekin_partials");
    return MPDPFT_OK;
}

int nait_dispersal_energy(mpdpft_ctx *c, int s, double *E) {
    NaitTables *t = c->nait;
    if (!t) return ctx_fail(c, MPDPFT_EINVAL, "nAiT tables not loaded");
    size_t G = c->G;
    int rc = ctx_delfield(c, c->d_V + s * G, s, c->p.DelFieldMethod);
    if (rc) return rc;
    c->delfield_valid = s;
    double alpha = c->p.Units == 2 ? 1. / 0.1312342130759153 : 1.;
    EkinArgs ea;
    ea.V = c->d_V + s * G;
    ea.grad2 = c->d_Grad2 + s * G;
    ea.lap = c->d_Lap + s * G;
    ea.mu = c->muVec[s];
    ea.thr = t->Mpp[s][2];
    ea.c0 = std::pow(alpha, 1. / 3.);
    ea.c1 = std::pow(alpha, -1.5);
    ea.c2 = std::pow(alpha, -0.5);
    ea.internal_acc = c->p.InternalAcc;
    ea.cmp = c->EnvE[s];
    double prefactor = c->p.degeneracy / (4. * kPI * kPI);
    ea.area_pref = std::pow(c->p.edge, (double)c->D) * prefactor;
    ea.neg_exp = -std::log(c->p.InternalAcc);
    ea.tail = ekin_tail();
    ea.G = (long long)G;
    double e = 0., prev = 0.;
    int Iteration = 0, MAXIteration = c->MaxIteration + 6;
    if (MAXIteration > 24) return ctx_fail(c, MPDPFT_EINVAL, "EkinAiT03D: MaxIteration above 18");
    for (;;) {
        ea.MaxIt = c->MaxIteration;
        ea.resume = Iteration > 0;  // MaxIt is one above the previous launch's
        stage_begin(c);
        hipLaunchKernelGGL(k_ekin_ait, dim3(xcd_blocks((long long)G)), dim3(64), 0, c->st, t->d_ap, ea, c->d_tmp1, t->d_ekin_v,
                           t->d_ekin_m);
        hipError_t he = hipGetLastError();
        if (he != hipSuccess) return hip_fail(c, he, "k_ekin_ait");
        stage_end(c, "ekin_ait", (double)G, 32. * (double)G, 0.);
        mk::EnergyParams ep{};
        ep.op = mk::E_PROD;  // integrand * 1
        he = mk::fill(G, 1., c->d_tmp2, c->st);
        if (he != hipSuccess) return hip_fail(c, he, "fill");
        he = mk::energy_integrand(c->g, ep, c->d_tmp1, c->d_tmp2, nullptr, nullptr, c->d_partials, c->st);
        if (he != hipSuccess) return hip_fail(c, he, "energy");
        double v;
        rc = ctx_reduce(c, 1, 1., &v);
        if (rc) return rc;
        double a = (c->p.edge / 2. - (-c->p.edge / 2.)) / ((double)c->EL - 1.);
        for (int d = 0; d < c->D; d++) v *= a;
        e = prefactor * v;
        double as = (e == 0. && prev == 0.) ? 0. : (e - prev) / (e + prev);
        if (Iteration > 0 && std::fabs(as) < std::max(c->p.RelAcc, kMP)) break;
        if (c->MaxIteration < MAXIteration) {
            Iteration++;
            c->MaxIteration++;
            prev = e;
        } else {
            c->warningCount++;
            c->control << "GetDispersalEnergy: Warning!!! Target accuracy for EkinAiT03D not reached for species " << s
                        << "\n";
            break;
        }
    }
    *E = e;
    t->ekin_s = s;
    if (int rc = v_fingerprint(c, s, t->ekin_fp)) return rc;
    t->ekin_mu = c->muVec[s];
    t->ekin_maxit = c->MaxIteration;
    return MPDPFT_OK;
}
