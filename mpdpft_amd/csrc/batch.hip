// batch.hip -- many independent self-consistent solves in one persistent
// launch (the runY jobs the reference runs one process at a time).
//
// One 256-thread workgroup owns one instance from the first density to the
// converged result: UpdateV, GetDensities with EnforceConstraints' bracketing
// and bisection, CheckConvergence, Intercept, AdmixDensities (Pulay mixer 1 with
// its small full-pivoting QR solve) and PrepareNextIteration all run inside the
// workgroup (mpDPFT.cpp:210-225, the routines cited below).  Every scalar
// decision is taken by all threads on block-reduced (broadcast) values, so the
// control flow stays uniform and no instance ever waits for another.  A
// workgroup that finishes an instance retires it (results and fields written
// out) and takes the next one from the queue, so the launch drains a queue far
// longer than the number of resident workgroups.
//
// Fields of the running instance live in the workgroup's slot; per-slot
// storage is Env, V, OldV, Den, OldDen and the Pulay history (n_in and the
// residual per history entry), [S][G] each.  Field storage type T is double
// (FP64 path) or float (FP32 bulk: integrals, mu, Pulay Gram matrix, QR and
// energies stay FP64); an FP32 solve ends with an FP64 polish from its result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "mpdpft.h"

namespace {

constexpr int kBT = 256;        // threads per instance workgroup
constexpr int kMaxScope = 16;   // Pulay history (entries) in the batch
constexpr int kBS = 2;          // species per instance (the species of a field index is idx >= G)
constexpr double kMPb = 1.0e-12;
constexpr int kWork = 5 * (kMaxScope + 1) * (kMaxScope + 1);  // >= row + A + qr + inv + vectors + coef + flag

struct BInst {
    int D, S, EL, env[kBS];
    long long G;
    double lat0, del, trapD;   // x_i = lat0 + del i; integral factor del^D
    double c0[kBS], c1[kBS];   // environment parameters
    double pref;               // TF prefactor
    double Ab[kBS], mu0[kBS], theta0[kBS];
    double RelAcc, DMM0, beta, incV0, sccrit, degeneracy, dispPref, aux0pref;
    int contact, maxsc, mixer, scope, convmode, ekin0, ekin1, units;
    long long out_off;         // [S][G] density then [S][G] potential in the output arena
};

struct BOut {
    int SCcount, converged, LinMix, PulMix, errorCount, warningCount, bisection, pad;
    double CC, DMM, Etot, Eint;
    double mu[kBS], N[kBS], Edisp[kBS], Eenv[kBS], Aux0[kBS];
};

template <typename T> struct BArgs {
    const BInst *inst;
    int n;
    int *next;            // queue head (atomic)
    T *slots;             // per slot: (5 + 2 (scope + 1)) S Gmax values
    long long slot_stride, Gmax;
    double *fields;       // output arena (FP64)
    BOut *out;
    int polish;           // 1: start each instance from its stored FP32 result (FP64 polish)
    const double *start;  // output arena holding the start densities/potentials (polish)
    double floor_rel;     // > 0: stop criterion floor (max|dV| relative to max|V|; its square for density mode)
    int stall_max;        // > 0: leave the loop after this many iterations without a new smallest CC
    double relacc_floor;  // lower bound of the mu search's relative particle-number accuracy
};

// ---- block primitives (deterministic, result broadcast to all threads) ----
__device__ __forceinline__ double bsum(double v, double *sh) {
    const int t = threadIdx.x;
    __syncthreads();
    sh[t] = v;
    __syncthreads();
    for (int w = kBT / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] += sh[t + w];
        __syncthreads();
    }
    return sh[0];
}
__device__ __forceinline__ double bmax(double v, double *sh) {
    const int t = threadIdx.x;
    __syncthreads();
    sh[t] = v;
    __syncthreads();
    for (int w = kBT / 2; w > 0; w >>= 1) {
        if (t < w) sh[t] = fmax(sh[t], sh[t + w]);
        __syncthreads();
    }
    return sh[0];
}

// Visit the points i = threadIdx.x + k kBT < n in increasing k, kU at a time:
// the kU loads of a step are issued before their values are used, so a lane
// keeps kU loads in flight (one workgroup per instance is latency bound
// otherwise).  Per-lane accumulation order is that of the plain strided loop.
constexpr int kU = 8;
template <class Ld, class Use>
__device__ __forceinline__ void stream(long long n, Ld ld, Use use) {
    long long i = threadIdx.x;
    for (; i + (kU - 1) * kBT < n; i += kU * kBT) {
        decltype(ld(0LL)) x[kU];
#pragma unroll
        for (int u = 0; u < kU; u++) x[u] = ld(i + u * kBT);
#pragma unroll
        for (int u = 0; u < kU; u++) use(i + u * kBT, x[u]);
    }
    for (; i < n; i += kBT) use(i, ld(i));
}

struct Pair {
    double a, b;
};

__device__ __forceinline__ double dEXP(double x) {  // EXP, mpDPFT.h:266
    if (x < -700.) return 0.;
    if (x < 700.) return exp(x);
    return 1.0142320547350045e+304;
}
__device__ __forceinline__ double dSign(double x) { return x < 0. ? -1. : 1.; }

// Grid indices fit 32 bits (G <= 257^3 < 2^32): 32-bit division only.
__device__ __forceinline__ double trapw(const BInst &in, long long i) {
    double w = 1.;
    unsigned rem = (unsigned)i;
    const unsigned EL = (unsigned)in.EL;
    for (int d = 0; d < in.D; d++) {
        unsigned q = rem / EL, c = rem - q * EL;
        rem = q;
        if (c == 0 || c == EL - 1) w *= 0.5;
    }
    return w;
}

__device__ __forceinline__ void coords(const BInst &in, long long i, double *r) {
    const unsigned EL = (unsigned)in.EL, u = (unsigned)i;
    if (in.D == 2) {
        unsigned a = u / EL, b = u - a * EL;
        r[0] = in.lat0 + in.del * (double)a;
        r[1] = in.lat0 + in.del * (double)b;
        r[2] = 0.;
    } else {
        unsigned a = u / (EL * EL), rem = u - a * EL * EL, b = rem / EL, c = rem - b * EL;
        r[0] = in.lat0 + in.del * (double)a;
        r[1] = in.lat0 + in.del * (double)b;
        r[2] = in.lat0 + in.del * (double)c;
    }
}

// Per-instance scalar state of the loop (identical in every thread).
struct Loop {
    double mu[kBS], PrevMu[kBS], OldMu[kBS], TmpAb[kBS], OldAb[kBS], theta[kBS], muEval[kBS];
    double DMM, RelAcc, incV, CC, OldCC, best;
    int SCcount, ExitSCcount, ExitSCloop, LinMix, PulMix, TrappedQ, warn, err, bis, pI, stall;
};

template <typename T> struct Inst {
    const BInst &in;
    T *Env, *V, *OldV, *Den, *OldDen, *nin, *res;  // nin/res: [scope + 1][S][G]
    double *sh;                                    // LDS scratch (kBT doubles)
    double *M;                                     // LDS Gram cache [(kMaxScope+1)^2]
    double *W;                                     // LDS work area [kWork] (Gram row, matrix, QR, coefficients)
    double floor_rel;                              // reduced-precision stop floor (0: the criterion as given)
    int stall_max;                                 // reduced-precision stall exit (0: none)
    double relacc;                                 // particle-number accuracy of the mu search
    __device__ long long SG() const { return (long long)in.S * in.G; }

    // N_s(mu) of the TF density (GetTFdensity :4415-4456 + Riemann :7981-8026)
    __device__ double evalN(int s, double mu) {
        const T *v = V + (long long)s * in.G;
        const double pref = in.pref;
        const bool d2 = in.D == 2;
        double acc = 0.;
        stream(
            in.G, [&](long long i) { return (double)v[i]; },
            [&](long long i, double vi) {
                double t = mu - vi;
                double val = t > 0. ? (d2 ? pref * t : pref * t * sqrt(t)) : 0.;
                acc += trapw(in, i) * val;
            });
        return bsum(acc, sh) * in.trapD;
    }
    __device__ void writeDen(int s, double mu) {
        T *n = Den + (long long)s * in.G;
        const T *v = V + (long long)s * in.G;
        const double pref = in.pref;
        const bool d2 = in.D == 2;
        stream(
            in.G, [&](long long i) { return (double)v[i]; },
            [&](long long i, double vi) {
                double t = mu - vi;
                n[i] = (T)(t > 0. ? (d2 ? pref * t : pref * t * sqrt(t)) : 0.);
            });
    }
    __device__ double N_at(Loop &L, int s, double mu) {
        L.muEval[s] = mu;
        return evalN(s, mu);
    }
    __device__ double N_eval(Loop &L, int s, double mu) {  // a density evaluation of the bisection
        L.bis++;
        return N_at(L, s, mu);
    }

    // EnforceConstraints (mpDPFT.cpp:4569-4743), the same step sequence as the
    // single-instance solver; the density of the last evaluated mu is kept.
    __device__ void enforce(Loop &L) {
        const int S = in.S;
        L.TrappedQ = 0;
        bool abortQ = false;
        if (L.RelAcc > kMPb) {
            double nrm = 0.;
            for (int s = 0; s < S; s++) {
                L.PrevMu[s] = L.mu[s];
                nrm += L.mu[s] * L.mu[s];
            }
            if (sqrt(nrm) < kMPb)
                for (int s = 0; s < S; s++) L.PrevMu[s] = dSign(L.PrevMu[s]) * kMPb;
            double lo[kBS], hi[kBS], prev[kBS], step[kBS];
            bool loF[kBS], hiF[kBS];
            for (int s = 0; s < S; s++) {
                lo[s] = hi[s] = L.mu[s];
                prev[s] = 0.;
                step[s] = L.mu[s];
                loF[s] = hiF[s] = false;
            }
            for (int s = 0; s < S && !abortQ; s++) {
                const double target = in.Ab[s], d0 = fabs(L.TmpAb[s] - target);
                if (!(d0 / fabs(2. * target) > L.RelAcc)) continue;
                int n = 0;
                while (L.TmpAb[s] > target && n < 20) {
                    hiF[s] = true;
                    hi[s] = fmin(hi[s], L.mu[s]);
                    step[s] = dEXP(L.DMM) * (1. + (double)n) * fabs(L.PrevMu[s]);
                    prev[s] = L.mu[s];
                    L.mu[s] -= step[s];
                    if (fabs(L.mu[s]) > 1.0e+20) {
                        L.warn++;
                        L.mu[s] = L.PrevMu[s];
                        break;
                    }
                    L.TmpAb[s] = N_eval(L, s, L.mu[s]);
                    n++;
                    if (n == 1) L.DMM -= dSign(fabs(L.TmpAb[s] - target) - d0) * 0.1 * fabs(L.DMM);
                    lo[s] = L.mu[s];
                    if (L.RelAcc < kMPb) L.RelAcc = relacc;
                }
                if (L.TmpAb[s] <= target) loF[s] = true;
                n = 0;
                if (!hiF[s]) {
                    while (L.TmpAb[s] < target && n < 20) {
                        loF[s] = true;
                        if (fabs(L.mu[s]) > 1.0e+100) {
                            L.err++;
                            L.mu[s] = L.PrevMu[s];
                            break;
                        }
                        lo[s] = fmax(lo[s], L.mu[s]);
                        step[s] = dEXP(L.DMM) * (1. + (double)n) * fabs(L.PrevMu[s]);
                        prev[s] = L.mu[s];
                        L.mu[s] += step[s];
                        if (fabs(L.mu[s]) > 1.0e+20) {
                            L.warn++;
                            L.mu[s] = L.PrevMu[s];
                            break;
                        }
                        L.TmpAb[s] = N_eval(L, s, L.mu[s]);
                        n++;
                        if (n == 1) L.DMM -= dSign(fabs(L.TmpAb[s] - target) - d0) * 0.1 * fabs(L.DMM);
                        hi[s] = L.mu[s];
                        if (L.RelAcc < kMPb) L.RelAcc = relacc;
                    }
                }
                if (L.TmpAb[s] >= target) hiF[s] = true;
                if (!hiF[s] || !loF[s]) {
                    L.mu[s] = L.PrevMu[s];
                    abortQ = true;
                    L.warn++;
                }
            }
            if (!abortQ) {
                for (int s = 0; s < S; s++) {
                    step[s] = (hi[s] - lo[s]) * 0.5;
                    L.mu[s] = (lo[s] + hi[s]) * 0.5;
                }
                for (int s = 0; s < S; s++) {
                    const double target = in.Ab[s];
                    if (!(fabs(L.TmpAb[s] - target) / fabs(2. * target) > relacc)) continue;
                    // the last five positive N values (hist5[4] the newest) and their count
                    double hist5[5] = {0., 0., 0., 0., 0.};
                    int nh = 0;
                    bool success = false, tooSmall = false;
                    while (!success && !tooSmall) {
                        success = true;
                        L.TmpAb[s] = N_eval(L, s, L.mu[s]);
                        if (L.TmpAb[s] > 0.) {
                            if (nh < 5) hist5[nh] = L.TmpAb[s];
                            else {
                                for (int h = 0; h < 4; h++) hist5[h] = hist5[h + 1];
                                hist5[4] = L.TmpAb[s];
                            }
                            nh++;
                        }
                        // trapped: two equal N among the four entries before the newest (:4690-4700)
                        bool trapped = false;
                        if (nh > 5)
                            for (int h = 0; h < 3 && !trapped; h++)
                                for (int hp = h + 1; hp < 4; hp++)
                                    if (fabs(hist5[h] - hist5[hp]) < relacc * target) {
                                        trapped = true;
                                        L.TrappedQ = 1;
                                        break;
                                    }
                        if (trapped) break;
                        for (int sg = -1; sg < 2; sg += 2) {
                            if ((double)sg * (L.TmpAb[s] - target) / fabs(2. * target) > relacc) {
                                prev[s] = L.mu[s];
                                L.mu[s] -= (double)sg * 0.5 * step[s];
                                if (fabs(L.mu[s]) > 1.0e+20) {
                                    L.warn++;
                                    L.mu[s] = L.PrevMu[s];
                                    break;
                                }
                                step[s] = fabs(L.mu[s] - prev[s]);
                                if (fabs(step[s]) < kMPb) {
                                    L.warn++;
                                    tooSmall = true;
                                    break;
                                }
                                success = false;
                            }
                        }
                    }
                }
                L.RelAcc = relacc;
            }
        }
        for (int s = 0; s < S; s++)
            if (!isfinite(L.mu[s])) L.mu[s] = L.PrevMu[s];
    }

    // GetDensities (:3306-3316): N at the current mu, EnforceConstraints, and
    // the density field of the last evaluated mu
    __device__ void densities(Loop &L) {
        for (int s = 0; s < in.S; s++) L.TmpAb[s] = N_at(L, s, L.mu[s]);
        enforce(L);
        for (int s = 0; s < in.S; s++) writeDen(s, L.muEval[s]);
        __syncthreads();
    }

    // UpdateV (:4815-4838): V_s = incV (beta sum_{s' != s} n_s') + Env_s from
    // the mixed density (OldDen after the swap)
    __device__ void updateV(Loop &L) {
        const long long SGn = SG(), G = in.G;
        const bool pair = in.contact && in.S == 2;  // S <= 2: the other species' density
        const double beta = in.beta, incV = L.incV;
        stream(
            SGn,
            [&](long long idx) {
                Pair x;
                x.a = (double)Env[idx];
                x.b = pair ? (double)OldDen[idx >= G ? idx - G : idx + G] : 0.;
                return x;
            },
            [&](long long idx, Pair x) {
                double acc = 0.;
                if (in.contact) acc += x.b * beta;
                if (incV < 1.) acc *= incV;
                V[idx] = (T)(acc + x.a);
            });
        __syncthreads();
        if (L.SCcount > 0) L.incV += in.incV0;
        L.incV = fmin(1., L.incV);
    }

    // CheckConvergence (:4759-4782) with RelDev (:4784-4802)
    __device__ void convergence(Loop &L) {
        const double mp = 100. * 1.0e-16;
        if (L.SCcount > 0) {
            L.CC = 0.;
            double tot = 0.;
            for (int s = 0; s < in.S; s++) tot += L.TmpAb[s];
            double crit = in.sccrit;
            if (in.convmode == 1) {
                double m = 0., vm = 0.;
                stream(
                    SG(),
                    [&](long long idx) {
                        Pair x;
                        x.a = (double)V[idx];
                        x.b = (double)OldV[idx];
                        return x;
                    },
                    [&](long long, Pair x) {
                        m = fmax(m, fabs(x.a - x.b));
                        vm = fmax(vm, fabs(x.a));
                    });
                L.CC = bmax(m, sh);
                if (floor_rel > 0.) crit = fmax(crit, floor_rel * bmax(vm, sh));
            } else {
                double maxtest = 0.;
                for (int s = 0; s < in.S; s++) {
                    const T *a = Den + (long long)s * in.G, *b = OldDen + (long long)s * in.G;
                    double aa = 0., bb = 0., ab = 0.;
                    stream(
                        in.G,
                        [&](long long i) {
                            Pair x;
                            x.a = (double)a[i];
                            x.b = (double)b[i];
                            return x;
                        },
                        [&](long long, Pair p) {
                            aa += p.a * p.a;
                            bb += p.b * p.b;
                            ab += p.a * p.b;
                        });
                    aa = bsum(aa, sh);
                    bb = bsum(bb, sh);
                    ab = bsum(ab, sh);
                    double rd = (aa + bb > kMPb) ? 1. - 2. * ab / (aa + bb) : 0.;
                    double test = rd * L.TmpAb[s] / tot;
                    if (test > maxtest) {
                        maxtest = test;
                        L.CC = maxtest;
                    }
                }
                if (floor_rel > 0.) crit = fmax(crit, floor_rel * floor_rel);
            }
            double dab = 0., dmu = 0., th = 0.;
            for (int s = 0; s < in.S; s++) {
                dab += (L.TmpAb[s] - L.OldAb[s]) * (L.TmpAb[s] - L.OldAb[s]);
                dmu += (L.mu[s] - L.OldMu[s]) * (L.mu[s] - L.OldMu[s]);
                th += L.theta[s];
            }
            th /= (double)in.S;
            bool steady = L.CC < mp && sqrt(dab) < mp && sqrt(dmu) < mp;
            if ((tot > mp || (tot <= mp && (double)L.SCcount * th > 1.)) && in.sccrit >= 0. &&
                (L.CC < crit || steady))
                L.ExitSCcount++;
            else
                L.ExitSCcount = 0;
            if (L.ExitSCcount >= 3) L.ExitSCloop = 1;
            if (stall_max > 0) {  // rounding noise has stopped the descent: hand over to the polish
                if (L.CC < L.best) {
                    L.best = L.CC;
                    L.stall = 0;
                } else if (++L.stall >= stall_max)
                    L.ExitSCloop = 1;
            }
        } else
            L.CC = 1.2345678901234567890;
    }

    // PulayMixer, mixer 1, Euclidean metric (:5005-5167); history slots are
    // addressed through a ring (logical entry k lives in physical slot
    // (head + k) mod (scope + 1)).  Returns whether coefficients were found.
    __device__ bool pulay(Loop &L, int &head, double *coef) {
        const int scope = in.scope, ld = scope + 1, nslot = scope + 1;
        const long long SGn = SG();
        int I = L.pI;
        {  // push n_in = OldDen, R = Den - OldDen into logical entry I
            int ps = (head + I) % nslot;
            T *a = nin + (long long)ps * SGn, *r = res + (long long)ps * SGn;
            stream(
                SGn,
                [&](long long idx) {
                    Pair x;
                    x.a = (double)OldDen[idx];
                    x.b = (double)Den[idx];
                    return x;
                },
                [&](long long idx, Pair x) {
                    a[idx] = (T)x.a;
                    r[idx] = (T)(x.b - x.a);
                });
        }
        I++;
        int shift = 0;
        if (I > scope) {  // drop the oldest entry
            head = (head + 1) % nslot;
            I--;
            shift = 1;
        }
        L.pI = I;
        __syncthreads();
        // the new Gram row: dots of the newest residual with every stored one
        double *row = W, *A = W + (kMaxScope + 1);
        const T *rn = res + (long long)((head + I - 1) % nslot) * SGn;
        for (int j = 0; j < I; j++) {
            const T *rj = res + (long long)((head + j) % nslot) * SGn;
            double acc = 0.;
            stream(
                SGn,
                [&](long long idx) {
                    Pair x;
                    x.a = (double)rn[idx];
                    x.b = (double)rj[idx];
                    return x;
                },
                [&](long long, Pair x) { acc += x.a * x.b; });
            acc = bsum(acc, sh);
            if (threadIdx.x == 0) row[j] = acc;
        }
        __syncthreads();
        // the Gram matrix from the cache (shifted when the oldest entry was dropped) and
        // the coefficients, by thread 0; flag and coefficients broadcast through LDS
        if (threadIdx.x == 0) {
            for (int i = 0; i < I; i++)
                for (int j = 0; j < I; j++) {
                    double v;
                    if (i == I - 1) v = row[j];
                    else if (j == I - 1) v = row[i];
                    else v = M[(i + shift) * ld + (j + shift)];
                    A[i * I + j] = v;
                }
            for (int i = 0; i < I; i++)
                for (int j = 0; j < I; j++) M[i * ld + j] = A[i * I + j];
            W[kWork - 1] = solve_coeff(I, A, W + 2 * (kMaxScope + 1) * (kMaxScope + 1), coef) ? 1. : 0.;
        }
        __syncthreads();
        return W[kWork - 1] != 0.;
    }

    // Full-pivoting Householder QR inverse (the decomposition Eigen's
    // FullPivHouseholderQR computes, with its rank rule), the 1e-8 validation
    // and the normalised coefficients c = A^-1 1 / sum(A^-1 1).
    __device__ __noinline__ static bool solve_coeff(int n, const double *A, double *wk, double *coef) {
        constexpr int NN = (kMaxScope + 1) * (kMaxScope + 1);
        double *qr = wk, *inv = wk + NN, *hc = wk + 2 * NN, *b = hc + (kMaxScope + 1), *y = b + (kMaxScope + 1),
               *x = y + (kMaxScope + 1);
        int *rt = (int *)(x + (kMaxScope + 1)), *ct = rt + (kMaxScope + 1), *perm = ct + (kMaxScope + 1);
        for (int k = 0; k < n * n; k++) qr[k] = A[k];
        const double eps = 2.220446049250313e-16;
        double prec = eps * n, biggest = 0., maxpivot = 0.;
        int nonzero = n;
        for (int k = 0; k < n; k++) {
            int rb = k, cb = k;
            double best = -1.;
            for (int j = k; j < n; j++)
                for (int i = k; i < n; i++)
                    if (fabs(qr[i * n + j]) > best) {
                        best = fabs(qr[i * n + j]);
                        rb = i;
                        cb = j;
                    }
            if (k == 0) biggest = best;
            if (best <= biggest * prec) {
                nonzero = k;
                for (int i = k; i < n; i++) {
                    rt[i] = ct[i] = i;
                    hc[i] = 0.;
                }
                break;
            }
            rt[k] = rb;
            ct[k] = cb;
            if (rb != k)
                for (int j = k; j < n; j++) {
                    double t = qr[k * n + j];
                    qr[k * n + j] = qr[rb * n + j];
                    qr[rb * n + j] = t;
                }
            if (cb != k)
                for (int i = 0; i < n; i++) {
                    double t = qr[i * n + k];
                    qr[i * n + k] = qr[i * n + cb];
                    qr[i * n + cb] = t;
                }
            double c0 = qr[k * n + k], tail = 0.;
            for (int i = k + 1; i < n; i++) tail += qr[i * n + k] * qr[i * n + k];
            double tau, beta;
            if (tail <= 2.2250738585072014e-308) {
                tau = 0.;
                beta = c0;
                for (int i = k + 1; i < n; i++) qr[i * n + k] = 0.;
            } else {
                beta = sqrt(c0 * c0 + tail);
                if (c0 >= 0.) beta = -beta;
                for (int i = k + 1; i < n; i++) qr[i * n + k] /= (c0 - beta);
                tau = (beta - c0) / beta;
            }
            qr[k * n + k] = beta;
            hc[k] = tau;
            maxpivot = fmax(maxpivot, fabs(beta));
            for (int j = k + 1; j < n; j++) {
                double dot = qr[k * n + j];
                for (int i = k + 1; i < n; i++) dot += qr[i * n + k] * qr[i * n + j];
                dot *= tau;
                qr[k * n + j] -= dot;
                for (int i = k + 1; i < n; i++) qr[i * n + j] -= dot * qr[i * n + k];
            }
        }
        double thr = maxpivot * eps * n;
        int rank = 0;
        for (int i = 0; i < nonzero; i++) rank += fabs(qr[i * n + i]) > thr;
        if (rank != n) return false;
        for (int i = 0; i < n; i++) perm[i] = i;
        for (int k = 0; k < n; k++) {
            int t = perm[k];
            perm[k] = perm[ct[k]];
            perm[ct[k]] = t;
        }
        for (int col = 0; col < n; col++) {
            for (int i = 0; i < n; i++) b[i] = i == col ? 1. : 0.;
            for (int k = 0; k < n; k++) {
                if (rt[k] != k) {
                    double t = b[k];
                    b[k] = b[rt[k]];
                    b[rt[k]] = t;
                }
                double dot = b[k];
                for (int i = k + 1; i < n; i++) dot += qr[i * n + k] * b[i];
                dot *= hc[k];
                b[k] -= dot;
                for (int i = k + 1; i < n; i++) b[i] -= dot * qr[i * n + k];
            }
            for (int i = n - 1; i >= 0; i--) {
                double acc = b[i];
                for (int j = i + 1; j < n; j++) acc -= qr[i * n + j] * y[j];
                y[i] = acc / qr[i * n + i];
            }
            for (int i = 0; i < n; i++) inv[perm[i] * n + col] = y[i];
        }
        double tv = 0.;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++) {
                double acc = 0.;
                for (int k = 0; k < n; k++) acc += inv[i * n + k] * A[k * n + j];
                acc -= i == j ? 1. : 0.;
                tv += acc * acc;
            }
        if (!(tv <= 1.0e-8)) return false;
        double Asum = 0.;
        for (int i = 0; i < n; i++) {
            double acc = 0.;
            for (int k = 0; k < n; k++) acc += inv[i * n + k];
            x[i] = acc;
        }
        for (int i = 0; i < n; i++) Asum += x[i];
        for (int i = 0; i < n; i++) coef[i] = x[i] / Asum;
        return true;
    }

    // AdmixDensities (:4921-4992)
    __device__ void admix(Loop &L, int &head) {
        const int S = in.S;
        double nA = 0., nD = 0.;
        for (int s = 0; s < S; s++) {
            nD += (in.Ab[s] - L.TmpAb[s]) * (in.Ab[s] - L.TmpAb[s]);
            nA += in.Ab[s] * in.Ab[s];
        }
        double rel = sqrt(nD) / sqrt(nA), mod = 1.;
        if (rel > kMPb) mod = fmin(1., pow(fabs(1. - rel), 30.));
        if (L.TrappedQ)
            for (int s = 0; s < S; s++) L.theta[s] *= mod;
        double *coef = W + kWork - 2 - (kMaxScope + 1);
        bool ok = false;
        if (in.mixer > 0 && L.SCcount > 0 && L.CC < 1.0e+300) ok = pulay(L, head, coef);
        const long long SGn = SG();
        if (ok) {
            L.PulMix++;
            const int nslot = in.scope + 1, pI = L.pI;
            const double th0 = L.theta[0], th1 = L.theta[kBS - 1];
            const long long G = in.G;
            for (long long idx = threadIdx.x; idx < SGn; idx += kBT) {
                const double th = idx >= G ? th1 : th0;  // S <= 2
                double acc = 0.;
                int slot = head;  // physical slot of logical entry p
#pragma unroll 4
                for (int p = 0; p < pI; p++) {
                    const long long off = (long long)slot * SGn + idx;
                    acc += coef[p] * ((double)nin[off] + th * (double)res[off]);
                    slot = slot + 1 == nslot ? 0 : slot + 1;
                }
                Den[idx] = (T)acc;
            }
        } else {
            L.LinMix++;
            const double th0 = L.theta[0], th1 = L.theta[kBS - 1];
            const long long G = in.G;
            stream(
                SGn,
                [&](long long idx) {
                    Pair x;
                    x.a = (double)Den[idx];
                    x.b = (double)OldDen[idx];
                    return x;
                },
                [&](long long idx, Pair x) {
                    const double th = idx >= G ? th1 : th0;  // S <= 2
                    Den[idx] = (T)(th * x.a + (1. - th) * x.b);
                });
        }
        __syncthreads();
        if (L.TrappedQ)
            for (int s = 0; s < S; s++) L.theta[s] *= 1. / mod;
    }
};

// One launch: each workgroup takes instances from the queue until it is empty.
template <typename T>
__global__ __launch_bounds__(kBT) void k_batch_solve(BArgs<T> a) {
    __shared__ double sh[kBT];
    __shared__ double Mcache[(kMaxScope + 1) * (kMaxScope + 1)];
    __shared__ double Work[kWork];
    __shared__ int nextIdx;
    T *base = a.slots + (long long)blockIdx.x * a.slot_stride;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) nextIdx = atomicAdd(a.next, 1);
        __syncthreads();
        const int k = nextIdx;
        if (k >= a.n) break;
        const BInst &in = a.inst[k];
        const long long SGn = (long long)in.S * in.G;
        Inst<T> I{in};
        I.Env = base;
        I.V = base + SGn;
        I.OldV = base + 2 * SGn;
        I.Den = base + 3 * SGn;
        I.OldDen = base + 4 * SGn;
        I.nin = base + 5 * SGn;
        I.res = I.nin + (long long)(in.scope + 1) * SGn;
        I.sh = sh;
        I.M = Mcache;
        I.W = Work;
        I.floor_rel = a.floor_rel;
        I.stall_max = a.stall_max;
        I.relacc = fmax(in.RelAcc, a.relacc_floor);
        double *outDen = a.fields + in.out_off, *outV = outDen + SGn;
        // InitializeEnvironments (:6362-6390): Env, and V <- Env (or the stored start)
        for (long long idx = threadIdx.x; idx < SGn; idx += kBT) {
            int s = idx >= in.G ? 1 : 0;  // S <= 2
            long long i = idx - (long long)s * in.G;
            double r[3];
            coords(in, i, r);
            double v = 0.;
            if (in.env[s] == 1) v = in.c0[s] * pow(r[0] * r[0] + r[1] * r[1] + r[2] * r[2], 1.);
            else if (in.env[s] == 3) v = in.c0[s] * ((1. - in.c1[s]) * r[0] * r[0] + (1. + in.c1[s]) * r[1] * r[1]);
            I.Env[idx] = (T)v;
            I.V[idx] = (T)(a.polish ? a.start[in.out_off + SGn + idx] : v);
        }
        __syncthreads();
        Loop L;
        for (int s = 0; s < in.S; s++) {
            L.mu[s] = a.polish ? a.out[k].mu[s] : in.mu0[s];
            L.theta[s] = in.theta0[s];
            L.TmpAb[s] = L.OldAb[s] = L.OldMu[s] = L.PrevMu[s] = 0.;
            L.muEval[s] = L.mu[s];
        }
        L.DMM = a.polish ? a.out[k].DMM : in.DMM0;
        L.RelAcc = I.relacc;
        L.incV = in.incV0;
        L.CC = L.OldCC = 0.;
        L.best = 1.0e+300;
        L.SCcount = L.ExitSCcount = L.ExitSCloop = L.LinMix = L.PulMix = L.TrappedQ = 0;
        L.warn = L.err = L.bis = L.pI = L.stall = 0;
        int head = 0;
        I.densities(L);  // initial densities (:194-201)
        while (!L.ExitSCloop && in.maxsc > 0) {
            // StoreCurrentIteration (:4804): swap the field pointers
            T *t = I.Den;
            I.Den = I.OldDen;
            I.OldDen = t;
            t = I.V;
            I.V = I.OldV;
            I.OldV = t;
            L.OldCC = L.CC;
            for (int s = 0; s < in.S; s++) {
                L.OldMu[s] = L.mu[s];
                L.OldAb[s] = L.TmpAb[s];
            }
            I.updateV(L);
            I.densities(L);
            I.convergence(L);
            if (L.SCcount == in.maxsc) L.ExitSCloop = 1;  // Intercept
            I.admix(L, head);
            L.SCcount++;  // PrepareNextIteration
        }
        // energies (GetEnergy :5414-5436): contact, environment, TF dispersal, AuxEkin[0]
        double Eint = 0.;
        if (in.contact && in.S > 1) {
            double acc = 0.;
            for (long long i = threadIdx.x; i < in.G; i += kBT)
                acc += trapw(in, i) * (double)I.Den[i] * (double)I.Den[in.G + i];
            Eint = in.beta * bsum(acc, sh) * in.trapD;
        }
        BOut o{};
        double Etot = Eint;
        const double DD = (double)in.D;
        for (int s = 0; s < in.S; s++) {
            const T *n = I.Den + (long long)s * in.G, *v = I.V + (long long)s * in.G, *e = I.Env + (long long)s * in.G;
            double ee = 0., ed = 0., ea = 0.;
            for (long long i = threadIdx.x; i < in.G; i += kBT) {
                double w = trapw(in, i), nn = (double)n[i];
                ee += w * (double)e[i] * nn;
                double tt = L.mu[s] - (double)v[i];
                ed += w * (tt > 0. ? pow(tt, (DD + 2.) / 2.) : 0.);
                ea += w * pow(nn > 0. ? nn : 0., (DD + 2.) / DD);
            }
            ee = bsum(ee, sh) * in.trapD;
            ed = bsum(ed, sh) * in.trapD;
            ea = bsum(ea, sh) * in.trapD;
            double aux0 = in.units > 0 ? in.aux0pref * ea : 0.;
            double disp = (in.ekin0 ? in.dispPref * ed : 0.) + (in.ekin1 ? aux0 : 0.);
            o.Eenv[s] = ee;
            o.Edisp[s] = disp;
            o.Aux0[s] = aux0;
            Etot += disp + ee;
        }
        // retire: fields and results out
        for (long long idx = threadIdx.x; idx < SGn; idx += kBT) {
            outDen[idx] = (double)I.Den[idx];
            outV[idx] = (double)I.V[idx];
        }
        if (threadIdx.x == 0) {
            o.SCcount = L.SCcount;
            o.converged = L.ExitSCcount >= 3;
            o.LinMix = L.LinMix;
            o.PulMix = L.PulMix;
            o.errorCount = L.err;
            o.warningCount = L.warn + (in.maxsc > 0 && L.SCcount >= in.maxsc ? 1 : 0);
            o.bisection = L.bis;
            o.CC = L.CC;
            o.DMM = L.DMM;
            o.Etot = Etot;
            o.Eint = Eint;
            for (int s = 0; s < in.S; s++) {
                o.mu[s] = L.mu[s];
                o.N[s] = L.TmpAb[s];
            }
            a.out[k] = o;
        }
    }
}

}  // namespace

struct mpdpft_batch {
    int device = 0, n = 0, flags = 0, nslots = 0, scope = 0;
    long long Gmax = 0, arena = 0;
    std::vector<BInst> inst;
    std::vector<mpdpft_params> params;
    BInst *d_inst = nullptr;
    BOut *d_out = nullptr;
    int *d_next = nullptr;
    void *d_slots = nullptr;
    double *d_fields = nullptr, *d_start = nullptr;
    std::vector<BOut> out;
    hipStream_t st = nullptr;
    std::string err;
    double seconds = 0., phase_seconds[2] = {0., 0.};
    long long phase_iterations[2] = {0, 0}, phase_evaluations[2] = {0, 0};
};

static int bfail(mpdpft_batch *b, int rc, const std::string &m) {
    b->err = m;
    return rc;
}

// mpdpft_params of one instance -> BInst (the set-up of mpdpft_create,
// InitializeHardCodedParameters :863-896, for the TF family).
static int prepare(mpdpft_batch *b, const mpdpft_params &p, BInst &o, int i) {
    auto bad = [&](const std::string &m) { return bfail(b, MPDPFT_EINVAL, "instance " + std::to_string(i) + ": " + m); };
    if (p.DIM != 2 && p.DIM != 3) return bad("DIM must be 2 or 3");
    if (p.S < 1 || p.S > kBS) return bad("S must be 1 or 2");
    if (p.DensityExpression != 7) return bad("the batch runs DensityExpression 7 (TF)");
    if (p.TaskType != 0 || p.Symmetry != 0 || p.Noise != 0. || p.InterpolVQ != 0 || p.method != 1)
        return bad("TaskType, Symmetry, Noise, InterpolVQ must be 0 and method 1");
    for (int k = 0; k < p.n_schedule; k++)
        if (p.Schedule[k] != 0) return bad("schedules are not run in the batch");
    int contact = 0;
    for (int a = 0; a < p.n_interactions; a++) {
        if (p.Interactions[a] == 1) contact = 1;
        else if (p.Interactions[a] != 0) return bad("the batch runs interactions 0 and 1 (contact)");
    }
    if ((int)(p.Mixer[0] + 0.5) > 1 || std::fabs(p.Mixer[2]) > kMPb) return bad("Pulay mixer 0/1, Euclidean metric");
    for (int e = 0; e < p.n_ekin; e++)
        if (p.EkinTypes[e] < 0 || p.EkinTypes[e] > (p.DIM == 2 ? 4 : 1))
            return bad("EkinTypes: 0 and 1 (3D), 0..4 (2D)");
    int scope = std::max(1, (int)p.Mixer[1]);
    if (scope > kMaxScope) return bad("Pulay scope above 16");
    if (p.steps < 2 || p.steps > 512) return bad("steps out of range");
    o.D = p.DIM;
    o.S = p.S;
    o.EL = p.steps + 1;
    o.G = 1;
    for (int d = 0; d < o.D; d++) o.G *= o.EL;
    const double x1 = -p.edge / 2., x2 = p.edge / 2.;
    o.lat0 = x1;
    o.del = (x2 - x1) / (double)p.steps;
    o.trapD = std::pow((p.edge / 2. - (-p.edge / 2.)) / ((double)o.EL - 1.), (double)o.D);
    for (int s = 0; s < o.S; s++) {
        int t = p.Environments[s];
        if (t != 0 && t != 1 && t != 3) return bad("environments 0, 1, 3");
        if ((t == 3) && p.n_mpp < 2 * p.S) return bad("mpp needs 2S values");
        if (t == 1 && p.n_mpp < p.S) return bad("mpp needs S values");
        o.env[s] = t;
        o.c0[s] = p.n_mpp > s ? p.mpp[s] : 0.;
        o.c1[s] = p.n_mpp > p.S + s ? p.mpp[p.S + s] : 0.;
        o.Ab[s] = p.Abundances[s];
        o.mu0[s] = p.muVec[s];
        o.theta0[s] = p.thetaVec[s];
    }
    const double g = p.degeneracy, PI = 3.141592653589793;
    o.pref = o.D == 2 ? g / (2. * PI) : (p.Units == 2 ? 0.0475412512 : 1.) * std::sqrt(2.) * g / (3. * PI * PI);
    o.RelAcc = p.RelAcc;
    o.DMM0 = p.DeltamuModifier;
    o.beta = p.beta;
    o.contact = contact;
    o.incV0 = p.incrementalV;
    o.sccrit = p.SCcriterion;
    o.degeneracy = g;
    o.maxsc = p.maxSCcount;
    o.mixer = (int)(p.Mixer[0] + 0.5);
    o.scope = scope;
    o.convmode = p.ConvergenceMode;
    o.units = p.Units;
    o.ekin0 = o.ekin1 = 0;
    for (int e = 0; e < p.n_ekin; e++) {
        if (p.EkinTypes[e] == 0) o.ekin0 = 1;
        else if (p.EkinTypes[e] == 1) o.ekin1 = 1;
    }
    const double DD = (double)o.D, U = p.Units == 2 ? 7.619964158443410 : 1.;
    const double af = o.D == 2 ? 2. * PI : 4. * PI;
    o.dispPref = g * af * std::pow(2., (DD + 2.) / 2.) / (std::pow(2. * PI * std::sqrt(U), DD) * (2. * DD + 4.));
    double ap = o.D == 2 ? PI / g : std::pow(243. * std::pow(PI, 4.) / (250. * g * g), 1. / 3.);
    if (p.Units == 2) ap *= 1. / 0.131163;
    o.aux0pref = ap;
    o.out_off = 0;
    return MPDPFT_OK;
}

#define BHIP(x)                                                                                     \
    do {                                                                                            \
        hipError_t e_ = (x);                                                                        \
        if (e_ != hipSuccess) return bfail(b, MPDPFT_EHIP, std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

extern "C" int mpdpft_batch_create(const mpdpft_params *p, int n, int device, int flags, mpdpft_batch **out) {
    if (!p || n < 1 || !out) return MPDPFT_EINVAL;
    mpdpft_batch *b = new mpdpft_batch();
    *out = b;
    b->n = n;
    b->flags = flags;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(b, MPDPFT_ENODEV, "no GPU visible");
    b->device = device < 0 ? 0 : device;
    BHIP(hipSetDevice(b->device));
    b->inst.resize(n);
    b->params.assign(p, p + n);
    long long off = 0;
    for (int i = 0; i < n; i++) {
        int rc = prepare(b, p[i], b->inst[i], i);
        if (rc) return rc;
        b->inst[i].out_off = off;
        off += 2LL * b->inst[i].S * b->inst[i].G;
        b->Gmax = std::max(b->Gmax, b->inst[i].G);
        b->scope = std::max(b->scope, b->inst[i].scope);
    }
    b->arena = off;
    // slots: resident workgroups; at most one per instance
    hipDeviceProp_t prop;
    BHIP(hipGetDeviceProperties(&prop, b->device));
    const char *env = std::getenv("MPDPFT_BATCH_SLOTS");
    int want = env ? std::atoi(env) : 2 * prop.multiProcessorCount;
    b->nslots = std::max(1, std::min(n, want));
    BHIP(hipStreamCreate(&b->st));
    BHIP(hipMalloc((void **)&b->d_inst, sizeof(BInst) * n));
    BHIP(hipMalloc((void **)&b->d_out, sizeof(BOut) * n));
    BHIP(hipMalloc((void **)&b->d_next, sizeof(int)));
    BHIP(hipMalloc((void **)&b->d_fields, sizeof(double) * (size_t)off));
    const long long per = (5LL + 2LL * (b->scope + 1)) * kBS * b->Gmax;
    BHIP(hipMalloc(&b->d_slots, sizeof(double) * (size_t)per * b->nslots));
    BHIP(hipMemcpy(b->d_inst, b->inst.data(), sizeof(BInst) * n, hipMemcpyHostToDevice));
    b->out.resize(n);
    return MPDPFT_OK;
}

// Reduced-precision stop: FP32 fields hold V to a few ulp of max|V|, so the
// FP32 loop stops at max|dV| < 1e-6 max|V| (density mode: 1e-12) or after 16
// iterations without progress; the FP64 polish then meets the instance's own
// criterion.  The FP32 loop's mu search stops at |N - N0| / 2N0 <= 1e-8: the
// rounding of the stored FP32 densities moves N by more than the usual 1e-11
// every iteration, which would cost a full bisection each time.
constexpr double kF32Floor = 1.0e-6;
constexpr int kF32Stall = 16;
constexpr double kF32RelAcc = 1.0e-8;

template <typename T> static int run(mpdpft_batch *b, int polish) {
    const long long per = (5LL + 2LL * (b->scope + 1)) * kBS * b->Gmax;
    BArgs<T> a;
    a.inst = b->d_inst;
    a.n = b->n;
    a.next = b->d_next;
    a.slots = (T *)b->d_slots;
    a.slot_stride = per;  // values of T per slot (the arena is sized for doubles)
    a.Gmax = b->Gmax;
    a.fields = b->d_fields;
    a.out = b->d_out;
    a.polish = polish;
    a.start = b->d_start;
    a.floor_rel = sizeof(T) == 4 ? kF32Floor : 0.;
    a.stall_max = sizeof(T) == 4 ? kF32Stall : 0;
    a.relacc_floor = sizeof(T) == 4 ? kF32RelAcc : 0.;
    BHIP(hipMemsetAsync(b->d_next, 0, sizeof(int), b->st));
    hipLaunchKernelGGL(k_batch_solve<T>, dim3(b->nslots), dim3(kBT), 0, b->st, a);
    BHIP(hipGetLastError());
    return MPDPFT_OK;
}

extern "C" int mpdpft_batch_solve(mpdpft_batch *b, mpdpft_result *res) {
    if (!b || !res || !b->st) return MPDPFT_EINVAL;
    BHIP(hipSetDevice(b->device));
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    const bool fp32 = (b->flags & MPDPFT_BATCH_FP32) != 0;
    b->phase_seconds[0] = b->phase_seconds[1] = 0.;
    b->phase_iterations[0] = b->phase_iterations[1] = 0;
    b->phase_evaluations[0] = b->phase_evaluations[1] = 0;
    auto count = [&](int ph) -> int {  // iterations of the phase just run (host copy of the results)
        BHIP(hipMemcpy(b->out.data(), b->d_out, sizeof(BOut) * b->n, hipMemcpyDeviceToHost));
        for (int i = 0; i < b->n; i++) {
            b->phase_iterations[ph] += b->out[i].SCcount;
            b->phase_evaluations[ph] += b->out[i].bisection;
        }
        return MPDPFT_OK;
    };
    if (fp32) {
        rc = run<float>(b, 0);
        if (rc) return rc;
        // FP64 polish from the FP32 result: its potential and mu start a full FP64 loop
        if (!b->d_start) BHIP(hipMalloc((void **)&b->d_start, sizeof(double) * (size_t)b->arena));
        BHIP(hipMemcpyAsync(b->d_start, b->d_fields, sizeof(double) * (size_t)b->arena, hipMemcpyDeviceToDevice, b->st));
        BHIP(hipStreamSynchronize(b->st));
        b->phase_seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = count(0))) return rc;
        auto t1 = std::chrono::steady_clock::now();
        rc = run<double>(b, 1);
        if (rc) return rc;
        BHIP(hipStreamSynchronize(b->st));
        b->phase_seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        b->seconds = b->phase_seconds[0] + b->phase_seconds[1];
        if ((rc = count(1))) return rc;
    } else {
        rc = run<double>(b, 0);
        if (rc) return rc;
        BHIP(hipStreamSynchronize(b->st));
        b->seconds = b->phase_seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = count(1))) return rc;
    }
    for (int i = 0; i < b->n; i++) {
        const BOut &o = b->out[i];
        mpdpft_result &r = res[i];
        std::memset(&r, 0, sizeof(r));
        r.S = b->inst[i].S;
        r.SCcount = o.SCcount;
        r.converged = o.converged;
        r.LinMix = o.LinMix;
        r.PulMix = o.PulMix;
        r.errorCount = o.errorCount;
        r.warningCount = o.warningCount;
        r.CC = o.CC;
        r.DeltamuModifier = o.DMM;
        r.TotalEnergy = o.Etot;
        r.bisection_steps = o.bisection;
        r.solve_seconds = b->seconds;
        const mpdpft_params &p = b->params[i];
        for (int s = 0; s < r.S; s++) {
            r.muVec[s] = o.mu[s];
            r.FinalAbundances[s] = o.N[s];
            r.DispersalEnergies[s] = o.Edisp[s];
            r.EnvironmentalEnergies[s] = o.Eenv[s];
            r.AuxEkin[s][0] = o.Aux0[s];
        }
        for (int a = 0; a < p.n_interactions; a++) r.InteractionEnergies[a] = p.Interactions[a] == 1 ? o.Eint : 0.;
    }
    return MPDPFT_OK;
}

extern "C" int mpdpft_batch_get_fields(mpdpft_batch *b, int i, double *den, double *V) {
    if (!b || i < 0 || i >= b->n) return MPDPFT_EINVAL;
    const BInst &in = b->inst[i];
    const size_t SG = (size_t)in.S * in.G;
    if (den) BHIP(hipMemcpy(den, b->d_fields + in.out_off, sizeof(double) * SG, hipMemcpyDeviceToHost));
    if (V) BHIP(hipMemcpy(V, b->d_fields + in.out_off + SG, sizeof(double) * SG, hipMemcpyDeviceToHost));
    return MPDPFT_OK;
}

extern "C" double mpdpft_batch_seconds(const mpdpft_batch *b) { return b ? b->seconds : 0.; }

extern "C" int mpdpft_batch_phase_stats(const mpdpft_batch *b, double *seconds, long long *iterations,
                                        long long *bisection_steps) {
    if (!b) return MPDPFT_EINVAL;
    for (int ph = 0; ph < 2; ph++) {
        if (seconds) seconds[ph] = b->phase_seconds[ph];
        if (iterations) iterations[ph] = b->phase_iterations[ph];
        if (bisection_steps) bisection_steps[ph] = b->phase_evaluations[ph];
    }
    return MPDPFT_OK;
}

extern "C" const char *mpdpft_batch_last_error(const mpdpft_batch *b) { return b ? b->err.c_str() : "null batch"; }

extern "C" void mpdpft_batch_destroy(mpdpft_batch *b) {
    if (!b) return;
    if (b->st) (void)hipStreamSynchronize(b->st);
    (void)hipFree(b->d_inst);
    (void)hipFree(b->d_out);
    (void)hipFree(b->d_next);
    (void)hipFree(b->d_slots);
    (void)hipFree(b->d_fields);
    (void)hipFree(b->d_start);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}
