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
// storage is Env, V, OldV and the Pulay history (n_in and the residual per
// history entry, plus the slot that holds the current mixed density), [S][G]
// each.  An iteration's unmixed density is never stored: it is the TF density
// of V at the evaluated mu, recomputed where it is used.  Field storage type T is double
// (FP64 path) or float (FP32 bulk: integrals, mu, Pulay Gram matrix, QR and
// energies stay FP64); an FP32 solve ends with an FP64 polish from its result.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "mpdpft.h"

const char *params_structure_error(const mpdpft_params *p);  // params.cpp

namespace {

constexpr int kBT = 256;        // threads per instance workgroup
constexpr int kMaxScope = 16;   // Pulay history (entries) in the batch
constexpr int kBS = 2;          // species per instance (the species of a field index is idx >= G)
constexpr double kMPb = 1.0e-12;
constexpr int kWork = 5 * (kMaxScope + 1) * (kMaxScope + 1);  // >= row + A + qr + inv + vectors + coef + flag
constexpr int kGC = 8;   // stored Pulay entries dotted per sweep of the new residual
constexpr int kMixC = 8; // Pulay entries combined per load batch of the mixing
constexpr int kUA = 2;   // grid points per lane visit in sweep A
constexpr int kTeamMax = 2;  // workgroups that share one instance (the largest instances)
constexpr int kRedMax = 16;  // values per team reduction
constexpr int kUB = 2;   // grid points per lane visit in sweep B

struct BInst {
    int D, S, EL, env[kBS];
    long long G;
    double invEL;              // 1 / EL (index arithmetic)
    double lat0, del, trapD;   // x_i = lat0 + del i; integral factor del^D
    double c0[kBS], c1[kBS];   // environment parameters
    double pref;               // TF prefactor
    double Ab[kBS], mu0[kBS], theta0[kBS];
    double RelAcc, DMM0, beta, incV0, sccrit, degeneracy, dispPref, aux0pref;
    int contact, maxsc, mixer, scope, convmode, ekin0, ekin1, units;
    long long out_off;         // [S][G] density then [S][G] potential in the output arena
};

struct BOut {
    int SCcount, converged, LinMix, PulMix, errorCount, warningCount, bisection, bracket_failures;
    int pulay_rejections, pad;
    double CC, DMM, Etot, Eint;
    double mu[kBS], N[kBS], Edisp[kBS], Eenv[kBS], Aux0[kBS];
    double bytes;                 // bytes the instance's fused sweeps moved (each field access once)
    double sbytes;                // the reference stages' (inputs + outputs) bytes of the same loop
    unsigned long long t0, t1;    // s_memrealtime (100 MHz) when the instance was taken and retired
    unsigned long long tph[4];    // time in sweep A, the mu search, sweep B, the Pulay solve (100 MHz ticks)
};

// Per-instance exchange area of a team: each member's block partials, by
// reduction generation parity, and the arrival counter of the team barrier.
struct TeamArea {
    unsigned count;
    unsigned pad[63];
    double part[2][kTeamMax][kRedMax];
};

// A queue entry: instance, rank in its team, team size.
struct Ticket {
    int inst, rank, size, pad;
};

template <typename T> struct BArgs {
    const BInst *inst;
    const Ticket *tickets;  // queue: one ticket per team member, largest instances first
    int n;                  // tickets
    TeamArea *team;         // [instances]
    int *team_err;          // set when a team barrier waited past its bound
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
// Wave level: quad and row butterflies (DPP), then the four row sums in a fixed
// order through readlane (wave-uniform result); block level: the four wave
// values through LDS, combined in wave order.  Two barriers per reduction.
template <int CTRL>
__device__ __forceinline__ double bdpp(double v) {
    int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
    int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double blane(double v, int l) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
// A wave-uniform double moved to scalar registers (the value is the same in
// every lane; the compiler then keeps what is derived from it scalar).
__device__ __forceinline__ double uni(double v) {
    int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wsum(double v) {
    v += bdpp<0xB1>(v);   // quad_perm [1,0,3,2]
    v += bdpp<0x4E>(v);   // quad_perm [2,3,0,1]
    v += bdpp<0x141>(v);  // row_half_mirror
    v += bdpp<0x140>(v);  // row_mirror
    return (blane(v, 0) + blane(v, 16)) + (blane(v, 32) + blane(v, 48));
}
__device__ __forceinline__ double wmax(double v) {
    v = fmax(v, bdpp<0xB1>(v));
    v = fmax(v, bdpp<0x4E>(v));
    v = fmax(v, bdpp<0x141>(v));
    v = fmax(v, bdpp<0x140>(v));
    return fmax(fmax(blane(v, 0), blane(v, 16)), fmax(blane(v, 32), blane(v, 48)));
}
constexpr int kWaves = kBT / 64;
static_assert(kWaves == 4, "the block combine below is written for four wavefronts");
__device__ __forceinline__ double bsum(double v, double *sh) {
    v = wsum(v);
    __syncthreads();  // sh is free (the previous reduction's readers are done)
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return uni((sh[0] + sh[1]) + (sh[2] + sh[3]));
}
__device__ __forceinline__ double bmax(double v, double *sh) {
    v = wmax(v);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    return uni(fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3])));
}
// n <= NV sums at once (the same order per value as bsum)
template <int NV>
__device__ __forceinline__ void bsum_n(double (&v)[NV], int n, double *sh) {
#pragma unroll
    for (int k = 0; k < NV; k++)
        if (k < n) v[k] = wsum(v[k]);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        const int w = threadIdx.x >> 6;
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < n) sh[w * NV + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++)
        if (k < n) v[k] = uni((sh[k] + sh[NV + k]) + (sh[2 * NV + k] + sh[3 * NV + k]));
}

// Visit the points i = lo + threadIdx.x + k kBT < n in increasing k, kU at a time:
// the kU loads of a step are issued before their values are used, so a lane
// keeps kU loads in flight (one workgroup per instance is latency bound
// otherwise).  Per-lane accumulation order is that of the plain strided loop.
constexpr int kU = 16;
template <class Ld, class Use>
__device__ __forceinline__ void stream(long long lo, long long n, Ld ld, Use use) {
    long long i = lo + threadIdx.x;
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

// Exact i / d for 32-bit i through the FP64 reciprocal inv = 1/d, corrected
// by one (the product can fall one below or above the quotient).
__device__ __forceinline__ unsigned udiv(unsigned i, unsigned d, double inv) {
    unsigned q = (unsigned)((double)i * inv);
    const unsigned r = i - q * d;
    if ((int)r < 0) q--;
    else if (r >= d) q++;
    return q;
}

// Trapezoid weight of grid point i (Riemann :7981-8026): 1/2 per axis on
// which the point sits on the first or last lattice plane.
__device__ __forceinline__ double trapw(const BInst &in, unsigned i) {
    const unsigned EL = (unsigned)in.EL;
    double w = 1.;
    unsigned rem = i;
    for (int d = 0; d < in.D; d++) {
        const unsigned q = udiv(rem, EL, in.invEL), c = rem - q * EL;
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
    double thMix[kBS], muMix[kBS];  // theta and mu_eval of the iteration whose mixing is pending
    double bytes;   // bytes the fused sweeps move: every field access of this kernel, counted once
    double sbytes;  // the loop's stage bytes: (inputs + outputs) of each reference stage per field
    int SCcount, ExitSCcount, ExitSCloop, LinMix, PulMix, TrappedQ, warn, err, bis, pI, stall, nobr, prej;
};

template <typename T> struct Inst {
    const BInst &in;
    T *Env, *V, *OldV, *nin, *res;                 // nin/res: [scope + 1][S][G]
    double *sh;                                    // LDS scratch (kBT doubles)
    double *W;                                     // LDS work area [kWork] (Gram row, matrix, QR, coefficients)
    double floor_rel;                              // reduced-precision stop floor (0: the criterion as given)
    int stall_max;                                 // reduced-precision stall exit (0: none)
    double relacc;                                 // particle-number accuracy of the mu search
    unsigned lo, hi;                               // this workgroup's grid points [lo, hi) of the instance
    int TS, rank;                                  // team size and rank (TS = 1: the whole instance)
    TeamArea *tm;                                  // the team's exchange area
    unsigned gen;                                  // team reductions done so far
    int *err;
    __device__ long long SG() const { return (long long)in.S * in.G; }

    // Team reductions: block values (identical in every thread), then, for a
    // team, every member's values through the exchange area summed in rank
    // order, so every member gets the same bits.  The members are resident
    // workgroups running the same control flow; the wait is bounded anyway.
    template <int NV>
    __device__ __forceinline__ void team_combine(double (&v)[NV], int n, bool is_max, int first_max = 1 << 30) {
        if (TS == 1) return;
        const unsigned g = gen++;
        double(*part)[kRedMax] = tm->part[g & 1];
        if (threadIdx.x == 0) {
            for (int k = 0; k < n; k++)
                __hip_atomic_store(&part[rank][k], v[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_fetch_add(&tm->count, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned target = (g + 1) * (unsigned)TS;
            unsigned spins = 0;
            while (__hip_atomic_load(&tm->count, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
                __builtin_amdgcn_s_sleep(1);
                if (++spins > (1u << 27)) {
                    __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NV; k++)
            if (k < n) {
                double acc = __hip_atomic_load(&part[0][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                for (int r = 1; r < TS; r++) {
                    const double x = __hip_atomic_load(&part[r][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    acc = (is_max || k >= first_max) ? fmax(acc, x) : acc + x;
                }
                v[k] = uni(acc);
            }
    }
    template <int NV>
    __device__ __forceinline__ void tsum_n(double (&v)[NV], int n) {
        bsum_n<NV>(v, n, sh);
        team_combine<NV>(v, n, false);
    }
    // v[0], v[1] summed, v[2], v[3] maximised over the team, in one exchange
    __device__ __forceinline__ void tsum2_max2(double (&v)[4]) {
        double sv[2] = {v[0], v[1]};
        bsum_n<2>(sv, 2, sh);
        const double m2 = bmax(v[2], sh), m3 = bmax(v[3], sh);
        v[0] = sv[0];
        v[1] = sv[1];
        v[2] = m2;
        v[3] = m3;
        team_combine<4>(v, 4, false, 2);
    }
    __device__ __forceinline__ double tsum(double x) {
        double v[1] = {bsum(x, sh)};
        team_combine<1>(v, 1, false);
        return v[0];
    }
    __device__ __forceinline__ double tmax(double x) {
        double v[1] = {bmax(x, sh)};
        team_combine<1>(v, 1, true);
        return v[0];
    }

    // N_s(mu) of the TF density (GetTFdensity :4415-4456 + Riemann :7981-8026)
    __device__ __forceinline__ double evalN(int s, double mu) {
        const T *v = V + (long long)s * in.G;
        const double pref = in.pref;
        const bool d2 = in.D == 2;
        double acc = 0.;
        stream(
            lo, hi, [&](long long i) { return (double)v[i]; },
            [&](long long i, double vi) {
                double t = mu - vi;
                double val = t > 0. ? (d2 ? pref * t : pref * t * sqrt(t)) : 0.;
                acc += trapw(in, i) * val;
            });
        return tsum(acc) * in.trapD;
    }
    // N_s at K chemical potentials in one sweep over V.  Per value the per-lane
    // accumulation and the block and team reductions are those of evalN.
    template <int K>
    __device__ __forceinline__ void evalN_multi(int s, const double (&mus)[K], double (&out)[K]) {
        const T *v = V + (long long)s * in.G;
        const double pref = in.pref;
        const bool d2 = in.D == 2;
        double acc[K];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = 0.;
        stream(
            lo, hi, [&](long long i) { return (double)v[i]; },
            [&](long long i, double vi) {
                const double w = trapw(in, i);
#pragma unroll
                for (int k = 0; k < K; k++) {
                    double t = mus[k] - vi;
                    double val = t > 0. ? (d2 ? pref * t : pref * t * sqrt(t)) : 0.;
                    acc[k] += w * val;
                }
            });
        tsum_n<K>(acc, K);
#pragma unroll
        for (int k = 0; k < K; k++) out[k] = acc[k] * in.trapD;
    }
    __device__ __forceinline__ double N_at(Loop &L, int s, double mu) {
        L.muEval[s] = mu;
        L.bytes += (double)in.G * sizeof(T);
        L.sbytes += (double)in.G * sizeof(T);
        return evalN(s, mu);
    }
    __device__ __forceinline__ double N_eval(Loop &L, int s, double mu) {  // a density evaluation of the bisection
        L.bis++;
        return N_at(L, s, mu);
    }

    // EnforceConstraints (mpDPFT.cpp:4569-4743), the same step sequence as the
    // single-instance solver; the density of the last evaluated mu is kept.
    __device__ __forceinline__ void enforce(Loop &L) {
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
                if (!hiF[s] || !loF[s]) {  // :4657-4661: printed there, not a warning
                    L.mu[s] = L.PrevMu[s];
                    abortQ = true;
                    L.nobr++;
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
                    bool success = false, tooSmall = false, done = false;
                    // Speculation over kSpec decisions: one sweep evaluates N at the current mu (node 0) and
                    // at every mu the next kSpec - 1 halvings can reach (node j: N too small -> 2j + 1 at
                    // mu + step/2, too large -> 2j + 2 at mu - step/2, each with its step |mu' - mu|); the
                    // step-by-step loop then runs on those values and takes its mu and step from the tree, so
                    // it evaluates at the same points in the same order, and counts what it uses.
                    constexpr int kSpec = 3, kCand = (1 << kSpec) - 1;
                    double cm[kCand], cst[kCand], cN[kCand];
                    while (!done && !success && !tooSmall) {
                        cm[0] = L.mu[s];
                        cst[0] = step[s];
#pragma unroll
                        for (int j = 0; 2 * j + 2 < kCand; j++) {
                            cm[2 * j + 1] = cm[j] + 0.5 * cst[j];
                            cst[2 * j + 1] = fabs(cm[2 * j + 1] - cm[j]);
                            cm[2 * j + 2] = cm[j] - 0.5 * cst[j];
                            cst[2 * j + 2] = fabs(cm[2 * j + 2] - cm[j]);
                        }
                        evalN_multi<kCand>(s, cm, cN);
                        L.bytes += (double)in.G * sizeof(T);
                        int j = 0;
                        for (int lvl = 0; lvl < kSpec; lvl++) {
                            success = true;
                            L.bis++;
                            L.sbytes += (double)in.G * sizeof(T);
                            L.muEval[s] = cm[j];
                            L.TmpAb[s] = cN[j];
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
                            if (trapped) {
                                done = true;
                                break;
                            }
                            int child = -1;
                            for (int sg = -1; sg < 2; sg += 2) {
                                if ((double)sg * (L.TmpAb[s] - target) / fabs(2. * target) > relacc) {
                                    // the child's mu and step: from the tree, or past its last level by
                                    // the same arithmetic
                                    const int c = sg < 0 ? 2 * j + 1 : 2 * j + 2;
                                    double nm, nst;
                                    if (c < kCand) {
                                        nm = cm[c];
                                        nst = cst[c];
                                    } else {
                                        nm = sg < 0 ? cm[j] + 0.5 * cst[j] : cm[j] - 0.5 * cst[j];
                                        nst = fabs(nm - cm[j]);
                                    }
                                    prev[s] = L.mu[s];
                                    L.mu[s] = nm;
                                    if (fabs(L.mu[s]) > 1.0e+20) {
                                        L.warn++;
                                        L.mu[s] = L.PrevMu[s];
                                        break;
                                    }
                                    step[s] = nst;
                                    if (fabs(step[s]) < kMPb) {
                                        L.warn++;
                                        tooSmall = true;
                                        break;
                                    }
                                    success = false;
                                    child = c;
                                }
                            }
                            if (success || tooSmall || child >= kCand) break;
                            j = child;
                        }
                    }
                }
                L.RelAcc = relacc;
            }
        }
        for (int s = 0; s < S; s++)
            if (!isfinite(L.mu[s])) L.mu[s] = L.PrevMu[s];
    }

    // TF density at one point (GetTFdensity :4415-4456): the value that enters
    // the particle-number integral, and the value a stored density field holds
    __device__ __forceinline__ double tfval(double mu, double v) const {
        const double t = mu - v;
        return t > 0. ? (in.D == 2 ? in.pref * t : in.pref * t * sqrt(t)) : 0.;
    }
    __device__ __forceinline__ double tfT(double mu, double v) const { return (double)(T)tfval(mu, v); }

    // Pulay combination of NP history entries (element offsets off[q] of
    // their slots) into the visit's accumulators: acc += c_q (n_in,q + theta
    // R_q).  Every load of the chunk is issued before the first use.
    template <int NS, int NP>
    __device__ __forceinline__ void mix_chunk(double (&acc)[kUA][kBS], const unsigned (&ii)[kUA], const unsigned *off,
                                              const double *c, double th0, double th1) const {
        const unsigned G = (unsigned)in.G;
        T xn[NP][kUA][NS], xr[NP][kUA][NS];
#pragma unroll
        for (int q = 0; q < NP; q++)
#pragma unroll
            for (int u = 0; u < kUA; u++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    xn[q][u][s] = nin[off[q] + s * G + ii[u]];
                    xr[q][u][s] = res[off[q] + s * G + ii[u]];
                }
#pragma unroll
        for (int q = 0; q < NP; q++)
#pragma unroll
            for (int u = 0; u < kUA; u++)
#pragma unroll
                for (int s = 0; s < NS; s++)
                    acc[u][s] += c[q] * ((double)xn[q][u][s] + (s == 0 ? th0 : th1) * (double)xr[q][u][s]);
    }
    template <int NS>
    __device__ __forceinline__ void mix_any(int np, double (&acc)[kUA][kBS], const unsigned (&ii)[kUA],
                                            const unsigned *off, const double *c, double th0, double th1) const {
        switch (np) {
        case 1: mix_chunk<NS, 1>(acc, ii, off, c, th0, th1); break;
        case 2: mix_chunk<NS, 2>(acc, ii, off, c, th0, th1); break;
        case 3: mix_chunk<NS, 3>(acc, ii, off, c, th0, th1); break;
        case 4: mix_chunk<NS, 4>(acc, ii, off, c, th0, th1); break;
        case 5: mix_chunk<NS, 5>(acc, ii, off, c, th0, th1); break;
        case 6: mix_chunk<NS, 6>(acc, ii, off, c, th0, th1); break;
        case 7: mix_chunk<NS, 7>(acc, ii, off, c, th0, th1); break;
        case 8: mix_chunk<NS, 8>(acc, ii, off, c, th0, th1); break;
        default: break;
        }
    }

    // Sweep A of an iteration, kUA grid points per lane visit (both species
    // of a point in the same lane): the previous iteration's mixed density m
    // (AdmixDensities :4921-4992: none at the first iteration, where m is the
    // initial density; linear theta n + (1 - theta) m_old with n = TF(V_old,
    // mu of that iteration); or the Pulay combination sum_p c_p (n_in,p +
    // theta R_p) over the history entries from the ring head), written to
    // the history slot mnew unless mix == 0; then -- unless this is the
    // loop's exit -- UpdateV (:4815-4838) V_s = incV beta m_{s'} + Env_s, the
    // particle numbers N_s(mu_s) of the TF density of V (GetDensities' first
    // evaluation, Riemann :7981-8026) and max|V - V_old|, max|V|
    // (CheckConvergence in potential mode).  Every load of a visit is issued
    // before its values are used (the tail visit clamps its indices), so a
    // lane keeps all of them in flight.
    template <int NS>
    __device__ __forceinline__ void sweepA_t(Loop &L, int mix, int mold_slot, int mnew_slot, int head, const double *coef,
                             bool update, double *N, double &dvmax, double &vmax) {
        const unsigned G = (unsigned)in.G;
        const long long SGn = SG();
        const int nslot = in.scope + 1, pI = L.pI;
        const T *mold = nin + (long long)mold_slot * SGn;
        T *mnew = nin + (long long)mnew_slot * SGn;
        const bool pair = in.contact && NS == 2;
        const double beta = in.beta, incV = L.incV, th0 = L.thMix[0], th1 = L.thMix[kBS - 1];
        const bool need_vold = update || mix == 1;
        const int np0 = mix == 2 ? min(pI, kMixC) : 0, np1 = mix == 2 ? pI - np0 : 0;
        unsigned off0[kMixC], off1[kMixC];
#pragma unroll
        for (int q = 0; q < kMixC; q++) {
            off0[q] = (unsigned)(((head + q) % nslot) * SGn);
            off1[q] = (unsigned)(((head + kMixC + q) % nslot) * SGn);
        }
        double n0 = 0., n1 = 0., dv = 0., vm = 0.;
        for (unsigned i0 = lo + threadIdx.x; i0 < hi; i0 += kUA * kBT) {
            unsigned ii[kUA];
            bool ok[kUA];
#pragma unroll
            for (int u = 0; u < kUA; u++) {
                const unsigned i = i0 + u * kBT;
                ok[u] = i < hi;
                ii[u] = ok[u] ? i : hi - 1;
            }
            double vo[kUA][kBS] = {}, m[kUA][kBS] = {}, env[kUA][kBS] = {};
            if (need_vold) {
#pragma unroll
                for (int u = 0; u < kUA; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++) vo[u][s] = (double)OldV[s * G + ii[u]];
            }
            if (mix != 2) {
#pragma unroll
                for (int u = 0; u < kUA; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++) m[u][s] = (double)mold[s * G + ii[u]];
            }
            if (update) {
#pragma unroll
                for (int u = 0; u < kUA; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++) env[u][s] = (double)Env[s * G + ii[u]];
            }
            if (mix == 1) {
#pragma unroll
                for (int u = 0; u < kUA; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++) {
                        const double th = s == 0 ? th0 : th1;
                        m[u][s] = (double)(T)(th * tfT(L.muMix[s], vo[u][s]) + (1. - th) * m[u][s]);
                    }
            } else if (mix == 2) {
                double acc[kUA][kBS] = {};
                mix_any<NS>(np0, acc, ii, off0, coef, th0, th1);
                if (np1 > 0) mix_any<NS>(np1, acc, ii, off1, coef + kMixC, th0, th1);
#pragma unroll
                for (int u = 0; u < kUA; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++) m[u][s] = (double)(T)acc[u][s];
            }
            if (mix != 0) {
#pragma unroll
                for (int u = 0; u < kUA; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++)
                        if (ok[u]) mnew[s * G + ii[u]] = (T)m[u][s];
            }
            if (update) {
#pragma unroll
                for (int u = 0; u < kUA; u++) {
                    if (!ok[u]) continue;
                    const double w = trapw(in, ii[u]);
#pragma unroll
                    for (int s = 0; s < NS; s++) {
                        double acc = 0.;
                        if (in.contact) acc += (pair ? m[u][1 - s] : 0.) * beta;
                        if (incV < 1.) acc *= incV;
                        const T vT = (T)(acc + env[u][s]);
                        V[s * G + ii[u]] = vT;
                        const double v = (double)vT;
                        const double nv = w * tfval(L.mu[s], v);
                        if (s == 0) n0 += nv;
                        else n1 += nv;
                        dv = fmax(dv, fabs(v - vo[u][s]));
                        vm = fmax(vm, fabs(v));
                    }
                }
            }
        }
        // algorithmic bytes of the sweep
        double fields = 0.;
        if (mix == 1) fields += 2.;                       // V_old, m_old
        else if (mix == 2) fields += 2. * (double)pI;    // n_in and R of every entry
        else if (update) fields += 1.;                   // m (first iteration)
        if (mix != 0) fields += 1.;                      // m written
        if (update) fields += (mix == 1 ? 0. : 1.) + 2.;  // V_old (if not read above), Env, V written
        L.bytes += fields * (double)SGn * sizeof(T);
        if (update) {
            double v[4] = {n0, n1, dv, vm};
            tsum2_max2(v);
            N[0] = v[0] * in.trapD;
            N[1] = v[1] * in.trapD;
            dvmax = v[2];
            vmax = v[3];
        }
    }
    __device__ __forceinline__ void sweepA(Loop &L, int mix, int mold_slot, int mnew_slot, int head, const double *coef,
                           bool update, double *N, double &dvmax, double &vmax) {
        if (in.S == 2) sweepA_t<2>(L, mix, mold_slot, mnew_slot, head, coef, update, N, dvmax, vmax);
        else sweepA_t<1>(L, mix, mold_slot, mnew_slot, head, coef, update, N, dvmax, vmax);
    }

    // Sweep B after the mu search, kUB grid points per lane visit: the density
    // n = TF(V, mu_eval) (not stored), the density-mode overlaps <n,n>, <m,m>,
    // <n,m> per species (CheckConvergence :4759-4782 with RelDev :4784-4802)
    // and, when Pulay mixing runs, the push of (n_in = m, R = n - m) into
    // history slot ps (m already lives there) with the new Gram row
    // <R_new, R_j> over NC retained entries (PulayMixer :5005-5167); every
    // load of a visit (V, m and the NC stored residuals) is issued first.  With
    // more than kGC retained entries the remaining chunks of kGC re-read R_new
    // (FIRST false).  acc: [0, kGC) the row, kGC <R_new, R_new>, then the
    // overlaps.
    template <int NS, int NC, bool FIRST>
    __device__ __forceinline__ void sweepB_t(Loop &L, bool push, const T *m, T *r, const unsigned *offj, double *acc) {
        const unsigned G = (unsigned)in.G;
        const bool dens_mode = in.convmode != 1;
        for (unsigned i0 = lo + threadIdx.x; i0 < hi; i0 += kUB * kBT) {
            unsigned ii[kUB];
            bool ok[kUB];
#pragma unroll
            for (int u = 0; u < kUB; u++) {
                const unsigned i = i0 + u * kBT;
                ok[u] = i < hi;
                ii[u] = ok[u] ? i : hi - 1;
            }
            T vv[kUB][NS], mv[kUB][NS], x[NC > 0 ? NC : 1][kUB][NS];
#pragma unroll
            for (int u = 0; u < kUB; u++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    if (FIRST) {
                        vv[u][s] = V[s * G + ii[u]];
                        mv[u][s] = m[s * G + ii[u]];
                    } else
                        vv[u][s] = r[s * G + ii[u]];
                }
#pragma unroll
            for (int j = 0; j < NC; j++)
#pragma unroll
                for (int u = 0; u < kUB; u++)
#pragma unroll
                    for (int s = 0; s < NS; s++) x[j][u][s] = res[offj[j] + s * G + ii[u]];
#pragma unroll
            for (int u = 0; u < kUB; u++)
#pragma unroll
                for (int s = 0; s < NS; s++) {
                    double rv;
                    if (FIRST) {
                        const double nv = tfT(L.muEval[s], (double)vv[u][s]), mm = (double)mv[u][s];
                        const T rT = (T)(nv - mm);
                        rv = (double)rT;
                        if (push && ok[u]) r[s * G + ii[u]] = rT;
                        if (dens_mode && ok[u]) {
                            acc[kGC + 1 + 3 * s] += nv * nv;
                            acc[kGC + 2 + 3 * s] += mm * mm;
                            acc[kGC + 3 + 3 * s] += nv * mm;
                        }
                        if (push && ok[u]) acc[kGC] += rv * rv;
                    } else
                        rv = (double)vv[u][s];
                    if (ok[u])
#pragma unroll
                        for (int j = 0; j < NC; j++) acc[j] += rv * (double)x[j][u][s];
                }
        }
    }
    template <int NS, bool FIRST>
    __device__ __forceinline__ void sweepB_nc(int nc, Loop &L, bool push, const T *m, T *r, const unsigned *offj, double *acc) {
        switch (nc) {
        case 0: sweepB_t<NS, 0, FIRST>(L, push, m, r, offj, acc); break;
        case 1: sweepB_t<NS, 1, FIRST>(L, push, m, r, offj, acc); break;
        case 2: sweepB_t<NS, 2, FIRST>(L, push, m, r, offj, acc); break;
        case 3: sweepB_t<NS, 3, FIRST>(L, push, m, r, offj, acc); break;
        case 4: sweepB_t<NS, 4, FIRST>(L, push, m, r, offj, acc); break;
        case 5: sweepB_t<NS, 5, FIRST>(L, push, m, r, offj, acc); break;
        case 6: sweepB_t<NS, 6, FIRST>(L, push, m, r, offj, acc); break;
        case 7: sweepB_t<NS, 7, FIRST>(L, push, m, r, offj, acc); break;
        default: sweepB_t<NS, 8, FIRST>(L, push, m, r, offj, acc); break;
        }
    }
    // Results: row[0..I-1] (thread 0), ov[3 s + {0,1,2}] (every thread).
    __device__ __forceinline__ void sweepB(Loop &L, bool push, int ps, int head, int I, double *row, double *ov) {
        const long long SGn = SG();
        const int nslot = in.scope + 1, nOld = push ? I - 1 : 0;
        const bool dens_mode = in.convmode != 1;
        const T *m = nin + (long long)ps * SGn;
        T *r = res + (long long)ps * SGn;
        for (int j0 = 0; j0 == 0 || j0 < nOld; j0 += kGC) {
            const int nc = min(kGC, nOld - j0);
            unsigned offj[kGC];
#pragma unroll
            for (int j = 0; j < kGC; j++) offj[j] = (unsigned)(((head + j0 + j) % nslot) * SGn);
            double acc[kGC + 1 + 3 * kBS];
#pragma unroll
            for (int j = 0; j < kGC + 1 + 3 * kBS; j++) acc[j] = 0.;
            if (j0 == 0) {
                if (in.S == 2) sweepB_nc<2, true>(nc, L, push, m, r, offj, acc);
                else sweepB_nc<1, true>(nc, L, push, m, r, offj, acc);
            } else {
                if (in.S == 2) sweepB_nc<2, false>(nc, L, push, m, r, offj, acc);
                else sweepB_nc<1, false>(nc, L, push, m, r, offj, acc);
            }
            if (j0 == 0) {
                tsum_n<kGC + 1 + 3 * kBS>(acc, kGC + 1 + 3 * kBS);
                if (threadIdx.x == 0) {
                    for (int j = 0; j < nc; j++) row[j] = acc[j];
                    if (push) row[nOld] = acc[kGC];
                }
                for (int k = 0; k < 3 * kBS; k++) ov[k] = acc[kGC + 1 + k];
                // V and m read; R written and the stored residuals read when pushing
                L.bytes += ((push ? 1. + (double)nc : 0.) + ((push || dens_mode) ? 2. : 0.)) * (double)SGn * sizeof(T);
            } else {
                tsum_n<kGC + 1 + 3 * kBS>(acc, nc);
                if (threadIdx.x == 0)
                    for (int j = 0; j < nc; j++) row[j0 + j] = acc[j];
                L.bytes += (1. + (double)nc) * (double)SGn * sizeof(T);
            }
        }
    }

    // CheckConvergence (:4759-4782): the exit counter from this iteration's
    // CC (potential mode: max|V - V_old|; density mode: max_s RelDev weighted
    // by the species' share of N)
    __device__ __forceinline__ void convergence(Loop &L, double dvmax, double vmax, const double *ov) {
        const double mp = 100. * 1.0e-16;
        if (L.SCcount > 0) {
            L.CC = 0.;
            double tot = 0.;
            for (int s = 0; s < in.S; s++) tot += L.TmpAb[s];
            double crit = in.sccrit;
            if (in.convmode == 1) {
                L.CC = dvmax;
                if (floor_rel > 0.) crit = fmax(crit, floor_rel * vmax);
            } else {
                double maxtest = 0.;
                for (int s = 0; s < in.S; s++) {
                    const double aa = ov[3 * s], bb = ov[3 * s + 1], ab = ov[3 * s + 2];
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

};

// One launch: each workgroup takes instances from the queue until it is empty.
// Per SC iteration of an instance (GetData :210-225) the fields are swept
// twice, plus one sweep per mu-bisection density evaluation: sweep A (the
// previous iteration's mixing, UpdateV, N at the current mu, potential-mode
// convergence), the mu search (EnforceConstraints), sweep B (the density,
// density-mode convergence, Pulay push and Gram row).  The density n of an
// iteration is never stored: it is TF(V, mu_eval), evaluated where it is
// used; the mixed density m lives in the Pulay history slot that the next
// push takes as its n_in.
template <typename T>
__global__ __launch_bounds__(kBT) void k_batch_solve(BArgs<T> a) {
    __shared__ double sh[kBT];
    __shared__ double Gcache[(kMaxScope + 1) * (kMaxScope + 1)];
    __shared__ double Work[kWork];
    __shared__ int nextIdx;
    for (;;) {
        __syncthreads();
        if (threadIdx.x == 0) nextIdx = atomicAdd(a.next, 1);
        __syncthreads();
        const int tix = __builtin_amdgcn_readfirstlane(nextIdx);
        if (tix >= a.n) break;
        // the ticket names the instance (largest first), this workgroup's rank in its
        // team and the team size; uniform, so the instance's parameters sit in scalar registers
        const int k = __builtin_amdgcn_readfirstlane(a.tickets[tix].inst);
        const int rank = __builtin_amdgcn_readfirstlane(a.tickets[tix].rank);
        const int TS = __builtin_amdgcn_readfirstlane(a.tickets[tix].size);
        const unsigned long long tstart = __builtin_amdgcn_s_memrealtime();
        const BInst &in = a.inst[k];
        const long long SGn = (long long)in.S * in.G, G = in.G;
        const int S = in.S, scope = in.scope, nslot = scope + 1, ld = scope + 1;
        Inst<T> I{in};
        I.TS = TS;
        I.rank = rank;
        I.tm = a.team + k;
        I.gen = 0;
        I.err = a.team_err;
        I.sh = sh;
        // a team works in its rank-0 member's field slot, each member on its own
        // contiguous, 64-aligned share of the grid points
        int slot_wg = blockIdx.x;
        if (TS > 1) {
            double v[1] = {rank == 0 ? (double)blockIdx.x : 0.};
            I.template team_combine<1>(v, 1, false);
            slot_wg = (int)v[0];
        }
        T *base = a.slots + (long long)slot_wg * a.slot_stride;
        I.lo = rank == 0 ? 0u : ((unsigned)(G * rank / TS) & ~63u);
        I.hi = rank == TS - 1 ? (unsigned)G : ((unsigned)(G * (rank + 1) / TS) & ~63u);
        I.Env = base;
        I.V = base + SGn;
        I.OldV = base + 2 * SGn;
        I.nin = base + 3 * SGn;
        I.res = I.nin + (long long)nslot * SGn;
        I.W = Work;
        I.floor_rel = a.floor_rel;
        I.stall_max = a.stall_max;
        I.relacc = fmax(in.RelAcc, a.relacc_floor);
        double *outDen = a.fields + in.out_off, *outV = outDen + SGn;
        double *row = Work, *A = Work + (kMaxScope + 1), *coef = Work + kWork - 2 - (kMaxScope + 1);
        Loop L;
        for (int s = 0; s < S; s++) {
            L.mu[s] = a.polish ? a.out[k].mu[s] : in.mu0[s];
            L.theta[s] = in.theta0[s];
            L.TmpAb[s] = L.OldAb[s] = L.OldMu[s] = L.PrevMu[s] = 0.;
            L.muEval[s] = L.muMix[s] = L.mu[s];
            L.thMix[s] = L.theta[s];
        }
        L.DMM = a.polish ? a.out[k].DMM : in.DMM0;
        L.RelAcc = I.relacc;
        L.incV = in.incV0;
        L.CC = L.OldCC = 0.;
        L.best = 1.0e+300;
        L.SCcount = L.ExitSCcount = L.ExitSCloop = L.LinMix = L.PulMix = L.TrappedQ = 0;
        L.warn = L.err = L.bis = L.pI = L.stall = L.nobr = L.prej = 0;
        // InitializeEnvironments (:6362-6390): Env, and V <- Env (or the stored start); the
        // particle numbers at the initial mu in the same sweep
        double n0 = 0., n1 = 0.;
        for (long long i = I.lo + threadIdx.x; i < I.hi; i += kBT) {
            double r[3];
            coords(in, i, r);
            const double w = trapw(in, i);
            for (int s = 0; s < S; s++) {
                const long long idx = s * G + i;
                double v = 0.;
                if (in.env[s] == 1) v = in.c0[s] * pow(r[0] * r[0] + r[1] * r[1] + r[2] * r[2], 1.);
                else if (in.env[s] == 3) v = in.c0[s] * ((1. - in.c1[s]) * r[0] * r[0] + (1. + in.c1[s]) * r[1] * r[1]);
                I.Env[idx] = (T)v;
                const T vT = (T)(a.polish ? a.start[in.out_off + SGn + idx] : v);
                I.V[idx] = vT;
                const double nv = w * I.tfval(L.mu[s], (double)vT);
                if (s == 0) n0 += nv;
                else n1 += nv;
            }
        }
        L.bytes = 2. * (double)SGn * sizeof(T) + (a.polish ? (double)SGn * 8. : 0.);
        L.sbytes = L.bytes + (double)SGn * sizeof(T);  // Env, V written; N at the initial mu
        {
            double v[3] = {n0, n1, 0.};
            I.template tsum_n<3>(v, 2);
            for (int s = 0; s < S; s++) {
                L.TmpAb[s] = v[s] * in.trapD;
                L.bytes += (double)G * sizeof(T);  // counted as a density evaluation
            }
        }
        // initial densities (:194-201): EnforceConstraints, then m = n(V, mu_eval) into slot 0
        I.enforce(L);
        int head = 0, mslot = 0;
        {
            T *m = I.nin;
            for (long long i = I.lo + threadIdx.x; i < I.hi; i += kBT)
                for (int s = 0; s < S; s++) m[s * G + i] = (T)I.tfval(L.muEval[s], (double)I.V[s * G + i]);
            L.bytes += 2. * (double)SGn * sizeof(T);
            L.sbytes += 2. * (double)SGn * sizeof(T);  // initial density: V in, n out
        }
        __syncthreads();
        int mix = 0, mold = 0;  // the pending mixing of the previous iteration (0: none)
        unsigned long long tph[4] = {0, 0, 0, 0};
        while (!L.ExitSCloop && in.maxsc > 0) {
            // StoreCurrentIteration (:4804): swap the potential buffers
            T *t = I.V;
            I.V = I.OldV;
            I.OldV = t;
            L.OldCC = L.CC;
            for (int s = 0; s < S; s++) {
                L.OldMu[s] = L.mu[s];
                L.OldAb[s] = L.TmpAb[s];
            }
            // sweep A: mixing of the previous iteration + UpdateV + N(mu) + max|dV|
            double N[kBS] = {0., 0.}, dvmax = 0., vmax = 0.;
            unsigned long long c0 = __builtin_amdgcn_s_memrealtime();
            I.sweepA(L, mix, mold, mslot, head, coef, true, N, dvmax, vmax);
            // stage bytes: UpdateV (Env and the other species' density in, V out), N at the current mu
            L.sbytes += ((in.contact && S == 2) ? 3. : 2.) * (double)SGn * sizeof(T) + (double)SGn * sizeof(T);
            unsigned long long c1 = __builtin_amdgcn_s_memrealtime();
            tph[0] += c1 - c0;
            if (L.SCcount > 0) L.incV += in.incV0;
            L.incV = fmin(1., L.incV);
            // GetDensities (:3306-3316): N at the current mu, EnforceConstraints
            for (int s = 0; s < S; s++) {
                L.TmpAb[s] = N[s];
                L.muEval[s] = L.mu[s];
            }
            I.enforce(L);
            c0 = __builtin_amdgcn_s_memrealtime();
            tph[1] += c0 - c1;
            // Pulay bookkeeping ahead of the sweep (PulayMixer :5005-5050): the new entry takes
            // slot mslot (its n_in, the mixed density, is already there); a full history drops
            // its oldest entry
            const bool tryPulay = in.mixer > 0 && L.SCcount > 0;
            int pI = L.pI, nhead = head, shift = 0;
            if (tryPulay) {
                pI++;
                if (pI > scope) {
                    nhead = (head + 1) % nslot;
                    pI--;
                    shift = 1;
                }
            }
            double ov[3 * kBS];
            I.sweepB(L, tryPulay, mslot, nhead, pI, row, ov);
            L.sbytes += 4. * (double)SGn * sizeof(T);  // density (V in, n out); convergence (two fields in)
            c1 = __builtin_amdgcn_s_memrealtime();
            tph[2] += c1 - c0;
            I.convergence(L, dvmax, vmax, ov);
            if (L.SCcount == in.maxsc) L.ExitSCloop = 1;  // Intercept
            // AdmixDensities (:4921-4992): the mixing itself runs in the next sweep A
            double nA = 0., nD = 0.;
            for (int s = 0; s < S; s++) {
                nD += (in.Ab[s] - L.TmpAb[s]) * (in.Ab[s] - L.TmpAb[s]);
                nA += in.Ab[s] * in.Ab[s];
            }
            double rel = sqrt(nD) / sqrt(nA), mod = 1.;
            if (rel > kMPb) mod = fmin(1., pow(fabs(1. - rel), 30.));
            if (L.TrappedQ)
                for (int s = 0; s < S; s++) L.theta[s] *= mod;
            for (int s = 0; s < S; s++) {
                L.thMix[s] = L.theta[s];
                L.muMix[s] = L.muEval[s];
            }
            bool ok = false;
            if (tryPulay && L.CC < 1.0e+300) {
                head = nhead;
                L.pI = pI;
                // the Gram matrix from the cache (shifted when the oldest entry was dropped) and
                // the coefficients, by thread 0; flag and coefficients broadcast through LDS
                if (threadIdx.x == 0) {
                    for (int i = 0; i < pI; i++)
                        for (int j = 0; j < pI; j++) {
                            double v;
                            if (i == pI - 1) v = row[j];
                            else if (j == pI - 1) v = row[i];
                            else v = Gcache[(i + shift) * ld + (j + shift)];
                            A[i * pI + j] = v;
                        }
                    for (int i = 0; i < pI; i++)
                        for (int j = 0; j < pI; j++) Gcache[i * ld + j] = A[i * pI + j];
                    Work[kWork - 1] =
                        Inst<T>::solve_coeff(pI, A, Work + 2 * (kMaxScope + 1) * (kMaxScope + 1), coef) ? 1. : 0.;
                }
                __syncthreads();
                ok = Work[kWork - 1] != 0.;
                if (!ok) L.prej++;
                tph[3] += __builtin_amdgcn_s_memrealtime() - c1;
                mold = mslot;
                mslot = (head + L.pI) % nslot;  // the free slot takes the next mixed density
            } else
                mold = mslot;  // linear mixing in place
            if (ok) L.PulMix++;
            else L.LinMix++;
            // stage bytes of AdmixDensities: Gram row (I + 1 fields) and Pulay combination (2I + 1),
            // or linear mixing (n, n_old in, n out)
            L.sbytes += (tryPulay && L.CC < 1.0e+300 ? (double)(L.pI + 1) : 0.) * (double)SGn * sizeof(T) +
                        (ok ? (double)(2 * L.pI + 1) : 3.) * (double)SGn * sizeof(T);
            mix = ok ? 2 : 1;
            if (L.TrappedQ)
                for (int s = 0; s < S; s++) L.theta[s] *= 1. / mod;
            L.SCcount++;  // PrepareNextIteration
        }
        // the last iteration's mixing (the loop exits after AdmixDensities)
        if (mix != 0) {
            double N[kBS], dvmax, vmax;
            T *t = I.V;  // sweep A reads the mixing's V_old from OldV
            I.V = I.OldV;
            I.OldV = t;
            I.sweepA(L, mix, mold, mslot, head, coef, false, N, dvmax, vmax);
            t = I.V;
            I.V = I.OldV;
            I.OldV = t;
        }
        __syncthreads();
        const T *Den = I.nin + (long long)mslot * SGn;  // the mixed density
        // energies (GetEnergy :5414-5436): contact, environment, TF dispersal, AuxEkin[0]
        double Eint = 0.;
        if (in.contact && S > 1) {
            double acc = 0.;
            for (long long i = I.lo + threadIdx.x; i < I.hi; i += kBT)
                acc += trapw(in, i) * (double)Den[i] * (double)Den[G + i];
            Eint = in.beta * I.tsum(acc) * in.trapD;
        }
        BOut o{};
        double Etot = Eint;
        const double DD = (double)in.D;
        for (int s = 0; s < S; s++) {
            const T *n = Den + (long long)s * G, *v = I.V + (long long)s * G, *e = I.Env + (long long)s * G;
            double ee = 0., ed = 0., ea = 0.;
            for (long long i = I.lo + threadIdx.x; i < I.hi; i += kBT) {
                double w = trapw(in, i), nn = (double)n[i];
                ee += w * (double)e[i] * nn;
                double tt = L.mu[s] - (double)v[i];
                ed += w * (tt > 0. ? pow(tt, (DD + 2.) / 2.) : 0.);
                ea += w * pow(nn > 0. ? nn : 0., (DD + 2.) / DD);
            }
            double v3[3] = {ee, ed, ea};
            I.template tsum_n<3>(v3, 3);
            ee = v3[0] * in.trapD;
            ed = v3[1] * in.trapD;
            ea = v3[2] * in.trapD;
            double aux0 = in.units > 0 ? in.aux0pref * ea : 0.;
            double disp = (in.ekin0 ? in.dispPref * ed : 0.) + (in.ekin1 ? aux0 : 0.);
            o.Eenv[s] = ee;
            o.Edisp[s] = disp;
            o.Aux0[s] = aux0;
            Etot += disp + ee;
        }
        L.bytes += ((in.contact && S > 1) ? 2. * (double)G : 0.) * sizeof(T) + 3. * (double)SGn * sizeof(T);
        L.sbytes += ((in.contact && S > 1) ? 2. * (double)G : 0.) * sizeof(T) + 3. * (double)SGn * sizeof(T) +
                    4. * (double)SGn * sizeof(T);  // energies; retire as a copy (two fields in, two out)
        // retire: fields and results out
        L.bytes += 2. * (double)SGn * sizeof(T) + 2. * (double)SGn * 8.;
        for (long long i = I.lo + threadIdx.x; i < I.hi; i += kBT)
            for (int s = 0; s < S; s++) {
                outDen[s * G + i] = (double)Den[s * G + i];
                outV[s * G + i] = (double)I.V[s * G + i];
            }
        if (TS > 1) {  // the rank-0 slot is free again only when every member has copied its share out
            double v[1] = {0.};
            I.template tsum_n<1>(v, 1);
        }
        if (threadIdx.x == 0 && rank == 0) {
            o.SCcount = L.SCcount;
            o.converged = L.ExitSCcount >= 3;
            o.LinMix = L.LinMix;
            o.PulMix = L.PulMix;
            o.errorCount = L.err;
            o.warningCount = L.warn + (in.maxsc > 0 && L.SCcount >= in.maxsc ? 1 : 0);
            o.bisection = L.bis;
            o.bracket_failures = L.nobr;
            o.pulay_rejections = L.prej;
            o.CC = L.CC;
            o.DMM = L.DMM;
            o.Etot = Etot;
            o.Eint = Eint;
            o.bytes = L.bytes;
            o.sbytes = L.sbytes;
            o.t0 = tstart;
            o.t1 = __builtin_amdgcn_s_memrealtime();
            for (int q = 0; q < 4; q++) o.tph[q] = tph[q];
            for (int s = 0; s < S; s++) {
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
    int *d_next = nullptr, *d_team_err = nullptr;
    Ticket *d_tickets = nullptr;
    TeamArea *d_team = nullptr;
    int n_tickets = 0, n_teams = 0;
    hipEvent_t ev[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [phase][start, stop] around k_batch_solve
    double kernel_ms[2] = {0., 0.}, kernel_bytes[2] = {0., 0.}, stage_bytes[2] = {0., 0.};
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
static int prepare(char (&msg)[256], const mpdpft_params &p, BInst &o, int i) {
    auto bad = [&](const char *m) {
        std::snprintf(msg, sizeof msg, "instance %d: %s", i, m);
        return MPDPFT_EINVAL;
    };
    if (const char *why = params_structure_error(&p)) return bad(why);
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
    o.invEL = 1. / (double)o.EL;
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

// Queue order and teams, shared by both solve entry points.  Order: largest
// instance first (S G decides the cost of an SC iteration), ties by index, so
// the longest solves start at once and the short ones fill in behind them
// (MPDPFT_BATCH_ORDER=index: the instances' own order).  Teams: an instance of
// at least team_points grid points is solved by kTeamMax workgroups (its SC
// loop would otherwise outlast the rest of the batch on one CU).  A team's
// tickets are consecutive and tickets are taken in order by running
// workgroups, so at most the workgroup holding the newest ticket waits for a
// partner that is not yet running, and that partner starts as soon as any
// other workgroup retires an instance; with fewer than kTeamMax slots there
// are no teams.  order: n ints of work space; tickets: room for kTeamMax n.
// Returns the number of tickets.
static int make_tickets(const BInst *inst, int n, int nslots, int *order, Ticket *tickets, int *n_teams) {
    for (int i = 0; i < n; i++) order[i] = i;
    const char *ord = std::getenv("MPDPFT_BATCH_ORDER");
    if (!(ord && std::strcmp(ord, "index") == 0))
        std::sort(order, order + n, [&](int x, int y) {
            const long long cx = (long long)inst[x].S * inst[x].G, cy = (long long)inst[y].S * inst[y].G;
            return cx != cy ? cx > cy : x < y;
        });
    const char *tp = std::getenv("MPDPFT_BATCH_TEAM_POINTS");  // 0: no teams
    const long long team_points = tp ? std::atoll(tp) : 40000;
    int team_inst = 0;
    for (int i = 0; i < n; i++)
        if (team_points > 0 && inst[i].G >= team_points) team_inst++;
    const bool teams = team_inst > 0 && nslots >= kTeamMax;
    int nt = 0;
    for (int q = 0; q < n; q++) {
        const int i = order[q];
        const int ts = (teams && inst[i].G >= team_points) ? kTeamMax : 1;
        for (int r = 0; r < ts; r++) tickets[nt++] = Ticket{i, r, ts, 0};
    }
    *n_teams = teams ? team_inst : 0;
    return nt;
}

// Resident workgroups of the persistent kernel on one device (both field
// types), optionally capped by MPDPFT_BATCH_SLOTS.
static hipError_t resident_slots(int device, int *slots) {
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return e;
    int per_cu = 0, per_cu32 = 0;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(&k_batch_solve<double>),
                                                     kBT, 0);
    if (e != hipSuccess) return e;
    e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu32, reinterpret_cast<const void *>(&k_batch_solve<float>),
                                                     kBT, 0);
    if (e != hipSuccess) return e;
    per_cu = std::max(1, std::min(per_cu, per_cu32));
    const char *env = std::getenv("MPDPFT_BATCH_SLOTS");
    const int all = per_cu * prop.multiProcessorCount;
    *slots = env ? std::max(1, std::min(std::atoi(env), all)) : all;
    return hipSuccess;
}

// Values of T per slot (the slot arena is sized for doubles): Env, V, OldV and
// the Pulay history of n_in and residual per entry plus the mixed density.
static long long slot_values(int scope, long long Gmax) { return (3LL + 2LL * (scope + 1)) * kBS * Gmax; }

// Device buffers of one solve (allocated by mpdpft_batch_create, or carved
// from the static scratch array by the allocation-free entry point).
struct DevBuffers {
    BInst *inst;
    BOut *out;
    Ticket *tickets;
    TeamArea *team;
    int *next, *team_err;
    void *slots;
    double *fields, *start;
    int n, n_tickets, nslots, scope;
    long long Gmax;
};

// Zero fill of 32-bit words: the queue head, the team areas and the team
// error flag are cleared by a kernel on the solve's stream before each launch.
namespace {
__global__ void k_clear_words(unsigned *p, long long nwords) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nwords; i += (long long)gridDim.x * blockDim.x)
        p[i] = 0u;
}
}  // namespace

static void clear_words(void *p, size_t bytes, hipStream_t st) {
    const long long nw = (long long)(bytes / sizeof(unsigned));
    const int blocks = (int)std::min<long long>(1024, (nw + 255) / 256);
    hipLaunchKernelGGL(k_clear_words, dim3(std::max(1, blocks)), dim3(256), 0, st, (unsigned *)p, nw);
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

// One persistent launch over the whole queue on stream st; ev0/ev1 (may be
// null) are recorded around the kernel.
template <typename T>
static hipError_t launch(const DevBuffers &d, int polish, hipStream_t st, hipEvent_t ev0, hipEvent_t ev1) {
    BArgs<T> a;
    a.inst = d.inst;
    a.tickets = d.tickets;
    a.n = d.n_tickets;
    a.team = d.team;
    a.team_err = d.team_err;
    a.next = d.next;
    a.slots = (T *)d.slots;
    a.slot_stride = slot_values(d.scope, d.Gmax);
    a.Gmax = d.Gmax;
    a.fields = d.fields;
    a.out = d.out;
    a.polish = polish;
    a.start = d.start;
    a.floor_rel = sizeof(T) == 4 ? kF32Floor : 0.;
    a.stall_max = sizeof(T) == 4 ? kF32Stall : 0;
    a.relacc_floor = sizeof(T) == 4 ? kF32RelAcc : 0.;
    clear_words(d.next, sizeof(int), st);
    clear_words(d.team, sizeof(TeamArea) * (size_t)d.n, st);
    clear_words(d.team_err, sizeof(int), st);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    if (ev0 && (e = hipEventRecord(ev0, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(k_batch_solve<T>, dim3(d.nslots), dim3(kBT), 0, st, a);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, st)) != hipSuccess) return e;
    return hipSuccess;
}

// BOut of each instance -> the mpdpft_result mpdpft_solve reports for it.
static void fill_results(const BOut *out, const BInst *inst, const mpdpft_params *params, int n, double seconds,
                         mpdpft_result *res) {
    for (int i = 0; i < n; i++) {
        const BOut &o = out[i];
        mpdpft_result &r = res[i];
        std::memset(&r, 0, sizeof(r));
        r.S = inst[i].S;
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
        r.bracket_failures = o.bracket_failures;
        r.pulay_rejections = o.pulay_rejections;
        r.solve_seconds = seconds;
        const mpdpft_params &p = params[i];
        for (int s = 0; s < r.S; s++) {
            r.muVec[s] = o.mu[s];
            r.FinalAbundances[s] = o.N[s];
            r.DispersalEnergies[s] = o.Edisp[s];
            r.EnvironmentalEnergies[s] = o.Eenv[s];
            r.AuxEkin[s][0] = o.Aux0[s];
        }
        for (int a = 0; a < p.n_interactions; a++) r.InteractionEnergies[a] = p.Interactions[a] == 1 ? o.Eint : 0.;
    }
}

extern "C" int mpdpft_batch_create(const mpdpft_params *p, int n, int device, int flags, mpdpft_batch **out) {
    if (!p || n < 1 || !out) return MPDPFT_EINVAL;
    mpdpft_batch *b = new mpdpft_batch();
    *out = b;
    b->n = n;
    b->flags = flags;
    b->inst.resize(n);
    b->params.assign(p, p + n);
    long long off = 0;
    for (int i = 0; i < n; i++) {  // every instance is checked before the device is touched
        char msg[256];
        int rc = prepare(msg, p[i], b->inst[i], i);
        if (rc) return bfail(b, rc, msg);
        b->inst[i].out_off = off;
        off += 2LL * b->inst[i].S * b->inst[i].G;
        b->Gmax = std::max(b->Gmax, b->inst[i].G);
        b->scope = std::max(b->scope, b->inst[i].scope);
    }
    b->arena = off;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return bfail(b, MPDPFT_ENODEV, "no GPU visible");
    b->device = device < 0 ? 0 : device;
    BHIP(hipSetDevice(b->device));
    // slots: the workgroups resident at once (a persistent launch larger than that only adds
    // workgroups that start after the queue is empty), at most kTeamMax per instance
    int want = 0;
    BHIP(resident_slots(b->device, &want));
    b->nslots = std::max(1, std::min(n * kTeamMax, want));
    BHIP(hipStreamCreate(&b->st));
    BHIP(hipMalloc((void **)&b->d_inst, sizeof(BInst) * n));
    BHIP(hipMalloc((void **)&b->d_out, sizeof(BOut) * n));
    BHIP(hipMalloc((void **)&b->d_next, sizeof(int)));
    BHIP(hipMalloc((void **)&b->d_fields, sizeof(double) * (size_t)off));
    BHIP(hipMalloc(&b->d_slots, sizeof(double) * (size_t)slot_values(b->scope, b->Gmax) * b->nslots));
    BHIP(hipMemcpy(b->d_inst, b->inst.data(), sizeof(BInst) * n, hipMemcpyHostToDevice));
    std::vector<int> order(n);
    std::vector<Ticket> tickets((size_t)kTeamMax * n);
    b->n_tickets = make_tickets(b->inst.data(), n, b->nslots, order.data(), tickets.data(), &b->n_teams);
    BHIP(hipMalloc((void **)&b->d_tickets, sizeof(Ticket) * b->n_tickets));
    BHIP(hipMemcpy(b->d_tickets, tickets.data(), sizeof(Ticket) * b->n_tickets, hipMemcpyHostToDevice));
    BHIP(hipMalloc((void **)&b->d_team, sizeof(TeamArea) * n));
    BHIP(hipMalloc((void **)&b->d_team_err, sizeof(int)));
    for (int ph = 0; ph < 2; ph++)
        for (int e = 0; e < 2; e++) BHIP(hipEventCreate(&b->ev[ph][e]));
    b->out.resize(n);
    return MPDPFT_OK;
}

template <typename T> static int run(mpdpft_batch *b, int polish, int phase) {
    DevBuffers d{b->d_inst, b->d_out, b->d_tickets, b->d_team, b->d_next, b->d_team_err, b->d_slots,
                 b->d_fields, b->d_start, b->n, b->n_tickets, b->nslots, b->scope, b->Gmax};
    BHIP(launch<T>(d, polish, b->st, b->ev[phase][0], b->ev[phase][1]));
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
    b->kernel_ms[0] = b->kernel_ms[1] = b->kernel_bytes[0] = b->kernel_bytes[1] = 0.;
    b->stage_bytes[0] = b->stage_bytes[1] = 0.;
    auto count = [&](int ph) -> int {  // iterations, bytes and kernel time of the phase just run
        int terr = 0;
        BHIP(hipMemcpy(&terr, b->d_team_err, sizeof(int), hipMemcpyDeviceToHost));
        if (terr) return bfail(b, MPDPFT_EHIP, "a team barrier of the batch kernel did not complete");
        BHIP(hipMemcpy(b->out.data(), b->d_out, sizeof(BOut) * b->n, hipMemcpyDeviceToHost));
        for (int i = 0; i < b->n; i++) {
            b->phase_iterations[ph] += b->out[i].SCcount;
            b->phase_evaluations[ph] += b->out[i].bisection;
            b->kernel_bytes[ph] += b->out[i].bytes;
            b->stage_bytes[ph] += b->out[i].sbytes;
        }
        float ms = 0.f;
        BHIP(hipEventElapsedTime(&ms, b->ev[ph][0], b->ev[ph][1]));
        b->kernel_ms[ph] = ms;
        return MPDPFT_OK;
    };
    if (fp32) {
        rc = run<float>(b, 0, 0);
        if (rc) return rc;
        // FP64 polish from the FP32 result: its potential and mu start a full FP64 loop
        if (!b->d_start) BHIP(hipMalloc((void **)&b->d_start, sizeof(double) * (size_t)b->arena));
        BHIP(hipMemcpyAsync(b->d_start, b->d_fields, sizeof(double) * (size_t)b->arena, hipMemcpyDeviceToDevice, b->st));
        BHIP(hipStreamSynchronize(b->st));
        b->phase_seconds[0] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = count(0))) return rc;
        auto t1 = std::chrono::steady_clock::now();
        rc = run<double>(b, 1, 1);
        if (rc) return rc;
        BHIP(hipStreamSynchronize(b->st));
        b->phase_seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t1).count();
        b->seconds = b->phase_seconds[0] + b->phase_seconds[1];
        if ((rc = count(1))) return rc;
    } else {
        rc = run<double>(b, 0, 1);
        if (rc) return rc;
        BHIP(hipStreamSynchronize(b->st));
        b->seconds = b->phase_seconds[1] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if ((rc = count(1))) return rc;
    }
    fill_results(b->out.data(), b->inst.data(), b->params.data(), b->n, b->seconds, res);
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

extern "C" int mpdpft_batch_kernel_stats(const mpdpft_batch *b, double *kernel_ms, double *stage_bytes,
                                         double *kernel_bytes, double *instance_us) {
    if (!b) return MPDPFT_EINVAL;
    for (int ph = 0; ph < 2; ph++) {
        if (kernel_ms) kernel_ms[ph] = b->kernel_ms[ph];
        if (stage_bytes) stage_bytes[ph] = b->stage_bytes[ph];
        if (kernel_bytes) kernel_bytes[ph] = b->kernel_bytes[ph];
    }
    if (instance_us)  // the last phase's per-instance time between being taken from the queue and retiring
        for (int i = 0; i < b->n; i++) {
            instance_us[i] = (double)(b->out[i].t1 - b->out[i].t0) * 0.01;
            if (std::getenv("MPDPFT_BATCH_PHASES"))  // diagnostics: sweep A, mu search, sweep B, Pulay solve
                std::fprintf(stderr, "batch phases %d %.1f %.1f %.1f %.1f us\n", i, b->out[i].tph[0] * 0.01,
                             b->out[i].tph[1] * 0.01, b->out[i].tph[2] * 0.01, b->out[i].tph[3] * 0.01);
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
    (void)hipFree(b->d_tickets);
    (void)hipFree(b->d_team);
    (void)hipFree(b->d_team_err);
    for (int ph = 0; ph < 2; ph++)
        for (int e = 0; e < 2; e++)
            if (b->ev[ph][e]) (void)hipEventDestroy(b->ev[ph][e]);
    (void)hipFree(b->d_slots);
    (void)hipFree(b->d_fields);
    (void)hipFree(b->d_start);
    if (b->st) (void)hipStreamDestroy(b->st);
    delete b;
}

// ---- allocation-free solve ------------------------------------------------
// The whole batch solve with no device allocation: instance descriptors,
// results, the queue, the team areas, the FP32 phase's start copy and every
// running workgroup's field slot are carved from `scratch`, one static array
// per GPU that the library holds until the process ends.  Host code reaches
// the array only through the address hipGetSymbolAddress returns on each
// device (kept in na_dev); the array's own name is a host-side stand-in.
constexpr size_t kScratchBytes = MPDPFT_NOALLOC_SCRATCH_BYTES;
constexpr int kNoAllocMaxInst = MPDPFT_NOALLOC_MAX_INSTANCES;
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct NoAllocDevice {
    unsigned char *ptr;  // this device's copy of scratch
    int slots;           // resident workgroups of k_batch_solve
};
static NoAllocDevice na_dev[kNoAllocMaxDevices];
// host staging of the calls (static: the entry point allocates nothing), and the
// lock that keeps calls from several threads of a process one at a time
static std::mutex na_lock;
static BInst na_inst[kNoAllocMaxInst];
static BOut na_out[kNoAllocMaxInst];
static Ticket na_tickets[kTeamMax * kNoAllocMaxInst];
static int na_order[kNoAllocMaxInst];

extern "C" int mpdpft_batch_solve_noalloc(const mpdpft_params *p, int n, int device, int flags, double *fields,
                                          size_t fields_len, mpdpft_result *res, char *err, size_t err_len) {
    char msg[256];
    auto fail = [&](int rc, const char *m) {
        if (err && err_len) std::snprintf(err, err_len, "%s", m);
        return rc;
    };
#define NAHIP(x)                                                                                   \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) {                                                                    \
            std::snprintf(msg, sizeof msg, "%s: %s", #x, hipGetErrorString(e_));                   \
            return fail(MPDPFT_EHIP, msg);                                                         \
        }                                                                                          \
    } while (0)
    if (!p || !fields || !res || n < 1) return fail(MPDPFT_EINVAL, "null instances, fields or results, or n < 1");
    if (n > kNoAllocMaxInst) {
        std::snprintf(msg, sizeof msg, "%d instances: the allocation-free solve takes at most %d", n, kNoAllocMaxInst);
        return fail(MPDPFT_EINVAL, msg);
    }
    if (flags & ~MPDPFT_BATCH_FP32) return fail(MPDPFT_EINVAL, "unknown flags");
    std::lock_guard<std::mutex> guard(na_lock);
    long long off = 0, Gmax = 0;
    int scope = 0;
    for (int i = 0; i < n; i++) {
        int rc = prepare(msg, p[i], na_inst[i], i);
        if (rc) return fail(rc, msg);
        na_inst[i].out_off = off;
        off += 2LL * na_inst[i].S * na_inst[i].G;
        Gmax = std::max(Gmax, na_inst[i].G);
        scope = std::max(scope, na_inst[i].scope);
    }
    if ((unsigned long long)off > (unsigned long long)fields_len) {
        std::snprintf(msg, sizeof msg, "fields holds %zu doubles; the instances' densities and potentials need %lld",
                      fields_len, off);
        return fail(MPDPFT_EINVAL, msg);
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MPDPFT_ENODEV, "no GPU visible");
    if (device < 0) device = 0;
    if (device >= ndev || device >= kNoAllocMaxDevices) return fail(MPDPFT_EINVAL, "device index out of range");
    NAHIP(hipSetDevice(device));
    NoAllocDevice &dv = na_dev[device];
    if (!dv.ptr) {  // first call on this device: its copy of scratch and the slot count
        void *ptr = nullptr;
        NAHIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        NAHIP(resident_slots(device, &dv.slots));
        dv.ptr = (unsigned char *)ptr;
    }
    // fields must be device memory of this GPU with room for off doubles
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, fields) != hipSuccess ||
        (attr.type != hipMemoryTypeDevice && attr.type != hipMemoryTypeManaged) || attr.device != device) {
        (void)hipGetLastError();
        return fail(MPDPFT_EINVAL, "fields is not device memory of the given GPU");
    }
    void *base = nullptr;
    size_t extent = 0;
    NAHIP(hipMemGetAddressRange(&base, &extent, fields));
    if ((unsigned char *)fields + sizeof(double) * (size_t)off > (unsigned char *)base + extent)
        return fail(MPDPFT_EINVAL, "fields ends inside the instances' densities and potentials");
    // the array's layout for this call: fixed parts, then as many field slots as fit
    const bool fp32 = (flags & MPDPFT_BATCH_FP32) != 0;
    size_t used = 0;
    auto take = [&](size_t bytes) {
        const size_t at = used;
        used += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t at_inst = take(sizeof(BInst) * n), at_out = take(sizeof(BOut) * n);
    const size_t at_tick = take(sizeof(Ticket) * kTeamMax * n), at_team = take(sizeof(TeamArea) * n);
    const size_t at_next = take(sizeof(int)), at_terr = take(sizeof(int));
    const size_t at_start = fp32 ? take(sizeof(double) * (size_t)off) : 0;
    const size_t slot_bytes = (sizeof(double) * (size_t)slot_values(scope, Gmax) + 255) & ~(size_t)255;
    const size_t fit = used < kScratchBytes ? (kScratchBytes - used) / slot_bytes : 0;
    if (fit < 1) {
        std::snprintf(msg, sizeof msg,
                      "the instances need %zu bytes of scratch for their descriptors%s and %zu per workgroup slot; "
                      "the allocation-free solve has %zu",
                      used, fp32 ? " and start fields" : "", slot_bytes, kScratchBytes);
        return fail(MPDPFT_EINVAL, msg);
    }
    const int nslots = (int)std::max<long long>(1, std::min<long long>({(long long)n * kTeamMax, dv.slots, (long long)fit}));
    int n_teams = 0;
    const int n_tickets = make_tickets(na_inst, n, nslots, na_order, na_tickets, &n_teams);
    unsigned char *sp = dv.ptr;
    NAHIP(hipMemcpy(sp + at_inst, na_inst, sizeof(BInst) * n, hipMemcpyHostToDevice));
    NAHIP(hipMemcpy(sp + at_tick, na_tickets, sizeof(Ticket) * n_tickets, hipMemcpyHostToDevice));
    DevBuffers d{(BInst *)(sp + at_inst), (BOut *)(sp + at_out), (Ticket *)(sp + at_tick), (TeamArea *)(sp + at_team),
                 (int *)(sp + at_next), (int *)(sp + at_terr), sp + used, fields,
                 fp32 ? (double *)(sp + at_start) : nullptr, n, n_tickets, nslots, scope, Gmax};
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t st = nullptr;  // the null stream: no stream is created
    int terr = 0;
    if (fp32) {
        NAHIP(launch<float>(d, 0, st, nullptr, nullptr));
        NAHIP(hipMemcpyAsync(d.start, fields, sizeof(double) * (size_t)off, hipMemcpyDeviceToDevice, st));
        NAHIP(hipStreamSynchronize(st));
        NAHIP(hipMemcpy(&terr, d.team_err, sizeof(int), hipMemcpyDeviceToHost));
        if (terr) return fail(MPDPFT_EHIP, "a team barrier of the batch kernel did not complete");
        NAHIP(launch<double>(d, 1, st, nullptr, nullptr));
    } else {
        NAHIP(launch<double>(d, 0, st, nullptr, nullptr));
    }
    NAHIP(hipStreamSynchronize(st));
    const double seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    NAHIP(hipMemcpy(&terr, d.team_err, sizeof(int), hipMemcpyDeviceToHost));
    if (terr) return fail(MPDPFT_EHIP, "a team barrier of the batch kernel did not complete");
    NAHIP(hipMemcpy(na_out, d.out, sizeof(BOut) * n, hipMemcpyDeviceToHost));
    fill_results(na_out, na_inst, p, n, seconds, res);
    if (err && err_len) err[0] = 0;
    return MPDPFT_OK;
#undef NAHIP
}
