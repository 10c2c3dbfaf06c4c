// kernels.hip -- gfx950 kernels of the SC loop: fused FFT load/store
// functors, pointwise stages and deterministic two-level reductions.
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "kernels.h"

namespace mk {

using fftk::launch_pass;

#define PI_D 3.141592653589793

// ---------------------------------------------------------------------------
// small device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double d_exp(double x) {  // the reference's EXP (mpDPFT.h:266)
    if (x < -700.) return 0.;
    if (x < 700.) return exp(x);
    return 1.0142320547350045e+304;
}
__device__ __forceinline__ double d_pos(double x) { return x > 0. ? x : 0.; }

// trapezoid weight of grid point idx (1/2 at index 0 and s per axis)
__device__ __forceinline__ double trap_weight(const Grid &g, int idx) {
    double w = 1.;
    int rem = idx;
    for (int d = 0; d < g.D; d++) {
        int c = rem % g.EL;
        rem /= g.EL;
        if (c == 0 || c == g.s) w *= 0.5;
    }
    return w;
}

// physical coordinates of grid point idx from the lattice table
__device__ __forceinline__ void coords(const Grid &g, int idx, double *r) {
    const double *lat = g.lat;
    if (g.D == 2) {
        int i = idx / g.EL, j = idx - i * g.EL;
        r[0] = lat[i];
        r[1] = lat[j];
        r[2] = 0.;
    } else {
        int i = idx / (g.EL * g.EL), rem = idx - i * g.EL * g.EL, j = rem / g.EL, k = rem - j * g.EL;
        r[0] = lat[i];
        r[1] = lat[j];
        r[2] = lat[k];
    }
}

// block sum of one double over 256 threads (4 wavefronts of 64)
__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    return r;
}
__device__ __forceinline__ double block_min(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_down(v, off, 64));
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0) r = fmin(fmin(sh[0], sh[1]), fmin(sh[2], sh[3]));
    return r;
}

__device__ double pw92_eps(double n, double *vc) {
    const double A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
    if (!(n > 0.)) {
        *vc = 0.;
        return 0.;
    }
    double rs = cbrt(3. / (4. * PI_D * n));
    double srs = sqrt(rs);
    double Q0 = -2. * A * (1. + a1 * rs);
    double Q1 = 2. * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
    double Q1p = A * (b1 / srs + 2. * b2 + 3. * b3 * srs + 4. * b4 * rs);
    double lg = log(1. + 1. / Q1);
    double ec = Q0 * lg;
    double dec = -2. * A * a1 * lg - Q0 * Q1p / (Q1 * Q1 + Q1);
    *vc = ec - rs / 3. * dec;
    return ec;
}

// d v_c / d n of PW92: v_c = eps - (rs/3) eps', drs/dn = -rs/(3n), so
// dv_c/dn = -(rs/(3n)) ((2/3) eps' - (rs/3) eps'').
__device__ double pw92_dvdn(double n) {
    const double A = 0.031091, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382, b4 = 0.49294;
    if (!(n > 0.)) return 0.;
    double rs = cbrt(3. / (4. * PI_D * n));
    double srs = sqrt(rs);
    double Q0 = -2. * A * (1. + a1 * rs), Q0p = -2. * A * a1;
    double Q1 = 2. * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
    double Q1p = A * (b1 / srs + 2. * b2 + 3. * b3 * srs + 4. * b4 * rs);
    double Q1pp = A * (-0.5 * b1 / (rs * srs) + 1.5 * b3 / srs + 4. * b4);
    double den = Q1 * Q1 + Q1;
    double lg = log(1. + 1. / Q1);
    double lgp = -Q1p / den;
    double lgpp = -Q1pp / den + Q1p * Q1p * (2. * Q1 + 1.) / (den * den);
    double dec = Q0p * lg + Q0 * lgp;
    double d2ec = 2. * Q0p * lgp + Q0 * lgpp;
    return -(rs / (3. * n)) * ((2. / 3.) * dec - (rs / 3.) * d2ec);
}

// LDA xc (id -1 exchange, -12 PW92) per particle energy and potential,
// with the Angstrom/eV conversion of Units 2.
__device__ void xc_lda(int id, int units, double n, double *e, double *v) {
    const double Hartree = 27.211386245988, Bohr = 0.529177210903;
    double cn = units == 2 ? Bohr * Bohr * Bohr : 1., ce = units == 2 ? Hartree : 1.;
    double rho = n > 0. ? cn * n : 0.;
    double ee = 0., vv = 0.;
    if (rho > 0.) {
        if (id == -1) {
            double c = cbrt(3. / PI_D);
            ee = -0.75 * c * cbrt(rho);
            vv = -c * cbrt(rho);
        } else {
            ee = pw92_eps(rho, &vv);
        }
    }
    *e = ce * ee;
    *v = ce * vv;
}

// FP32 forms of the pointwise potentials (the precision variant)
__device__ float pw92_v_f(float n) {
    const float A = 0.031091f, a1 = 0.21370f, b1 = 7.5957f, b2 = 3.5876f, b3 = 1.6382f, b4 = 0.49294f;
    // below 1e-30 (r_s > 6e9) v_c ~ -a1 / (b4 r_s) < 1e-10, and r_s would overflow the FP32 forms
    if (!(n > 1e-30f)) return 0.f;
    float rs = cbrtf(3.f / (4.f * (float)PI_D * n));
    float srs = sqrtf(rs);
    float Q0 = -2.f * A * (1.f + a1 * rs);
    float Q1 = 2.f * A * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
    float Q1p = A * (b1 / srs + 2.f * b2 + 3.f * b3 * srs + 4.f * b4 * rs);
    float lg = logf(1.f + 1.f / Q1);
    float ec = Q0 * lg;
    float dec = -2.f * A * a1 * lg - Q0 * Q1p / (Q1 * Q1 + Q1);
    return ec - rs / 3.f * dec;
}
__device__ float xc_lda_v_f(int id, int units, float n) {
    const float Hartree = 27.211386245988f, Bohr = 0.529177210903f;
    float cn = units == 2 ? Bohr * Bohr * Bohr : 1.f, ce = units == 2 ? Hartree : 1.f;
    float rho = n > 0.f ? cn * n : 0.f;
    if (!(rho > 0.f)) return 0.f;
    if (id == -1) return -ce * cbrtf(3.f / (float)PI_D) * cbrtf(rho);
    return ce * pw92_v_f(rho);
}

// d v / d n of the LDA potentials above (Units 2 conversion included)
__device__ double xc_lda_dvdn(int id, int units, double n) {
    const double Hartree = 27.211386245988, Bohr = 0.529177210903;
    double cn = units == 2 ? Bohr * Bohr * Bohr : 1., ce = units == 2 ? Hartree : 1.;
    double rho = n > 0. ? cn * n : 0.;
    if (!(rho > 0.)) return 0.;
    if (id == -1) {
        double c = cbrt(3. / PI_D);
        return ce * cn * (-c / 3.) * cbrt(rho) / rho;
    }
    return ce * cn * pw92_dvdn(rho);
}

// ---------------------------------------------------------------------------
// FFT functors
// ---------------------------------------------------------------------------
struct LdGrid {  // real grid fields -> complex (im may be null)
    const double *re, *im;
    int EL;
    __device__ double2 operator()(int x, int y, int z) const {
        int i = (x * EL + y) * EL + z;
        double2 v;
        v.x = re[i];
        v.y = im ? im[i] : 0.;
        return v;
    }
};
// DelField's input with the TaperOff shift and taper fused: (V - max) T with
// the per-point factor table T of k_taper_factor (1 inside the taper radius)
struct LdGridTaper {
    const double *re, *T;
    const double *shift;  // device slot holding -max over the shell
    int EL;
    __device__ double2 operator()(int x, int y, int z) const {
        int i = (x * EL + y) * EL + z;
        double sft = -(*shift);
        if (sft < -1e307 || sft > 1e307) sft = 0.;
        return double2{(re[i] - sft) * T[i], 0.};
    }
};
// |grad V|^2 = Dx^2 + Dy^2 (+ Dz^2) formed while loading (sum_squares3 fused)
struct LdGridSumSq {
    const double *a, *b, *c;
    int EL;
    __device__ double2 operator()(int x, int y, int z) const {
        int i = (x * EL + y) * EL + z;
        double v = a[i] * a[i];
        if (b) v += b[i] * b[i];
        if (c) v += c[i] * c[i];
        return double2{v, 0.};
    }
};
struct LdBuf {
    const double2 *p;
    int n1, n2;
    __device__ double2 operator()(int x, int y, int z) const { return p[((size_t)x * n1 + y) * n2 + z]; }
};
struct StBuf {
    double2 *p;
    int n1, n2;
    __device__ void operator()(int x, int y, int z, double2 v) const { p[((size_t)x * n1 + y) * n2 + z] = v; }
};
struct StBufF0 {  // also captures dx^D * DC term
    double2 *p;
    int n1, n2;
    double *f0slot;
    double dxD;
    __device__ void operator()(int x, int y, int z, double2 v) const {
        p[((size_t)x * n1 + y) * n2 + z] = v;
        if (f0slot && x == 0 && y == 0 && z == 0) *f0slot = dxD * v.x;
    }
};

// multiplier applied on the first backward load (D a compile-time constant, so
// the per-axis lookups of a line's fixed coordinates are shared by its loads)
// KIND >= 0: the multiplier kind fixed at compile time; those instantiations are
// launched only with the (k, smoothing) table kek, so they read nothing else
// (with the table choice left to run time the compiler loads both forms).
template <int D, int KIND = -1> struct LdMulD {
    const double2 *p;
    int n1, n2, Dr, s;
    MulParams m;
    __device__ __forceinline__ double smooth(const int *q) const {
        if (!m.ek) return 1.;
        double f = m.ek[q[0]];
        for (int d = 1; d < D; d++) f *= m.ek[q[d]];
        return f;
    }
    __device__ double2 operator()(int x, int y, int z) const {
        double2 v = p[((size_t)x * n1 + y) * n2 + z];
        int q3[3] = {x, y, z};
        int q[3];
        for (int d = 0; d < D; d++) q[d] = q3[3 - D + d];
        double k[3] = {0., 0., 0.}, k2 = 0., fk = 1.;
        if (KIND >= 0 || m.kek) {  // k and the smoothing factor from one table entry per axis
            for (int d = 0; d < D; d++) {
                double2 t = m.kek[q[d]];
                k[d] = t.x;
                k2 += k[d] * k[d];
                fk = d == 0 ? t.y : fk * t.y;
            }
        } else {
            for (int d = 0; d < D; d++) {
                k[d] = m.kq[q[d]];
                k2 += k[d] * k[d];
            }
        }
        double2 o;
        switch (KIND >= 0 ? KIND : m.kind) {
        case MUL_SMOOTH: {
            double f = (KIND >= 0 || m.kek) ? fk : smooth(q);
            o.x = v.x * f;
            o.y = v.y * f;
            return o;
        }
        case MUL_DERIV_PAIR:
        case MUL_DERIV_LAP: {
            double f = (KIND >= 0 || m.kek) ? fk : smooth(q);
            // A = i ka f v  (Nyquist plane of axis a dropped)
            double2 A = {0., 0.}, B = {0., 0.};
            if (m.axis_a >= 0 && q[m.axis_a] != s / 2) {
                double c = k[m.axis_a] * f;
                A.x = -c * v.y;
                A.y = c * v.x;
            }
            if ((KIND >= 0 ? KIND : m.kind) == MUL_DERIV_PAIR) {
                if (q[m.axis_b] != s / 2) {
                    double c = k[m.axis_b] * f;
                    B.x = -c * v.y;
                    B.y = c * v.x;
                }
            } else {
                double c = -k2 * f * f;
                B.x = c * v.x;
                B.y = c * v.y;
            }
            // A + i B
            o.x = A.x - B.y;
            o.y = A.y + B.x;
            return o;
        }
        case MUL_SECOND: {
            double c = -k[m.axis_a] * k[m.axis_a];
            o.x = c * v.x;
            o.y = c * v.y;
            return o;
        }
        case MUL_HARTREE_PER: {
            double kk2 = 0.;
            int par = 0;
            for (int d = 0; d < D; d++) {
                double kh = m.kq_h[q[d]];
                kk2 += kh * kh;
                par += q[d];
            }
            double f0 = *m.f0slot;
            double re = m.dxD * v.x, im = m.dxD * v.y;
            if (kk2 > m.dkzero) {
                double sg = (par & 1) ? -1. : 1.;
                if (m.regularize) o.x = (re - sg * f0 * d_exp(-kk2)) / kk2;
                else o.x = re / kk2;
                o.y = im / kk2;
            } else {
                o.x = m.regularize ? f0 : 0.;
                o.y = 0.;
            }
            return o;
        }
        case MUL_HARTREE_2D: {  // HartreeRepulsion 2D (mpDPFT.cpp:2627-2645); dkzero is the |k| threshold
            double kk2 = 0.;
            int par = 0;
            for (int d = 0; d < D; d++) {
                double kh = m.kq_h[q[d]];
                kk2 += kh * kh;
                par += q[d];
            }
            const double k = sqrt(kk2), f0 = *m.f0slot;
            if (k > m.dkzero) {
                double sg = (par & 1) ? -1. : 1.;
                o.x = (m.dxD * v.x - sg * f0 * d_exp(-k * k)) / k;
                o.y = m.dxD * v.y / k;
            } else {
                o.x = 0.;
                o.y = 0.;
            }
            return o;
        }
        case MUL_DIPOLAR: {  // DipolarInteractionPos (mpDPFT.cpp:2735-2746), 3D; dkzero = 0.01 dk^2
            double kk2 = 0.;
            int par = 0;
            for (int d = 0; d < D; d++) {
                double kh = m.kq_h[q[d]];
                kk2 += kh * kh;
                par += q[d];
            }
            const double kz = m.kq_h[q[D - 1]], ql2 = kz * kz;
            if (kk2 > m.dkzero) {
                const double f = (ql2 / kk2) - 1. / 3. + (8. * PI_D) / 3.;
                const double sg = (par & 1) ? -1. : 1.;
                o.x = (m.dxD * v.x - sg * m.n0 * d_exp(-kk2)) * f;
                o.y = m.dxD * v.y * f;
            } else {
                o.x = 0.;
                o.y = 0.;
            }
            return o;
        }
        default:
            return v;
        }
    }
};

// e^-x I0(x), x >= 0 (the 2D Hartree compensation; the reference calls
// gsl_sf_bessel_I0): the power series sum_k ((x/2)^k / k!)^2 of positive terms
// for x <= 30, the asymptotic series (2 pi x)^-1/2 sum_k a_k x^-k with a_k =
// a_{k-1} (2k-1)^2 / (8k) beyond, each summed until the next term is below 1e-17
// of the sum (within 2e-15 of the exact value).
__device__ double exp_i0(double x) {
    if (x <= 30.) {
        double q = 0.25 * x * x, t = 1., sum = 1.;
        for (int k = 1; k < 400; k++) {
            t *= q / ((double)k * (double)k);
            sum += t;
            if (t < 1e-17 * sum) break;
        }
        return exp(-x) * sum;
    }
    double a = 1., sum = 1., xi = 1. / x;
    for (int k = 1; k < 200; k++) {
        double nx = a * (2. * k - 1.) * (2. * k - 1.) / (8. * k) * xi;
        if (nx < 1e-17 * sum || nx > a) break;
        a = nx;
        sum += a;
    }
    return sum / sqrt(2. * PI_D * x);
}

// store the real/imag parts onto the (s+1)^D grid with the periodic images
struct StGrid {
    double *re, *im;
    double scale;
    int post;  // 0 plain, 1 POS(re), 2 Hartree compensation on re (3D erf form), 3 (2D Bessel form)
    int EL, s, D;
    const double *lat;
    const double *f0slot;
    double c1;  // 4 pi gammaH for post 2, 2 pi gammaH for post 3
    // post 3: 2 pi gammaH (val + f0 e^{-z^2/8} I0(z^2/8) / (4 sqrt(pi))) (mpDPFT.cpp:2647-2655)
    __device__ double hart2d(int idx, double val) const {
        int a = idx / EL, b = idx - a * EL;
        double z = sqrt(lat[a] * lat[a] + lat[b] * lat[b]), e = z * z / 8.;
        double ei0 = e < 1.0e+300 ? exp_i0(e) : 1. / sqrt(2. * PI_D * e);
        return c1 * (val + (*f0slot) * ei0 / (4. * sqrt(PI_D)));
    }
    __device__ double hart(int idx, double val) const {
        double r[3];
        int i = idx;
        if (D == 2) {
            int a = i / EL, b = i - a * EL;
            r[0] = lat[a];
            r[1] = lat[b];
            r[2] = 0.;
        } else {
            int a = i / (EL * EL), rem = i - a * EL * EL, b = rem / EL, c = rem - b * EL;
            r[0] = lat[a];
            r[1] = lat[b];
            r[2] = lat[c];
        }
        double z = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        double zf = z > 1.0e-10 ? PI_D * erf(z / 2.) / z : sqrt(PI_D);
        return c1 * (val + (*f0slot) * zf / (4. * PI_D * PI_D));
    }
    __device__ void put(int idx, double2 v) const {
        if (re) {
            double a = scale * v.x;
            if (post == 1) a = d_pos(a);
            else if (post == 2) a = hart(idx, a);
            else if (post == 3) a = hart2d(idx, a);
            re[idx] = a;
        }
        if (im) im[idx] = scale * v.y;
    }
    __device__ void operator()(int x, int y, int z, double2 v) const {
        // the point and its periodic images (index 0 repeated at s along each axis)
        const int i0 = (x * EL + y) * EL + z, sy = s * EL;
        put(i0, v);
        if (z == 0) put(i0 + s, v);
        if (y == 0) {
            put(i0 + sy, v);
            if (z == 0) put(i0 + sy + s, v);
        }
        if (D == 3 && x == 0) {
            const int ix = i0 + sy * EL;
            put(ix, v);
            if (z == 0) put(ix + s, v);
            if (y == 0) {
                put(ix + sy, v);
                if (z == 0) put(ix + sy + s, v);
            }
        }
    }
};

// zero-padded loads of the isolated Hartree transform
struct LdGridPadZ {  // z < lim of the grid field, zeros beyond
    const double *n;
    int EL, lim;
    __device__ double2 operator()(int x, int y, int z) const {
        double2 v = {0., 0.};
        if (z < lim) v.x = n[(x * EL + y) * EL + z];
        return v;
    }
};
struct LdBufPad {  // virtual zeros beyond index s along 'axis'
    const double2 *p;
    int n1, n2, axis, s;
    __device__ double2 operator()(int x, int y, int z) const {
        int c = axis == 0 ? x : (axis == 1 ? y : z);
        if (c >= s) return double2{0., 0.};
        return p[((size_t)x * n1 + y) * n2 + z];
    }
};
struct StBufMulK {  // store v * khat
    double2 *p;
    const double *khat;
    int n1, n2;
    __device__ void operator()(int x, int y, int z, double2 v) const {
        size_t i = ((size_t)x * n1 + y) * n2 + z;
        double k = khat[i];
        p[i] = double2{v.x * k, v.y * k};
    }
};
struct StGridIso {  // keep [0,s]^D of the circular convolution
    double *out;
    double scale;
    int EL, s, D, accumulate;
    int images0;  // write 0 at the periodic images (a coordinate equal to s)
    __device__ void operator()(int x, int y, int z, double2 v) const {
        if (z > s) return;
        int idx = (x * EL + y) * EL + z;
        double a = scale * v.x;
        if (images0 && (z == s || y == s || (D == 3 && x == s))) a = 0.;
        if (accumulate) out[idx] += a;
        else out[idx] = a;
    }
};
struct LdKernel {  // 1/r on the (2s)^D torus, g0 at the origin
    int M, s;
    double dx, g0;
    int D;
    __device__ double2 operator()(int x, int y, int z) const {
        int c[3] = {x, y, z};
        double r2 = 0.;
        for (int d = 3 - D; d < 3; d++) {
            int m = c[d] <= s ? c[d] : c[d] - M;
            double t = (double)m * dx;
            r2 += t * t;
        }
        double2 v;
        v.x = r2 > 0. ? 1. / sqrt(r2) : g0;
        v.y = 0.;
        return v;
    }
};

// ---------------------------------------------------------------------------
// axis-pass dispatch over the supported lengths
// ---------------------------------------------------------------------------
template <int AXIS, class Ld, class St>
static hipError_t pass(int N, Ld ld, St st, int L0, int L1, int sign, const double2 *tw, hipStream_t stream) {
    switch (N) {
    case 16: return launch_pass<16, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 32: return launch_pass<32, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 64: return launch_pass<64, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 96: return launch_pass<96, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 128: return launch_pass<128, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 192: return launch_pass<192, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 256: return launch_pass<256, AXIS>(ld, st, L0, L1, sign, tw, stream);
    case 512: return launch_pass<512, AXIS>(ld, st, L0, L1, sign, tw, stream);
    default: return hipErrorInvalidValue;
    }
}

#define CHK(x)                               \
    do {                                     \
        hipError_t e_ = (x);                 \
        if (e_ != hipSuccess) return e_;     \
    } while (0)

hipError_t fft_forward_real(const Grid &g, const double *re, const double *im, double2 *spec,
                            const double2 *tw, double *f0slot, hipStream_t st) {
    int s = g.s;
    double dxD = g.D == 3 ? g.dx * g.dx * g.dx : g.dx * g.dx;
    LdGrid lg{re, im, g.EL};
    StBuf sb{spec, s, s};
    LdBuf lb{spec, s, s};
    StBufF0 sf{spec, s, s, f0slot, dxD};
    if (g.D == 3) {
        CHK(pass<2>(s, lg, sb, s, s, -1, tw, st));
        CHK(pass<1>(s, lb, sb, s, s, -1, tw, st));
        CHK(pass<0>(s, lb, sf, s, s, -1, tw, st));
    } else {
        CHK(pass<2>(s, lg, sb, 1, s, -1, tw, st));
        CHK(pass<1>(s, lb, sf, 1, s, -1, tw, st));
    }
    return hipSuccess;
}


hipError_t fft_backward_real(const Grid &g, const double2 *spec, double2 *work, const MulParams &m, double *re,
                             double *im, double scale, int post, double c1, const double2 *tw, hipStream_t st) {
    int s = g.s;
    StBuf sb{work, s, s};
    LdBuf lb{work, s, s};
    StGrid sg{re, im, scale, post, g.EL, s, g.D, g.lat, m.f0slot, c1};
    if (g.D == 3) {
        // DelField's multipliers (with the (k, smoothing) table) compiled per kind
        if (m.kind == MUL_DERIV_PAIR && m.kek) {
            LdMulD<3, MUL_DERIV_PAIR> lm{spec, s, s, 3, s, m};
            CHK(pass<0>(s, lm, sb, s, s, +1, tw, st));
        } else if (m.kind == MUL_DERIV_LAP && m.kek) {
            LdMulD<3, MUL_DERIV_LAP> lm{spec, s, s, 3, s, m};
            CHK(pass<0>(s, lm, sb, s, s, +1, tw, st));
        } else if (m.kind == MUL_SMOOTH && m.kek) {
            LdMulD<3, MUL_SMOOTH> lm{spec, s, s, 3, s, m};
            CHK(pass<0>(s, lm, sb, s, s, +1, tw, st));
        } else {
            LdMulD<3> lm{spec, s, s, 3, s, m};
            CHK(pass<0>(s, lm, sb, s, s, +1, tw, st));
        }
        CHK(pass<1>(s, lb, sb, s, s, +1, tw, st));
        CHK(pass<2>(s, lb, sg, s, s, +1, tw, st));
    } else {
        LdMulD<2> lm{spec, s, s, 2, s, m};
        CHK(pass<1>(s, lm, sb, 1, s, +1, tw, st));
        CHK(pass<2>(s, lb, sg, 1, s, +1, tw, st));
    }
    return hipSuccess;
}

// Time each block shape of a transform and keep the fastest (once per length
// and precision per process).  run() enqueues one transform on st.
template <class Run> hipError_t tune_one(int N, bool fp32, Run run, hipStream_t st) {
    static bool done[2][10] = {};
    int k = N == 16 ? 0 : N == 32 ? 1 : N == 64 ? 2 : N == 96 ? 3 : N == 128 ? 4 : N == 192 ? 5 : N == 256 ? 6
          : N == 512 ? 7 : 8;
    if (done[fp32 ? 1 : 0][k]) return hipSuccess;
    hipEvent_t e0, e1;
    CHK(hipEventCreate(&e0));
    CHK(hipEventCreate(&e1));
    float best = 1e30f;
    int bestSel = 1;
    for (int sel = 0; sel < 3; sel++) {
        fftk::bsel_slot(N, fp32) = sel;
        CHK(run());  // warm-up (first launch of this shape)
        CHK(hipEventRecord(e0, st));
        for (int r = 0; r < 3; r++) CHK(run());
        CHK(hipEventRecord(e1, st));
        CHK(hipEventSynchronize(e1));
        float ms = 0.f;
        CHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) {
            best = ms;
            bestSel = sel;
        }
    }
    fftk::bsel_slot(N, fp32) = bestSel;
    done[fp32 ? 1 : 0][k] = true;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return hipSuccess;
}

hipError_t fft_autotune(const Grid &g, const double *field, double2 *spec, const double2 *tw, double2 *pad,
                        const double *khat, const double2 *tw2, double *out, hipStream_t st) {
    CHK(tune_one(g.s, false, [&]() { return fft_forward_real(g, field, nullptr, spec, tw, nullptr, st); }, st));
    if (pad && g.D == 3)
        CHK(tune_one(2 * g.s, false, [&]() { return hartree_isolated(g, field, pad, khat, tw2, 1., out, 0, st); },
                     st));
    return hipSuccess;
}

hipError_t fft_forward_taper(const Grid &g, const double *V, const double *T, const double *shift_dev, double2 *spec,
                             const double2 *tw, hipStream_t st) {
    int s = g.s;
    LdGridTaper lg{V, T, shift_dev, g.EL};
    StBuf sb{spec, s, s};
    LdBuf lb{spec, s, s};
    if (g.D == 3) {
        CHK(pass<2>(s, lg, sb, s, s, -1, tw, st));
        CHK(pass<1>(s, lb, sb, s, s, -1, tw, st));
        CHK(pass<0>(s, lb, sb, s, s, -1, tw, st));
    } else {
        CHK(pass<2>(s, lg, sb, 1, s, -1, tw, st));
        CHK(pass<1>(s, lb, sb, 1, s, -1, tw, st));
    }
    return hipSuccess;
}

hipError_t fft_forward_sumsq(const Grid &g, const double *a, const double *b, const double *c, double2 *spec,
                             const double2 *tw, hipStream_t st) {
    int s = g.s;
    LdGridSumSq lg{a, b, c, g.EL};
    StBuf sb{spec, s, s};
    LdBuf lb{spec, s, s};
    if (g.D == 3) {
        CHK(pass<2>(s, lg, sb, s, s, -1, tw, st));
        CHK(pass<1>(s, lb, sb, s, s, -1, tw, st));
        CHK(pass<0>(s, lb, sb, s, s, -1, tw, st));
    } else {
        CHK(pass<2>(s, lg, sb, 1, s, -1, tw, st));
        CHK(pass<1>(s, lb, sb, 1, s, -1, tw, st));
    }
    return hipSuccess;
}

hipError_t build_isolated_kernel(const Grid &g, double2 *pad, const double2 *tw2, double g0, double *khat,
                                 hipStream_t st);

hipError_t hartree_isolated(const Grid &g, const double *n, double2 *pad, const double *khat,
                            const double2 *tw2, double scale, double *out, int accumulate, hipStream_t st,
                            int transpose) {
    // forward: sources [0,s)^D (the periodic images carry no charge), potential on [0,s]^D;
    // transpose: sources [0,s]^D, result on [0,s)^D and 0 at the images
    int s = g.s, M = 2 * s, lim = transpose ? s + 1 : s;
    StBuf sb{pad, M, M};
    LdBuf lb{pad, M, M};
    StBufMulK sk{pad, khat, M, M};
    StGridIso so{out, scale, g.EL, s, g.D, accumulate, transpose};
    if (g.D == 3) {
        // forward, pruned: only lines that can hold nonzeros
        CHK(pass<2>(M, LdGridPadZ{n, g.EL, lim}, sb, lim, lim, -1, tw2, st));
        CHK(pass<1>(M, LdBufPad{pad, M, M, 1, lim}, sb, lim, M, -1, tw2, st));
        CHK(pass<0>(M, LdBufPad{pad, M, M, 0, lim}, sk, M, M, -1, tw2, st));
        // backward, pruned: only lines that reach [0,s]^3
        CHK(pass<0>(M, lb, sb, M, M, +1, tw2, st));
        CHK(pass<1>(M, lb, sb, s + 1, M, +1, tw2, st));
        CHK(pass<2>(M, lb, so, s + 1, s + 1, +1, tw2, st));
    } else {
        CHK(pass<2>(M, LdGridPadZ{n, g.EL, lim}, sb, 1, lim, -1, tw2, st));
        CHK(pass<1>(M, LdBufPad{pad, M, M, 1, lim}, sk, 1, M, -1, tw2, st));
        CHK(pass<1>(M, lb, sb, 1, M, +1, tw2, st));
        CHK(pass<2>(M, lb, so, 1, s + 1, +1, tw2, st));
    }
    return hipSuccess;
}

__global__ void k_copy_re(size_t n, const double2 *a, double *b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i].x;
}

hipError_t build_isolated_kernel(const Grid &g, double2 *pad, const double2 *tw2, double g0, double *khat,
                                 hipStream_t st) {
    int s = g.s, M = 2 * s;
    StBuf sb{pad, M, M};
    LdBuf lb{pad, M, M};
    LdKernel lk{M, s, g.dx, g0, g.D};
    size_t tot = (size_t)M * M * (g.D == 3 ? M : 1);
    if (g.D == 3) {
        CHK(pass<2>(M, lk, sb, M, M, -1, tw2, st));
        CHK(pass<1>(M, lb, sb, M, M, -1, tw2, st));
        CHK(pass<0>(M, lb, sb, M, M, -1, tw2, st));
    } else {
        CHK(pass<2>(M, lk, sb, 1, M, -1, tw2, st));
        CHK(pass<1>(M, lb, sb, 1, M, -1, tw2, st));
    }
    hipLaunchKernelGGL(k_copy_re, dim3(1024), dim3(256), 0, st, tot, (const double2 *)pad, khat);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// reductions
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_finish_sums(const double *partials, int nres, double *out, double scale) {
    __shared__ double sh[4];
    for (int r = 0; r < nres; r++) {
        double v = 0.;
        for (int i = threadIdx.x; i < kRedBlocks; i += 256) v += partials[(size_t)r * kRedBlocks + i];
        double t = block_sum(v, sh);
        if (threadIdx.x == 0) out[r] = t * scale;
        __syncthreads();
    }
}
hipError_t finish_sums(const double *partials, int nres, double *out, double scale, hipStream_t st) {
    hipLaunchKernelGGL(k_finish_sums, dim3(1), dim3(256), 0, st, partials, nres, out, scale);
    return hipGetLastError();
}
__global__ __launch_bounds__(256) void k_finish_min(const double *partials, int nres, double *out) {
    __shared__ double sh[4];
    for (int r = 0; r < nres; r++) {
        double v = 1.0e308;
        for (int i = threadIdx.x; i < kRedBlocks; i += 256) v = fmin(v, partials[(size_t)r * kRedBlocks + i]);
        double t = block_min(v, sh);
        if (threadIdx.x == 0) out[r] = t;
        __syncthreads();
    }
}
hipError_t finish_minmax(const double *partials, int nres, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_finish_min, dim3(1), dim3(256), 0, st, partials, nres, out);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// densities
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_tf_density(Grid g, const double *__restrict__ V, double mu, double pref,
                                                    double *__restrict__ n, double *partials) {
    __shared__ double sh[4];
    double acc = 0.;
    int G = (int)g.G;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < G; i += gridDim.x * 256) {
        double t = mu - V[i];
        double val = 0.;
        if (t > 0.) val = g.D == 2 ? pref * t : pref * t * sqrt(t);
        if (n) n[i] = val;
        acc += trap_weight(g, i) * val;
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}
hipError_t tf_density(const Grid &g, const double *V, double mu, double pref, double *n, double *partials,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_tf_density, dim3(kRedBlocks), dim3(256), 0, st, g, V, mu, pref, n, partials);
    return hipGetLastError();
}

constexpr int kMaxMu = 32;
__global__ __launch_bounds__(256) void k_tf_multi(Grid g, const double *__restrict__ V, const double *mus, int nmu,
                                                  double pref, double *partials) {
    __shared__ double sh[4];
    double acc[kMaxMu];
    double mu[kMaxMu];
#pragma unroll
    for (int j = 0; j < kMaxMu; j++) {
        acc[j] = 0.;
        mu[j] = j < nmu ? mus[j] : -1.0e300;
    }
    int G = (int)g.G;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < G; i += gridDim.x * 256) {
        double v = V[i], w = trap_weight(g, i);
#pragma unroll
        for (int j = 0; j < kMaxMu; j++) {
            double t = mu[j] - v;
            double val = 0.;
            if (t > 0.) val = g.D == 2 ? pref * t : pref * t * sqrt(t);
            acc[j] += w * val;
        }
    }
    for (int j = 0; j < nmu; j++) {
        double r = block_sum(acc[j], sh);
        if (threadIdx.x == 0) partials[(size_t)j * kRedBlocks + blockIdx.x] = r;
        __syncthreads();
    }
}
hipError_t tf_density_multi(const Grid &g, const double *V, const double *mus_dev, int nmu, double pref,
                            double *partials, hipStream_t st) {
    if (nmu > kMaxMu) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_tf_multi, dim3(kRedBlocks), dim3(256), 0, st, g, V, mus_dev, nmu, pref, partials);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Gombas and VWN correlation (GombasCorrelation / VWNCorrelation,
// mpDPFT.cpp:2774-2828), densities in 1/Angstrom^3, energies in eV.
// Gombas: e(n) = -gammaH n (a1 t / (a2 + c t) + (b1 / c) ln(1 + b2 c t)),
// t = POS(n)^(1/3); v = de/dn.  VWN: e(n) = n A F(x) for n > 1e-15, x =
// (3 / (4 pi c^3 n))^(1/6) = sqrt(r_s), F = Y + (2b/Q) T - (b x0 / X0)(Z + 2(b +
// 2 x0) T / Q) with X = x^2 + b x + c', Y = ln(x^2/X), Z = ln((x - x0)^2 / X),
// T = atan(Q / (2x + b)); v = de/dn = A (F - (x/6) dF/dx) for n > 1e-12.  The
// potentials are used in 3D with Units 2 only (0 otherwise), as in the reference.
// ---------------------------------------------------------------------------
constexpr double kGa1 = 0.0357, kGa2 = 0.0562, kGb1 = 0.0311, kGb2 = 2.39, kBohrA = 0.52917721067;
__device__ __forceinline__ void gombas(double n, double gammaH, double *e, double *v) {
    double t = pow(d_pos(n), 1. / 3.), dn = kGa2 + kBohrA * t, la = 1. + kGb2 * kBohrA * t;
    double f = kGa1 * t / dn + (kGb1 / kBohrA) * log(la);
    *e = -gammaH * n * f;
    *v = -gammaH * (f + (t / 3.) * (kGa1 * kGa2 / (dn * dn) + kGb1 * kGb2 / la));
}
constexpr double kVA = 0.846021, kVb = 13.072, kVc = 42.7198, kVQ = 0.0448998, kVx0 = -0.409286, kVX0 = 37.53713;
__device__ __forceinline__ void vwn_F(double x, double *F, double *dF) {
    double X = x * x + kVb * x + kVc, dX = 2. * x + kVb;
    double Y = log(x * x / X), Z = log((x - kVx0) * (x - kVx0) / X), T = atan(kVQ / (2. * x + kVb));
    double dT = -2. * kVQ / ((2. * x + kVb) * (2. * x + kVb) + kVQ * kVQ);
    const double g = kVb * kVx0 / kVX0, h = 2. * (kVb + 2. * kVx0) / kVQ;
    *F = Y + 2. * kVb / kVQ * T - g * (Z + h * T);
    *dF = (2. / x - dX / X) + 2. * kVb / kVQ * dT - g * ((2. / (x - kVx0) - dX / X) + h * dT);
}
__device__ __forceinline__ double vwn_x(double n) {
    return pow(3. / (4. * PI_D * kBohrA * kBohrA * kBohrA * n), 1. / 6.);
}
// RenormalizedContact (mpDPFT.cpp:2891-2904), 2D, two species: with n' = POS(n) +
// 1e-16 and eta = -2 / ln(4 pi a^2 n'), the energy density pi n'_0 n'_1 (b(eta_0) +
// b(eta_1)) and the potential of species s, pi n'_o (b(eta_s) + b(eta_o) + b'(eta_s)
// eta_s^2 / 2) (o the other species).  b and b' are the reference's separately
// fitted functions (PiotrBeta, :4458-4472): fit coefficients kept as the reference
// has them, since the results are defined by the fit; both branches meet at eta =
// 1.55 to 1e-12 in value and in b'.
__device__ __forceinline__ double qmc_beta(int der, double x) {
    const double x2 = x * x, x4 = x2 * x2;
    if (x < 0.) return 100000.;
    if (der == 0) {
        if (x > 1.55) return -0.4633747986519757 / x + 1.443565409987811 - 0.20465388562309392 / x2;
        return -0.0004147527851895318 + 1.0062109233358525 * x - 0.0584541517808854 * x2 -
               0.3073866986724988 * x4 + 0.25331504986348463 * x4 * x - 0.06139783376450124 * x4 * x2;
    }
    if (x > 1.55)
        return -0.0011588876670877157 + 0.0663320388441727 / x + 1.745208782122636 / (x2 * x) -
               1.25878015302857 / x4;
    return 1 - 0.028100195459653712 * x - 0.3963015574364809 * x2 - 0.3445575747110073 * x2 * x -
           0.0987982521417292 * x4 + 1.2591046941895094 * x4 * x - 1.2654959902790368 * x4 * x2 +
           0.5086340329349458 * x4 * x2 * x - 0.07525628170004087 * x4 * x4;
}
__device__ __forceinline__ double renorm_eta(double n, double a) {
    return -2. / log(4. * PI_D * a * a * (d_pos(n) + 1.0e-16));
}

// the pointwise potentials for k_update_v, out of line: inlined, their
// transcendentals set the register count (and occupancy) of every update
__device__ __noinline__ double xc_v(int t, int units, double n) {
    double e, v;
    xc_lda(t, units, n, &e, &v);
    return v;
}
__device__ __noinline__ double xc_v_f(int t, int units, double n) { return (double)xc_lda_v_f(t, units, (float)n); }
// DiracExchange (mpDPFT.cpp:2766-2769): -gammaH 2 (D n / pi)^{1/D} / (D - 1), D > 1;
// the D-th root as cbrt (3D) or sqrt (2D) rather than pow(., 1/D) (1/3 is not a
// double, so the two differ by about an ulp)
__device__ __forceinline__ double dirac_root(double x, int D) { return D == 3 ? cbrt(x) : sqrt(x); }
__device__ __noinline__ double dirac_v(double n, int D, double gammaH) {
    const double DD = (double)D;
    return D > 1 ? -gammaH * (2. * dirac_root(d_pos(n) * DD / PI_D, D)) / (DD - 1.) : 0.;
}
__device__ __noinline__ double dirac_v_f(double n, int D, double gammaH) {
    float nf = (float)d_pos(n), Df = (float)D, x = nf * Df / (float)PI_D;
    return D > 1 ? (double)(-(float)gammaH * (2.f * (D == 3 ? cbrtf(x) : sqrtf(x))) / (Df - 1.f)) : 0.;
}
// DeltaEkin2D3Dcrossover (mpDPFT.cpp:2906-2923), 2D, Units 1: l = floor(-1/2 +
// sqrt(1/4 + 4 pi n / alpha)) filled transverse levels; for l > 0 the energy
// density E0 + n V0 + n^2 V1 (V0 = l alpha / 2, V1 = -pi l / (l + 1), E0 =
// -alpha^2 l (l + 1)(l + 2) / (48 pi)) and the potential V0 + 2 n V1.
__device__ __noinline__ double crossover(double n, double a, int energy) {
    const double l = floor(-0.5 + sqrt(0.25 + 4. * PI_D * n / a));
    if (!(l > 1.0e-12)) return 0.;
    const double V0 = 0.5 * l * a, V1 = -PI_D * l / (l + 1.);
    if (energy) return -a * a * l * (l + 1.) * (l + 2.) / (48. * PI_D) + n * V0 + n * n * V1;
    return V0 + n * 2. * V1;
}
// Quasi2Dcontact (mpDPFT.cpp:2952-2967), 2D, two species, Units 1, beta != 0:
// levels l_s = max(0, floor(-1/2 + sqrt(1/4 + 4 pi n_s / alpha))), G = beta
// sqrt(alpha), f1 = alpha / (2 pi); potential of the focal species (n1)
// G (n2 eta1(l1,l2) + f1 eta2(l1,l2)), energy density G (n1 n2 eta1(l1,l2) +
// f1 (n1 eta2(l1,l2) + n2 eta2(l2,l1)) + f1^2 eta3(l1,l2)); eta = the reference's
// 31 x 31 table (NaN for levels beyond the rows given, where the reference reads
// past its table).
__device__ __noinline__ double quasi2d(double n1, double n2, double a, double beta, const double *eta, int rows,
                                       int energy) {
    int l1 = (int)floor(-0.5 + sqrt(0.25 + 4. * PI_D * n1 / a)), l2 = (int)floor(-0.5 + sqrt(0.25 + 4. * PI_D * n2 / a));
    l1 = l1 < 0 ? 0 : l1;
    l2 = l2 < 0 ? 0 : l2;
    const int k12 = l1 * 31 + l2, k21 = l2 * 31 + l1;
    if (l1 > 30 || l2 > 30 || k12 >= rows || k21 >= rows) return __builtin_nan("");
    const double f1 = a / (2. * PI_D), f2 = f1 * f1, G = beta * sqrt(a);
    if (energy)
        return G * (n1 * n2 * eta[3 * k12] + f1 * (n1 * eta[3 * k12 + 1] + n2 * eta[3 * k21 + 1]) + f2 * eta[3 * k12 + 2]);
    return G * (n2 * eta[3 * k12] + f1 * eta[3 * k12 + 1]);
}
__device__ __noinline__ double renorm_v(double n_s, double n_o, double a) {
    const double es = renorm_eta(n_s, a), eo = renorm_eta(n_o, a);
    return PI_D * (d_pos(n_o) + 1.0e-16) * (qmc_beta(0, es) + qmc_beta(0, eo) + 0.5 * qmc_beta(1, es) * es * es);
}
__device__ __noinline__ double gombas_v(double n, double gammaH) {
    double e, v;
    gombas(n, gammaH, &e, &v);
    return v;
}
__device__ __noinline__ double vwn_v(double n) {
    const double d = d_pos(n);
    if (!(d > 1.0e-12)) return 0.;
    double x = vwn_x(d), F, dF;
    vwn_F(x, &F, &dF);
    return kVA * (F - x / 6. * dF);
}

// ---------------------------------------------------------------------------
// potentials
// ---------------------------------------------------------------------------
// one species per blockIdx.y
__global__ __launch_bounds__(256) void k_update_v(Grid g, VintParams vp, const double *__restrict__ den,
                                                  const double *__restrict__ env, double *V) {
    const size_t G = g.G;
    const int s = blockIdx.y;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < G; i += (size_t)gridDim.x * 256) {
        const size_t idx = (size_t)s * G + i;
        double ns = den[idx];
        double acc = 0.;
        for (int a = 0; a < vp.n_inter; a++) {
            int t = vp.inter[a];
            double v = 0.;
            if (t == 1) {
                for (int sp = 0; sp < vp.S; sp++)
                    if (sp != s) v += den[(size_t)sp * G + i];
                v *= vp.beta;
            } else if (t == 8) {
                v = vp.hartree_in_V ? V[idx] : 0.;
            } else if (t == 10) {
                v = den[i] * (2. * vp.beta);
            } else if (t == 11) {
                v = vp.fp32 ? dirac_v_f(ns, g.D, vp.gammaH) : dirac_v(ns, g.D, vp.gammaH);
            } else if (t == -1 || t == -12) {
                v = vp.fp32 ? xc_v_f(t, vp.units, ns) : xc_v(t, vp.units, ns);
            } else if (t == 9) {  // two species (checked at create)
                v = renorm_v(ns, den[(size_t)(1 - s) * G + i], vp.beta);
            } else if (t == 12 && g.D == 3 && vp.units == 2) {
                v = gombas_v(ns, vp.gammaH);
            } else if (t == 13 && g.D == 3 && vp.units == 2) {
                v = vwn_v(ns);
            } else if (t == 14 && vp.units == 1) {
                v = crossover(ns, vp.alpha, 0);
            } else if (t == 24) {  // two species, Units 1, beta != 0 (checked at create)
                v = quasi2d(ns, den[(size_t)(1 - s) * G + i], vp.alpha, vp.beta, vp.eta, vp.eta_rows, 0);
            } else if (t == 15) {
                v = vp.dip_k ? vp.dip_k[idx] : 0.;
            } else if (t == 17) {
                v = vp.dip_r ? vp.dip_r[idx] : 0.;
            }
            acc += v;
        }
        if (vp.inc < 1.) acc *= vp.inc;
        V[idx] = acc + env[idx];
    }
}
hipError_t update_v(const Grid &g, const VintParams &vp, const double *den_all, const double *env, double *V,
                    hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<size_t>(2048 / (size_t)vp.S + 1, (g.G + 255) / 256);
    hipLaunchKernelGGL(k_update_v, dim3(blocks, vp.S), dim3(256), 0, st, g, vp, den_all, env, V);
    return hipGetLastError();
}

// d V_s / d n_s of the pointwise interactions of k_update_v (one species)
__global__ __launch_bounds__(256) void k_vint_dvdn(Grid g, VintParams vp, const double *__restrict__ den,
                                                   double *out) {
    const double DD = (double)g.D;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < g.G; i += (size_t)gridDim.x * 256) {
        double ns = den[i], acc = 0.;
        for (int a = 0; a < vp.n_inter; a++) {
            int t = vp.inter[a];
            if (t == 10) {
                acc += 2. * vp.beta;
            } else if (t == 11) {  // v = -gammaH 2 (D n / pi)^{1/D} / (D - 1)
                if (g.D > 1 && ns > 0.)
                    acc += -vp.gammaH * 2. / (DD - 1.) * dirac_root(ns * DD / PI_D, g.D) / (DD * ns);
            } else if (t == -1 || t == -12) {
                acc += xc_lda_dvdn(t, vp.units, ns);
            }
        }
        out[i] = vp.inc * acc;
    }
}
hipError_t vint_dvdn(const Grid &g, const VintParams &vp, const double *den, double *out, hipStream_t st) {
    hipLaunchKernelGGL(k_vint_dvdn, dim3(2048), dim3(256), 0, st, g, vp, den, out);
    return hipGetLastError();
}

// DipolarInteractionPos (mpDPFT.cpp:2749-2755), 3D: v = alpha (v + n0 zf(|r|) / (8 pi^3))
// with zf the transform of the Gaussian the k-space multiplier subtracted.  A
// pass of its own after the backward transform: inside the transform's store
// the two extra transcendentals put a stack frame on every backward transform.
__global__ __launch_bounds__(256) void k_dipolar_comp(Grid g, double n0, double alpha, double *v) {
    const int EL = g.EL;
    const double sp = sqrt(PI_D);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.G; i += (size_t)gridDim.x * blockDim.x) {
        const int a = (int)(i / ((size_t)EL * EL)), rem = (int)(i - (size_t)a * EL * EL), b = rem / EL,
                  c = rem - b * EL;
        const double z = sqrt(g.lat[a] * g.lat[a] + g.lat[b] * g.lat[b] + g.lat[c] * g.lat[c]);
        double zf;
        if (z > 1.0e-12)
            zf = (PI_D / (z * z * z)) * (d_exp(-z * z / 4.) * sp * z * (4. + z * z) - 4. * PI_D * erf(z / 2.)) +
                 PI_D * sp * d_exp(-z * z / 4.) * (-1. / 3. + 8. * PI_D / 3.);
        else
            zf = PI_D * sp * (8. * PI_D / 3.);
        v[i] = alpha * (v[i] + (n0 / (8. * PI_D * PI_D * PI_D)) * zf);
    }
}
hipError_t dipolar_comp(const Grid &g, double n0, double alpha, double *v, hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)std::min<size_t>((g.G + 255) / 256, 4096);
    hipLaunchKernelGGL(k_dipolar_comp, dim3(blocks), dim3(256), 0, st, g, n0, alpha, v);
    return hipGetLastError();
}

// DipolarInteraction (mpDPFT.cpp:2969-2990), 3D, momentum space: one block per
// grid point q_i sums, over the grid points k_j in a fixed order, the trapezoid
// weight times k_z/|k|^2 - 1/3 where p = q_i + k_j lies inside |p| < edge/2 and
// n(p) > n(q_i), p rounded to its grid index as IndexOfVec (:3083) does.
__global__ __launch_bounds__(256) void k_dipolar_k(Grid g, const double *__restrict__ n, double scale,
                                                   double *out) {
    __shared__ double sh[4];
    const int EL = g.EL;
    const double x0 = -g.edge / 2.;
    for (size_t i = blockIdx.x; i < g.G; i += gridDim.x) {
        double q[3];
        coords(g, (int)i, q);
        const double rhop = n[i];
        double acc = 0.;
        for (size_t j = threadIdx.x; j < g.G; j += 256) {
            double k[3], pk[3];
            coords(g, (int)j, k);
            for (int d = 0; d < 3; d++) pk[d] = q[d] + k[d];
            double tf = 0.;
            if (sqrt(pk[0] * pk[0] + pk[1] * pk[1] + pk[2] * pk[2]) < 0.5 * g.edge) {
                size_t idx = 0;
                for (int d = 0; d < 3; d++) {
                    const int ci = (int)(0.5 + (double)(EL - 1) * ((pk[d] - x0) / g.edge));
                    idx = idx * (size_t)EL + (size_t)ci;
                }
                if (n[idx] - rhop > 0.) tf = k[2] / (k[0] * k[0] + k[1] * k[1] + k[2] * k[2]) - 1. / 3.;
            }
            acc += trap_weight(g, (int)j) * tf;
        }
        const double r = block_sum(acc, sh);
        if (threadIdx.x == 0) out[i] = scale * r;
        __syncthreads();
    }
}
hipError_t dipolar_k(const Grid &g, const double *n, double scale, double *out, hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    const unsigned blocks = (unsigned)std::min<size_t>(g.G, 65536);
    hipLaunchKernelGGL(k_dipolar_k, dim3(blocks), dim3(256), 0, st, g, n, scale, out);
    return hipGetLastError();
}

// ---- the n3' densities (DensityExpression 6 and 10) -------------------------
// Bessel J_n(x), n = 0..3, x >= 0: the power series for x < 2, Miller's
// backward recurrence from order 2 floor(x/2) + 44 normalised by
// J_0 + 2 sum J_2k = 1 for 2 <= x < 25, Hankel's asymptotic expansion with the
// phase x - (2n+1) pi/4 taken through cos x, sin x for x >= 25 (absolute error
// below 5e-16 against an independent implementation).
__device__ double bessel_jn(int n, double x) {
    if (x < 2.) {
        double h = 0.5 * x, t = 1., q = -h * h, sum = 0.;
        for (int k = 1; k <= n; k++) t *= h / (double)k;
        for (int k = 0; k < 40; k++) {
            sum += t;
            t *= q / ((double)(k + 1) * (double)(k + 1 + n));
            if (fabs(t) < 1e-18 * fabs(sum)) break;
        }
        return sum;
    }
    if (x < 25.) {
        const int N = 2 * (int)(0.5 * x) + 44;
        double jp = 0., jc = 1e-30, norm = 0., jn = 0.;
        for (int k = N; k >= 1; k--) {
            const double jm = (2. * (double)k / x) * jc - jp;
            jp = jc;
            jc = jm;
            if (k - 1 == n) jn = jc;
            if (k - 1 > 0 && ((k - 1) & 1) == 0) norm += 2. * jc;
        }
        norm += jc;
        return jn / norm;
    }
    const double mu = 4. * (double)n * (double)n;
    double P = 0., Q = 0., a = 1., xk = 1.;
    for (int k = 0; k < 60; k++) {
        const double term = a / xk;
        if (k > 0 && fabs(term) < 1e-17) break;
        if ((k & 1) == 0) P += ((k / 2) & 1) ? -term : term;
        else Q += (((k - 1) / 2) & 1) ? -term : term;
        a *= (mu - (double)(2 * k + 1) * (double)(2 * k + 1)) / ((double)(k + 1) * 8.);
        xk *= x;
        if (fabs(a / xk) > fabs(term) && k > 2) break;
    }
    const double c0 = (n == 0 || n == 3) ? 1. : -1., c1 = (n <= 1) ? 1. : -1.;  // cos, sin of (2n+1) pi/4, over 1/sqrt(2)
    double sx, cx;
    sincos(x, &sx, &cx);
    const double r2 = 0.70710678118654752440;
    const double cw = r2 * (c0 * cx + c1 * sx), sw = r2 * (c0 * sx - c1 * cx);
    return sqrt(2. / (PI_D * x)) * (P * cw - Q * sw);
}

// Getn3pFFT (mpDPFT.cpp:3625-3684): one block per standard-order spectrum entry
// f; the entry holds the centred momentum q at index (f + s/2) mod s per axis,
// C(q) = pref sum_r e^{-i q.r} (Uf (mu - V(r)) - q^2/8)_+^{D/2} over the (s+1)^D
// points r = x0 + index dx, with the centring sign of the backward transform.
__global__ __launch_bounds__(256) void k_n3p_spec(Grid g, const double *__restrict__ V, double mu, double Uf,
                                                 double pref, double gph, double2 *__restrict__ spec) {
    __shared__ double sh[4];
    const int s = g.s, EL = g.EL, D = g.D;
    const size_t f = blockIdx.x;
    int qi[3] = {0, (int)(f / s) % s, (int)(f % s)};
    if (D == 3) qi[0] = (int)(f / ((size_t)s * s));
    const double dx = g.edge / (double)s, x0 = -g.edge / 2., k0 = -PI_D / dx, dk = 2. * PI_D / g.edge;
    double q[3] = {0., 0., 0.}, q2 = 0.;
    int par = 0;
    for (int d = 3 - D; d < 3; d++) {
        par += qi[d];
        q[d] = k0 + (double)((qi[d] + s / 2) % s) * dk;
        q2 += q[d] * q[d];
    }
    const double q2term = 0.125 * q2;
    double re = 0., im = 0.;
    for (size_t r = threadIdx.x; r < g.G; r += 256) {
        const double t = Uf * (mu - V[r]) - q2term;
        if (!(t > 0.)) continue;
        int ri[3] = {0, 0, 0};
        size_t rem = r;
        for (int d = 2; d >= 3 - D; d--) {
            ri[d] = (int)(rem % EL);
            rem /= EL;
        }
        const double amp = D == 2 ? t : t * sqrt(t);
        double arg = -q[3 - D] * (x0 + (double)ri[3 - D] * dx);
        for (int d = 4 - D; d < 3; d++) arg -= q[d] * (x0 + (double)ri[d] * dx);
        double sn, cs;
        sincos(arg, &sn, &cs);
        re += cs * amp;
        im += sn * amp;
    }
    re = block_sum(re, sh);
    im = block_sum(im, sh);
    if (threadIdx.x == 0) {
        const double sg = (par & 1) ? -gph : gph;
        spec[f] = make_double2(sg * pref * re, sg * pref * im);
    }
}
hipError_t n3p_spec(const Grid &g, const double *V, double mu, double Uf, double pref, double2 *spec,
                    hipStream_t st) {
    const double gph = (((long long)g.D * g.s / 2) & 1) ? -1. : 1.;
    hipLaunchKernelGGL(k_n3p_spec, dim3((unsigned)g.F), dim3(256), 0, st, g, V, mu, Uf, pref, gph, spec);
    return hipGetLastError();
}

// Getn3p (mpDPFT.cpp:3702-3729): one block per grid point i, n_i = g dx^D sum_j
// w_j (PV_j / (2 pi r_ij))^D J_D(2 r_ij PV_j), PV_j = sqrt(2 m (mu - V_j)) where
// mu > V_j, and (PV_j^2 / pi)^D / (8, 48) (D = 2, 3) for r_ij <= 1e-16.
__global__ __launch_bounds__(256) void k_n3p_bessel(Grid g, const double *__restrict__ V, double mu, double mass,
                                                   double scale, double *__restrict__ n) {
    __shared__ double sh[4];
    const int D = g.D;
    const int i = (int)blockIdx.x;
    double ri[3], rj[3];
    coords(g, i, ri);
    const double gam = D == 2 ? 0.125 : 0.02083333333333333;
    double acc = 0.;
    for (int j = threadIdx.x; j < (int)g.G; j += 256) {
        double PV = mu - V[j];
        if (!(PV > 0.)) continue;
        PV = sqrt(2. * mass * PV);
        coords(g, j, rj);
        double rp = (rj[0] - ri[0]) * (rj[0] - ri[0]) + (rj[1] - ri[1]) * (rj[1] - ri[1]);
        if (D == 3) rp += (rj[2] - ri[2]) * (rj[2] - ri[2]);
        double si;
        if (rp > 1e-16) {
            rp = sqrt(rp);
            const double z = 2. * rp * PV;
            const double J = z > 2.0e+12 ? sqrt(2. / (PI_D * z)) * cos(z - 0.5 * (double)D * PI_D - 0.25 * PI_D)
                                         : bessel_jn(D, z);
            const double b = PV / (2. * PI_D * rp);
            si = (D == 2 ? b * b : b * b * b) * J;
        } else {
            const double b = PV * PV / PI_D;
            si = (D == 2 ? b * b : b * b * b) * gam;
        }
        acc += trap_weight(g, j) * si;
    }
    acc = block_sum(acc, sh);
    if (threadIdx.x == 0) n[i] = scale * acc;
}
hipError_t n3p_bessel(const Grid &g, const double *V, double mu, double mass, double scale, double *n,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_n3p_bessel, dim3((unsigned)g.G), dim3(256), 0, st, g, V, mu, mass, scale, n);
    return hipGetLastError();
}

// trapezoid-weighted partial sums of f into partials[0..kRedBlocks)
__global__ __launch_bounds__(256) void k_trap_sum(Grid g, const double *__restrict__ f, double *partials) {
    __shared__ double sh[4];
    double acc = 0.;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < (int)g.G; i += gridDim.x * 256) acc += trap_weight(g, i) * f[i];
    const double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}
hipError_t trap_sum(const Grid &g, const double *f, double *partials, hipStream_t st) {
    hipLaunchKernelGGL(k_trap_sum, dim3(kRedBlocks), dim3(256), 0, st, g, f, partials);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// convergence, stats
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_conv_dots(Grid g, const double *a, const double *b, double *partials) {
    __shared__ double sh[4];
    int s = blockIdx.y;
    const double *pa = a + (size_t)s * g.G, *pb = b + (size_t)s * g.G;
    double aa = 0., bb = 0., ab = 0.;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < g.G; i += (size_t)gridDim.x * 256) {
        double x = pa[i], y = pb[i];
        aa += x * x;
        bb += y * y;
        ab += x * y;
    }
    double r0 = block_sum(aa, sh);
    __syncthreads();
    double r1 = block_sum(bb, sh);
    __syncthreads();
    double r2 = block_sum(ab, sh);
    if (threadIdx.x == 0) {
        partials[(size_t)(3 * s + 0) * kRedBlocks + blockIdx.x] = r0;
        partials[(size_t)(3 * s + 1) * kRedBlocks + blockIdx.x] = r1;
        partials[(size_t)(3 * s + 2) * kRedBlocks + blockIdx.x] = r2;
    }
}

__global__ __launch_bounds__(256) void k_field_minmax(Grid g, const double *den, const double *V, double *partials) {
    __shared__ double sh[4];
    int s = blockIdx.y;
    const double *pd = den + (size_t)s * g.G, *pv = V + (size_t)s * g.G;
    double m0 = 1e308, m1 = 1e308, m2 = 1e308, m3 = 1e308;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < g.G; i += (size_t)gridDim.x * 256) {
        double x = pd[i], y = pv[i];
        m0 = fmin(m0, x);
        m1 = fmin(m1, -x);
        m2 = fmin(m2, y);
        m3 = fmin(m3, -y);
    }
    double r[4];
    r[0] = block_min(m0, sh);
    __syncthreads();
    r[1] = block_min(m1, sh);
    __syncthreads();
    r[2] = block_min(m2, sh);
    __syncthreads();
    r[3] = block_min(m3, sh);
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; k++) partials[(size_t)(4 * s + k) * kRedBlocks + blockIdx.x] = r[k];
}

__global__ __launch_bounds__(256) void k_max_abs_diff(size_t n, const double *a, const double *b, double *partials) {
    __shared__ double sh[4];
    double m = 1e308;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) m = fmin(m, -fabs(a[i] - b[i]));
    double r = block_min(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// ---------------------------------------------------------------------------
// Pulay
// ---------------------------------------------------------------------------
__global__ void k_pulay_push(size_t n, const double *nin, const double *nout, double *snin, double *sres) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double a = nin[i], b = nout[i];
        snin[i] = a;
        sres[i] = b - a;
    }
}

struct PtrList {
    const double *p[64];
};
// Pulay Gram row <y_j, x> (j < nh) on the FP64 matrix cores: per tile of 16
// history rows, v_mfma_f64_16x16x4f64 multiplies A (16 rows x 4 elements, lane
// l holding row l%16's element at slot l/16) by B (4 x 16, only column 0 = x
// nonzero: lanes 0, 16, 32, 48) into C; C[i][0] accumulates <y_i, x>.  The 4
// element slots of one MFMA are the lane groups' own contiguous runs of kMdT
// elements (lane group g takes [c*4kMdT + g*kMdT, +kMdT) of chunk c), so every
// element of x meets the same element of each row exactly once.  C[i][j] sits in
// lane j + 16*(i%4), register i/4 (A[i][k] in lane i + 16k, B[k][j] in lane
// j + 16k), so C[i][0] is in lane 16*(i%4).  Per block the four waves' rows are
// summed in a fixed order.
typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));
constexpr int kMdT = 4;
__global__ __launch_bounds__(256) void k_multi_dot(size_t n, const double *__restrict__ x, PtrList ys, int nh,
                                                   double *partials) {
    __shared__ double sh[4][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const size_t chunk = 4 * kMdT;
    for (int base = 0; base < nh; base += 16) {
        const double *row = base + r < nh ? ys.p[base + r] : nullptr;
        mfma_f64x4 acc = {0., 0., 0., 0.};
        for (size_t c = (size_t)blockIdx.x * 4 + wave; c * chunk < n; c += (size_t)gridDim.x * 4) {
            const size_t k0 = c * chunk + (size_t)g * kMdT;
            double a[kMdT], b[kMdT];
#pragma unroll
            for (int t = 0; t < kMdT; t++) {
                const size_t k = k0 + t;
                a[t] = (row && k < n) ? row[k] : 0.;
                b[t] = (r == 0 && k < n) ? x[k] : 0.;
            }
#pragma unroll
            for (int t = 0; t < kMdT; t++) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[t], b[t], acc, 0, 0, 0);
        }
        if (r == 0) {
#pragma unroll
            for (int v = 0; v < 4; v++) sh[wave][g + 4 * v] = acc[v];
        }
        __syncthreads();
        if (threadIdx.x < 16 && base + (int)threadIdx.x < nh) {
            const int i = threadIdx.x;
            partials[(size_t)(base + i) * kRedBlocks + blockIdx.x] = (sh[0][i] + sh[1][i]) + (sh[2][i] + sh[3][i]);
        }
        __syncthreads();
    }
}

// x' = w(q) x over S packed spectra of F points: w = 2^(number of axes at the
// Nyquist index s/2), the multiplicity of that mode among the (s+1)^D entries of
// the reference's wrap-filled centred spectrum.
__global__ void k_weight_nyquist(Grid g, int S, const double2 *x, double2 *out) {
    const size_t F = g.F, tot = F * (size_t)S;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
        size_t q = idx % F;
        double w = 1.;
        for (int d = 0; d < g.D; d++) {
            if ((int)(q % (size_t)g.s) == g.s / 2) w *= 2.;
            q /= (size_t)g.s;
        }
        const double2 v = x[idx];
        out[idx] = double2{w * v.x, w * v.y};
    }
}

__global__ __launch_bounds__(256) void k_pulay_mix(size_t G, int S, int I, PtrList nin, PtrList res,
                                                   const double *coeff, const double *theta, double *den) {
    size_t tot = G * (size_t)S;
    for (size_t idx = blockIdx.x * 256 + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * 256) {
        int s = (int)(idx / G);
        double th = theta[s];
        double acc = 0.;
        for (int p = 0; p < I; p++) acc += coeff[p] * (nin.p[p][idx] + th * res.p[p][idx]);
        den[idx] = acc;
    }
}

// ImposeNoise (mpDPFT.cpp:4490-4504): den *= the host-drawn factors 1 + Noise RN
__global__ void k_mul_fields(size_t n, const double *f, double *den) {
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < n; idx += (size_t)gridDim.x * blockDim.x)
        den[idx] *= f[idx];
}

__global__ void k_linear_mix(size_t G, int S, const double *theta, const double *old, double *den) {
    size_t tot = G * (size_t)S;
    for (size_t idx = blockIdx.x * (size_t)blockDim.x + threadIdx.x; idx < tot; idx += (size_t)gridDim.x * blockDim.x) {
        int s = (int)(idx / G);
        double th = theta[s];
        den[idx] = th * den[idx] + (1. - th) * old[idx];
    }
}

// ---------------------------------------------------------------------------
// DelField helpers, environments, energies
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_shell_max(Grid g, const double *f, double rt,
                                                   double *partials) {
    __shared__ double sh[4];
    double m = 1e308;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < g.G; i += (size_t)gridDim.x * 256) {
        double r[3];
        coords(g, (int)i, r);
        double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        if (fabs(rn - rt) < 2. * g.dx) m = fmin(m, -f[i]);
    }
    double r = block_min(m, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}


__global__ void k_sum_squares3(size_t G, const double *a, const double *b, const double *c, double *out) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < G; i += (size_t)gridDim.x * blockDim.x) {
        double v = a[i] * a[i];
        if (b) v += b[i] * b[i];
        if (c) v += c[i] * c[i];
        out[i] = v;
    }
}

__global__ void k_env(Grid g, EnvParams e, double *env) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < g.G; i += (size_t)gridDim.x * blockDim.x) {
        double r[3];
        coords(g, (int)i, r);
        double r2 = 0.;
        for (int d = 0; d < g.D; d++) r2 += r[d] * r[d];
        double v = 0.;
        if (e.type == 1) v = e.c0 * pow(r2, 1.);
        else if (e.type == 3) v = e.c0 * ((1. - e.c1) * r[0] * r[0] + (1. + e.c1) * r[1] * r[1]);
        else if (e.type == 40) v = -e.c0 / sqrt(r2 + e.c1 * e.c1);
        else if (e.type == 41) {
            for (int c = 0; c < e.ncentres; c++) {
                const double *C = e.centres + 5 * c;
                double dd = 0.;
                for (int d = 0; d < g.D; d++) dd += (r[d] - C[d]) * (r[d] - C[d]);
                v += -C[3] / sqrt(dd + C[4] * C[4]);
            }
        }
        env[i] = v;
    }
}

__global__ __launch_bounds__(256) void k_energy(Grid g, EnergyParams ep, const double *a,
                                                const double *b, const double *c, const double *d, double *partials) {
    __shared__ double sh[4];
    double acc = 0., DD = (double)g.D;
    for (size_t i = blockIdx.x * 256 + threadIdx.x; i < g.G; i += (size_t)gridDim.x * 256) {
        double v = 0.;
        switch (ep.op) {
        case E_PROD: v = a[i] * b[i]; break;
        case E_HALF_PROD: v = 0.5 * a[i] * b[i]; break;
        case E_DIRAC: v = ep.p0 * pow(d_pos(a[i]), 1. + 1. / DD); break;
        case E_XC: {
            double e, vv;
            xc_lda(ep.xc_id, ep.units, a[i], &e, &vv);
            v = d_pos(a[i]) * e;
            break;
        }
        case E_TF_DISP: {
            double t = ep.p0 - a[i];
            v = t > 0. ? pow(t, (DD + 2.) / 2.) : 0.;
            break;
        }
        case E_POW_POS: v = pow(d_pos(a[i]), ep.p0); break;
        case E_GOMBAS: {
            double vv;
            gombas(a[i], ep.p0, &v, &vv);
            break;
        }
        case E_RENORM: {
            const double e0 = renorm_eta(a[i], ep.p0), e1 = renorm_eta(b[i], ep.p0);
            v = PI_D * (d_pos(a[i]) + 1.0e-16) * (d_pos(b[i]) + 1.0e-16) * (qmc_beta(0, e0) + qmc_beta(0, e1));
            break;
        }
        case E_QT2: {
            const double t = ep.p0 - a[i];
            const double q = 2. * 0.131163 * t;
            v = t > 0. ? q * q : 0.;
            break;
        }
        case E_VWN: {
            const double d = d_pos(a[i]);
            v = 0.;
            if (d > 1.0e-15) {
                double F, dF;
                vwn_F(vwn_x(d), &F, &dF);
                v = d * kVA * F;
            }
            break;
        }
        case E_CROSSOVER: v = crossover(a[i], ep.p0, 1); break;
        case E_QUASI2D: v = quasi2d(a[i], b[i], ep.p0, ep.p1, ep.eta, ep.eta_rows, 1); break;
        case E_GRADCORR: {
            double rho = fabs(a[i]);
            v = rho > ep.p0 ? b[i] / rho : 0.;
            break;
        }
        case E_VIRIAL: {
            double r[3];
            coords(g, (int)i, r);
            v = a[i] * (r[0] * b[i] + r[1] * c[i] + r[2] * d[i]);
            break;
        }
        }
        acc += trap_weight(g, (int)i) * v;
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void k_fill(size_t n, double v, double *a) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) a[i] = v;
}

// ---------------------------------------------------------------------------
// host wrappers
// ---------------------------------------------------------------------------
static inline unsigned gblocks(size_t n) {
    size_t b = (n + 255) / 256;
    return (unsigned)(b > 4096 ? 4096 : (b < 1 ? 1 : b));
}

hipError_t conv_dots_S(const Grid &g, int S, const double *a, const double *b, double *partials, hipStream_t st) {
    hipLaunchKernelGGL(k_conv_dots, dim3(kRedBlocks, S), dim3(256), 0, st, g, a, b, partials);
    return hipGetLastError();
}
hipError_t field_minmax_S(const Grid &g, int S, const double *den, const double *V, double *partials,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_field_minmax, dim3(kRedBlocks, S), dim3(256), 0, st, g, den, V, partials);
    return hipGetLastError();
}
hipError_t max_abs_diff(size_t n, const double *a, const double *b, double *partials, hipStream_t st) {
    hipLaunchKernelGGL(k_max_abs_diff, dim3(kRedBlocks), dim3(256), 0, st, n, a, b, partials);
    return hipGetLastError();
}
hipError_t pulay_push(size_t n, const double *nin, const double *nout, double *slot_nin, double *slot_res,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_pulay_push, dim3(gblocks(n)), dim3(256), 0, st, n, nin, nout, slot_nin, slot_res);
    return hipGetLastError();
}
hipError_t multi_dot(size_t n, const double *x, const double *const *ys, int nh, double *partials,
                     hipStream_t st) {
    if (nh > 64) return hipErrorInvalidValue;
    PtrList pl;
    for (int j = 0; j < 64; j++) pl.p[j] = j < nh ? ys[j] : nullptr;
    hipLaunchKernelGGL(k_multi_dot, dim3(kRedBlocks), dim3(256), 0, st, n, x, pl, nh, partials);
    return hipGetLastError();
}
hipError_t weight_nyquist(const Grid &g, int S, const double2 *x, double2 *out, hipStream_t st) {
    hipLaunchKernelGGL(k_weight_nyquist, dim3(gblocks(g.F * S)), dim3(256), 0, st, g, S, x, out);
    return hipGetLastError();
}
hipError_t pulay_mix(const Grid &g, int S, int I, const double *const *nin, const double *const *res,
                     const double *coeff_dev, const double *theta_dev, double *den, hipStream_t st) {
    if (I > 64) return hipErrorInvalidValue;
    PtrList a, b;
    for (int j = 0; j < 64; j++) {
        a.p[j] = j < I ? nin[j] : nullptr;
        b.p[j] = j < I ? res[j] : nullptr;
    }
    hipLaunchKernelGGL(k_pulay_mix, dim3(gblocks(g.G * S)), dim3(256), 0, st, g.G, S, I, a, b, coeff_dev,
                       theta_dev, den);
    return hipGetLastError();
}
hipError_t mul_fields(size_t n, const double *factors, double *den, hipStream_t st) {
    hipLaunchKernelGGL(k_mul_fields, dim3(gblocks(n)), dim3(256), 0, st, n, factors, den);
    return hipGetLastError();
}
hipError_t linear_mix(size_t n_per_species, int S, const double *theta_dev, const double *old, double *den,
                      hipStream_t st) {
    hipLaunchKernelGGL(k_linear_mix, dim3(gblocks(n_per_species * S)), dim3(256), 0, st, n_per_species, S,
                       theta_dev, old, den);
    return hipGetLastError();
}
__global__ void k_taper_factor(Grid g, double rt, double *T) {
    double a = -2500. / (g.edge * g.edge);
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < g.G; i += (size_t)gridDim.x * blockDim.x) {
        double r[3];
        coords(g, (int)i, r);
        double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        T[i] = rn > rt ? d_exp(a * (rn - rt) * (rn - rt)) : 1.;
    }
}
__global__ void k_shell_mark(Grid g, double rt, int *flag) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < g.G; i += (size_t)gridDim.x * blockDim.x) {
        double r[3];
        coords(g, (int)i, r);
        double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        flag[i] = fabs(rn - rt) < 2. * g.dx ? 1 : 0;
    }
}
// Block minima, then the last block to finish (ticket counter) reduces them and
// resets the counter: one launch.
__global__ __launch_bounds__(256) void k_shell_max_list(const double *f, const int *idx, int n, double *partials,
                                                        unsigned *ticket, double *out) {
    __shared__ double sh[4];
    __shared__ bool last;
    double m = 1e308;
    // eight list entries per thread and round: their index loads, then their
    // value loads, are in flight together (the gathers are latency bound)
    constexpr int U = 8;
    const int stride = gridDim.x * 256;
    for (int j0 = blockIdx.x * 256 + threadIdx.x; j0 < n; j0 += U * stride) {
        int id[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int j = j0 + u * stride;
            id[u] = j < n ? idx[j] : -1;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
            if (id[u] >= 0) m = fmin(m, -f[id[u]]);
    }
    double r = block_min(m, sh);
    if (threadIdx.x == 0) {
        partials[blockIdx.x] = r;
        __threadfence();
        last = atomicAdd(ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last) {
        __threadfence();
        double v = 1e308;
        for (int j = threadIdx.x; j < (int)gridDim.x; j += 256) v = fmin(v, partials[j]);
        double t = block_min(v, sh);
        if (threadIdx.x == 0) {
            *out = t;
            *ticket = 0u;
        }
    }
}
hipError_t taper_factor(const Grid &g, double rt, double *T, hipStream_t st) {
    hipLaunchKernelGGL(k_taper_factor, dim3(gblocks(g.G)), dim3(256), 0, st, g, rt, T);
    return hipGetLastError();
}
hipError_t shell_mark(const Grid &g, double rt, int *flag, hipStream_t st) {
    hipLaunchKernelGGL(k_shell_mark, dim3(gblocks(g.G)), dim3(256), 0, st, g, rt, flag);
    return hipGetLastError();
}
hipError_t shell_max_list(const double *f, const int *idx, int n, double *partials, unsigned *ticket, double *out,
                          hipStream_t st) {
    constexpr int kBlocks = 64;
    hipLaunchKernelGGL(k_shell_max_list, dim3(kBlocks), dim3(256), 0, st, f, idx, n, partials, ticket, out);
    return hipGetLastError();
}
hipError_t shell_max(const Grid &g, const double *f, double rt, double *partials, hipStream_t st) {
    hipLaunchKernelGGL(k_shell_max, dim3(kRedBlocks), dim3(256), 0, st, g, f, rt, partials);
    return hipGetLastError();
}
hipError_t sum_squares3(const Grid &g, const double *a, const double *b, const double *c, double *out,
                        hipStream_t st) {
    hipLaunchKernelGGL(k_sum_squares3, dim3(gblocks(g.G)), dim3(256), 0, st, g.G, a, b, c, out);
    return hipGetLastError();
}
hipError_t environment(const Grid &g, const EnvParams &e, double *env, hipStream_t st) {
    hipLaunchKernelGGL(k_env, dim3(gblocks(g.G)), dim3(256), 0, st, g, e, env);
    return hipGetLastError();
}
hipError_t energy_integrand(const Grid &g, const EnergyParams &ep, const double *a, const double *b,
                            const double *c, const double *d, double *partials, hipStream_t st) {
    hipLaunchKernelGGL(k_energy, dim3(kRedBlocks), dim3(256), 0, st, g, ep, a, b, c, d, partials);
    return hipGetLastError();
}
hipError_t fill(size_t n, double v, double *a, hipStream_t st) {
    hipLaunchKernelGGL(k_fill, dim3(gblocks(n)), dim3(256), 0, st, n, v, a);
    return hipGetLastError();
}

}  // namespace mk
