// kernels.h -- host-callable launchers of the gfx950 kernels (kernels.hip).
// All launchers enqueue on the given stream and return the launch status.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace mk {

// Geometry of one (s+1)^D grid and its s^D periodic cell.
struct Grid {
    int D;        // 2 or 3
    int s;        // steps (FFT length per axis)
    int EL;       // s + 1
    size_t G;     // EL^D
    size_t F;     // s^D
    double edge, dx;
    const double *lat;  // device lattice table x_i = -edge/2 + i*edge/s, length EL
};

// ---- FFT: centred transforms between real grid fields and a packed spectrum
// (standard DFT order, s^D complex).  spec = DFT(f) (unnormalised, no
// centring signs); the reference's centred spectrum is
// dx^D (-1)^{qx+qy+qz} spec[q].
hipError_t fft_forward_real(const Grid &g, const double *re, const double *im, double2 *spec,
                            const double2 *tw, double *f0slot, hipStream_t st);

// Backward multipliers applied while loading the spectrum (first pass).
enum MulKind {
    MUL_NONE = 0,        // plain
    MUL_SMOOTH = 1,      // exp(sigma |k|^2)
    MUL_DERIV_PAIR = 2,  // re <- i k_a e^{sigma k^2}, im <- i k_b e^{sigma k^2} (Nyquist plane of the axis zeroed)
    MUL_DERIV_LAP = 3,   // re <- i k_z e^{sigma k^2}, im <- -|k|^2 e^{2 sigma k^2}
    MUL_HARTREE_PER = 4, // reference periodic Hartree with Gaussian compensation
    MUL_SECOND = 5,      // re <- -k_a^2 (method 4 Laplacian pieces)
    MUL_HARTREE_2D = 6,  // 2D periodic Hartree: (f - f0 e^{-k^2}) / |k|, 0 for |k| <= dkzero
    MUL_DIPOLAR = 7      // 3D dipolar (position space): (f - n0 e^{-k^2}) (k_z^2/k^2 - 1/3 + 8 pi/3), 0 for k^2 <= dkzero
};
struct MulParams {
    int kind;
    int axis_a, axis_b;   // axes for MUL_DERIV_PAIR (0..2 in 3D; 1..2 in 2D)
    double sigma;
    const double *ek;     // EXP(sigma k^2) per standard index (the smoothing factor is separable:
                          // exp(sigma |k|^2) = prod_d ek[q_d]); null: sigma = 0
    const double *kq;     // k per standard index (kLattice values), length s
    const double *kq_h;   // k per standard index (k0 + p dk), Hartree
    const double *f0slot; // device scalar: dx^D sum n (Hartree)
    const double2 *kek;   // optional (k, EXP(sigma k^2)) per standard index: replaces kq and ek lookups
    double dxD;           // dx^D
    double dkzero;        // 1e-6 dk^2
    int regularize;       // RegularizeCoulombSingularity
    double n0;            // MUL_DIPOLAR: the subtracted Gaussian's amplitude
};
// Backward transform: f = L^-D * IDFT(M * spec) written to re/im real grid
// fields (either may be null) with periodic-image fill, out = scale*value
// (+ out if accumulate).  post: 0 none, 1 POS clip of re.
// post 2 adds the Gaussian compensation f0 erf(r/2)/(4 pi r) and multiplies
// by c1 = 4 pi gammaH (needs lat, the lattice table, and m.f0slot).
// The first pass reads 'spec' and writes 'work' (may equal spec), so a
// spectrum can feed several backward transforms.
hipError_t fft_backward_real(const Grid &g, const double2 *spec, double2 *work, const MulParams &m, double *re,
                             double *im, double scale, int post, double c1, const double2 *tw, hipStream_t st);
// Forward transforms with DelField's pointwise preparation fused into the first
// pass: the TaperOff shift (-*shift_dev holds the shell maximum) and taper of V
// (T: taper_factor's table), or the sum of squares of the derivative fields
// (sum_squares3; c may be null).
hipError_t fft_forward_taper(const Grid &g, const double *V, const double *T, const double *shift_dev, double2 *spec,
                             const double2 *tw, hipStream_t st);
hipError_t fft_forward_sumsq(const Grid &g, const double *a, const double *b, const double *c, double2 *spec,
                             const double2 *tw, hipStream_t st);

// Isolated-boundary Hartree: zero-padded (2s)^D circular convolution with the
// precomputed real kernel spectrum khat; out[i] += gammaH dx^D conv (i in [0,s]^D)
// of the sources on [0,s)^D.  transpose = 1 applies the transposed operator
// (sources on [0,s]^D, result on [0,s)^D, 0 at the periodic images).
hipError_t hartree_isolated(const Grid &g, const double *n, double2 *pad, const double *khat,
                            const double2 *tw2, double gamma_dxD_over_M, double *out, int accumulate,
                            hipStream_t st, int transpose = 0);
// Block-shape autotuner: times each lines-per-block choice of the s-point transform
// (a forward transform of 'field') and, with pad given, of the 2s-point isolated
// Hartree convolution, and keeps the fastest for the process (results do not
// depend on the choice: every line is transformed by the same arithmetic).
hipError_t fft_autotune(const Grid &g, const double *field, double2 *spec, const double2 *tw, double2 *pad,
                        const double *khat, const double2 *tw2, double *out, hipStream_t st);
hipError_t fft32_autotune(const Grid &g, const double *field, float2 *spec, const float2 *tw, float2 *pad,
                          const float *khat, const float2 *tw2, double *out, hipStream_t st);
// Build khat = DFT((2s)^D kernel 1/r, g0 at r=0) into pad then copy real part to khat.
hipError_t build_isolated_kernel(const Grid &g, double2 *pad, const double2 *tw2, double g0, double *khat,
                                 hipStream_t st);

// ---- reductions: partial sums per block, then a fixed-order final sum.
constexpr int kRedBlocks = 1024;
constexpr int kRedThreads = 256;
// out[k] for k < nres; 'partials' holds kRedBlocks*nres doubles.
hipError_t finish_sums(const double *partials, int nres, double *out, double scale, hipStream_t st);
hipError_t finish_minmax(const double *partials, int nres, double *out, hipStream_t st);

// TF density n = pref (mu - V)_+^{D/2} with fused trapezoid weights; writes n
// if n != nullptr; partial sums of w*n into partials[0..kRedBlocks).
hipError_t tf_density(const Grid &g, const double *V, double mu, double pref, double *n, double *partials,
                      hipStream_t st);
// Evaluate N(mu_j) for nmu chemical potentials in one pass over V (no n written).
hipError_t tf_density_multi(const Grid &g, const double *V, const double *mus_dev, int nmu, double pref,
                            double *partials, hipStream_t st);

// V_s = inc * (H_s + sum of pointwise interactions) + Env_s, interactions in
// list order; 'hartree_in_V' says V already holds the Hartree term.
struct VintParams {
    int n_inter;
    int inter[16];
    int S, s;
    double beta, gammaH, inc;
    int units;
    int hartree_in_V;
    int fp32;  // pointwise interaction potentials evaluated in FP32
    double alpha;                 // omega_z / mu0 mu^2 prefactor (interactions 14, 24)
    const double *eta;            // [eta_rows][3] eta triplets of interaction 24
    int eta_rows;
    const double *dip_r, *dip_k;  // [S][G] dipolar potentials of 17 and 15 (computed beforehand)
};
hipError_t update_v(const Grid &g, const VintParams &vp, const double *den_all, const double *env, double *V,
                    hipStream_t st);
// out = inc * sum_a dV_a/dn of the pointwise interactions (10, 11, -1, -12) of one
// species at density den (the diagonal of the interaction Jacobian without Hartree).
hipError_t vint_dvdn(const Grid &g, const VintParams &vp, const double *den, double *out, hipStream_t st);
// DipolarInteraction (momentum space, 3D): out[i] = scale sum_j w_j (k_z/|k|^2 - 1/3) over the grid
// points k_j with |q_i + k_j| < edge/2 and n(q_i + k_j) > n(q_i); w_j the trapezoid weights.
hipError_t dipolar_k(const Grid &g, const double *n, double scale, double *out, hipStream_t st);
// DipolarInteractionPos's last step (3D): v = alpha (v + n0 zf(|r|)/(8 pi^3)) in place, v the
// backward transform of the multiplied spectrum.
hipError_t dipolar_comp(const Grid &g, double n0, double alpha, double *v, hipStream_t st);
// Getn3pFFT's momentum field (DensityExpression 6) as the standard-order spectrum that the
// backward transform (scale L^-D, plain store) turns into the density.
hipError_t n3p_spec(const Grid &g, const double *V, double mu, double Uf, double pref, double2 *spec,
                    hipStream_t st);
// Getn3p (DensityExpression 10): n_i = scale sum_j w_j K(r_ij, PV_j), scale = g dx^D.
hipError_t n3p_bessel(const Grid &g, const double *V, double mu, double mass, double scale, double *n,
                      hipStream_t st);
// trapezoid-weighted partial sums of f into partials[0..kRedBlocks)
hipError_t trap_sum(const Grid &g, const double *f, double *partials, hipStream_t st);

// RelDev dots per species + field stats: partials layout per species s:
// [<a,a>, <b,b>, <a,b>] sums and [denmin, denmax, vmin, vmax].
hipError_t conv_dots_S(const Grid &g, int S, const double *a, const double *b, double *partials, hipStream_t st);
hipError_t field_minmax_S(const Grid &g, int S, const double *den, const double *V, double *partials,
                          hipStream_t st);
hipError_t max_abs_diff(size_t n, const double *a, const double *b, double *partials, hipStream_t st);

// Pulay helpers over n = S*G entries.
hipError_t pulay_push(size_t n, const double *nin, const double *nout, double *slot_nin, double *slot_res,
                      hipStream_t st);
// partials[j*kRedBlocks + blk] = partial <res_new, res_j>, j < nh (FP64 matrix cores)
hipError_t multi_dot(size_t n, const double *x, const double *const *ys, int nh, double *partials,
                     hipStream_t st);
// out = w(q) x for S packed spectra, w = 2^(axes at the Nyquist index): the multiplicity of
// each mode among the (s+1)^D entries of the wrap-filled centred spectrum (Pulay momentum metric)
hipError_t weight_nyquist(const Grid &g, int S, const double2 *x, double2 *out, hipStream_t st);
// den[s] = sum_p c_p (nin_p + theta_s res_p)
hipError_t pulay_mix(const Grid &g, int S, int I, const double *const *nin, const double *const *res,
                     const double *coeff_dev, const double *theta_dev, double *den, hipStream_t st);
hipError_t linear_mix(size_t n_per_species, int S, const double *theta_dev, const double *old, double *den,
                      hipStream_t st);
// den[i] *= factors[i], i < n (ImposeNoise with host-drawn factors)
hipError_t mul_fields(size_t n, const double *factors, double *den, hipStream_t st);

// DelField's TaperOff shift by the shell maximum (mpDPFT.cpp:7251-7264); the taper itself is
// applied in fft_forward_taper's first pass.
hipError_t shell_max(const Grid &g, const double *f, double rt, double *partials, hipStream_t st);
// The taper's geometry, once per grid: T[i] = EXP(-2500/L^2 (r - rt)^2) beyond rt, else 1
// (DelField's taper, mpDPFT.cpp:7251-7264); flag[i] = 1 on the shell |r - rt| < 2 dx (k_shell_max's test).
hipError_t taper_factor(const Grid &g, double rt, double *T, hipStream_t st);
hipError_t shell_mark(const Grid &g, double rt, int *flag, hipStream_t st);
// *out = min over j < n of -f[idx[j]] (k_shell_max over a precomputed shell list) in one launch;
// partials: 64 doubles; ticket: a device counter that starts at 0 (left at 0)
hipError_t shell_max_list(const double *f, const int *idx, int n, double *partials, unsigned *ticket, double *out,
                          hipStream_t st);
hipError_t sum_squares3(const Grid &g, const double *a, const double *b, const double *c, double *out,
                        hipStream_t st);

// Environment fields.
struct EnvParams {
    int type;
    double c0, c1;           // mpp[s], mpp[S+s]
    int ncentres;
    const double *centres;   // device [n][5]
};
hipError_t environment(const Grid &g, const EnvParams &e, double *env, hipStream_t st);

// Energy integrands (weighted trapezoid sums into partials[0..kRedBlocks)).
enum EnergyOp {
    E_PROD = 0,          // a*b
    E_DIRAC = 1,         // dirac energy density of a (scale passed)
    E_XC = 2,            // POS(a) * eps_xc(a)
    E_TF_DISP = 3,       // (mu - a)_+^{(D+2)/2}
    E_POW_POS = 4,       // POS(a)^p
    E_GRADCORR = 5,      // b/|a| if |a| > thr else 0   (a = n, b = grad2)
    E_VIRIAL = 6,        // a * (x b + y c + z d)
    E_HALF_PROD = 7,     // 0.5 a b
    E_GOMBAS = 8,        // Gombas correlation energy density of a, scale p0 (= gammaH)
    E_VWN = 9,           // VWN correlation energy density of a (eV, density in 1/Angstrom^3)
    E_QT2 = 10,          // (2 * 0.131163 * (p0 - a))^2 where p0 > a, else 0 (a = V, p0 = mu)
    E_RENORM = 11,       // renormalised 2D contact energy density of (a, b) = (n_0, n_1), p0 = scattering length
    E_CROSSOVER = 12,    // 2D-3D crossover energy density of a, p0 = alpha
    E_QUASI2D = 13       // quasi-2D contact energy density of (a, b) = (n_0, n_1), p0 = alpha, p1 = beta, eta
};
struct EnergyParams {
    int op;
    double p0, p1;       // op parameters (mu, power, threshold, scale)
    int xc_id, units;
    const double *eta;   // E_QUASI2D: [eta_rows][3]
    int eta_rows;
};
hipError_t energy_integrand(const Grid &g, const EnergyParams &ep, const double *a, const double *b,
                            const double *c, const double *d, double *partials, hipStream_t st);

// ---- FP32 transforms of the precision variant (kernels_fp32.hip; 3D only).
// Spectra, twiddles, multipliers and accumulation in FP32; grid fields FP64.
// fft32_forward: V, or (V - max) T when T is given (shift_dev as fft_forward_taper).
hipError_t fft32_forward(const Grid &g, const double *V, const double *T, const double *shift_dev, float2 *spec,
                         const float2 *tw, hipStream_t st);
hipError_t fft32_forward_sumsq(const Grid &g, const double *a, const double *b, const double *c, float2 *spec,
                               const float2 *tw, hipStream_t st);
// kind: MUL_DERIV_PAIR / MUL_DERIV_LAP / MUL_SMOOTH / MUL_NONE with kek[q] = (k, smoothing factor)
hipError_t fft32_backward(const Grid &g, const float2 *spec, float2 *work, int kind, int axis_a, int axis_b,
                          const float2 *kek, double *re, double *im, float scale, int post, const float2 *tw,
                          hipStream_t st);
hipError_t hartree_isolated32(const Grid &g, const double *n, float2 *pad, const float *khat, const float2 *tw2,
                              float scale, double *out, hipStream_t st);
hipError_t to_float(size_t n, const double *a, float *b, hipStream_t st);
hipError_t to_float2(size_t n, const double2 *a, float2 *b, hipStream_t st);

// Fill a device array with a value.
hipError_t fill(size_t n, double v, double *a, hipStream_t st);

}  // namespace mk
