/*
 * dpft_oracle.h -- ORACLE: test infrastructure only.
 *
 * A plain-C FP64 CPU restatement of the reference's self-consistent DPFT loop
 * (huehou/mpDPFT, mpDPFT.cpp GetData :135-244 and the routines it calls).  It
 * exists to check the HIP product and to time a CPU baseline; only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product never links or calls it.
 */
#ifndef DPFT_ORACLE_H
#define DPFT_ORACLE_H
#include "../include/mpdpft.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_state orc_state;

/* Full solve: GetData with TaskType 0.  den_out/v_out are [S][G] (may be NULL). */
int oracle_solve(const mpdpft_params *p, mpdpft_result *r, double *den_out, double *v_out);

/* Stage entry points on a fresh state built from p. */
int oracle_fft_centered(const mpdpft_params *p, int sign, double *cfield /* [(s+1)^D][2] */);
int oracle_dft_brute(int n, int sign, const double *in, double *out); /* 1D, unnormalised */
int oracle_fft1d(int n, int sign, const double *in, double *out);     /* 1D, unnormalised */
int oracle_delfield(const mpdpft_params *p, const double *f, int method, double *dx, double *dy,
                    double *dz, double *gradsq, double *lap);
int oracle_vint(const mpdpft_params *p, int type, int s, const double *den, double *vint);
int oracle_density(const mpdpft_params *p, int s, double mu, const double *v, double *den,
                   double *N);
int oracle_env(const mpdpft_params *p, double *env /* [S][G] */);
double oracle_integrate(const mpdpft_params *p, const double *f);
void oracle_pw92(double n, double *eps_c, double *v_c);
/* Pulay small solve: full-pivoting Householder QR inverse of the I x I Gram
 * matrix A (row-major), validation and normalised coefficients.  Returns 1 on
 * success, 0 when the reference would fall back to linear mixing. */
int oracle_pulay_coefficients(int I, const double *A, double *coeff, double *testval);
/* Hartree by direct O(G^2) summation of the isolated-boundary convolution
 * (checks the zero-padded FFT path on small grids). */
int oracle_hartree_isolated_direct(const mpdpft_params *p, const double *den, double *vh);
/* Cell-averaged 1/r at the origin of the isolated Hartree kernel, in units of
 * 1/dx: (1/dx^3) * integral over the cube [-dx/2,dx/2]^3 of 1/|r|. */
double oracle_cube_inverse_r(void);

/* nAiT special functions: CurlyA(y) = int_y^inf Ai, Li_{3/2-j}(-e^x) (j = 0..3),
 * Ai(y), as the density and energy use them. */
double oracle_curlyA(double y);
double oracle_polylog(int j, double x);
double oracle_airy(double y);
/* The nAiT density at one point (Units 1, degeneracy 1): U = V - mu, ar =
 * |grad V|^(2/3)/2, lap = Laplacian of V, temperature T; finite-T TF form
 * below the gradient threshold thr, else the adaptive Boole Airy average
 * with BP panels, maxIter levels and relative accuracy acc (no
 * negligibility exit). */
double oracle_nait_point(double U, double ar, double lap, double T, double thr, int BP, int maxIter, double acc);
/* EkinAiT03D's kinetic energy density at one point (Units 1): U = V - mu,
 * a = |grad V|^(2/3)/2, lap = Laplacian of V; TF form below the gradient
 * threshold thr, else the refined Airy average (no negligibility exit). */
double oracle_ekin_point(double U, double a, double lap, double thr, double acc, int MaxIt);
/* the fitted b(eta) (der 0) and b'(eta) (der 1) of the renormalised 2D contact */
double oracle_qmc_beta(int der, double x);
/* e^-x I0(x) for x >= 0 (2D periodic Hartree compensation) */
double oracle_exp_i0(double x);
/* Bessel J_n(x), n = 0..3, x >= 0 (Getn3p's kernel) */
double oracle_bessel_jn(int n, double x);
/* TabFunc_2D3D_eta.dat (eta triplets of the quasi-2D contact, interaction 24) */
int oracle_load_eta(const char *path);
/* energy integrand of interaction 14 (species s) or 24 (species s focal) */
int oracle_energy_density(const mpdpft_params *p, int type, int s, const double *den, double *out);
/* Gram matrix of I residuals [I][S][G] in the Pulay momentum-space metric (Mixer[2] != 0, DIM > 1) */
int oracle_pulay_momentum_gram(const mpdpft_params *p, int I, const double *res, double *M);

/* Timed partial workload for the CPU baseline: n_iter SC iterations (no
 * convergence test) starting from the initial density; returns seconds. */
int oracle_enforce_trace(const mpdpft_params *p, const double *v, double *mu, double *N, double *den,
                         int *nobracket, double *trace_mu, double *trace_N, int *trace_phase, int trace_max,
                         int *trace_n);
int oracle_convergence_sequence(const mpdpft_params *p, int n, const double *dens, const double *Vs, const double *Ns,
                                const double *mus, double *CC);
int oracle_enforce(const mpdpft_params *p, const double *v, double *mu, double *N, double *den, int *nobracket);
int oracle_mix_sequence(const mpdpft_params *p, int nsteps, const double *n_in0, const double *n_outs, double *out,
                        int *rejections);
double oracle_time_iterations(const mpdpft_params *p, int n_iter, int *bisection_steps);
/* ImposeNoise's generator: MT19937-64 and the uniform draw on [-1, 1) */
void oracle_mt64_seed(unsigned long long *mt, int *mti, unsigned long long seed);
unsigned long long oracle_mt64_next(unsigned long long *mt, int *mti);
double oracle_rn(unsigned long long *mt, int *mti);

#ifdef __cplusplus
}
#endif
#endif
