/*
 * mpdpft.h -- C ABI of the MI355X-native self-consistent DPFT loop.
 *
 * Plain C: pointers, sizes and POD structs only.  Every entry point returns an
 * int status (MPDPFT_OK or a negative MPDPFT_E* code) and never exits the
 * process.  A context owns its device buffers and one HIP stream; it is not
 * thread-safe, distinct contexts may be driven from distinct host threads.
 *
 * Field layout (host and device): row-major over the (steps+1)^DIM grid with x
 * slowest, index = i*EL*EL + j*EL + k, EL = steps+1; per-species fields are
 * stacked species-major ([S][G]).  This is the layout of the reference's
 * datastruct fields (mpDPFT.h:71-89, mpDPFT.cpp:7169-7186).
 */
#ifndef MPDPFT_H
#define MPDPFT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: mpdpft_result gained AuxEint; 3: mpdpft_result gained bracket_failures and pulay_rejections */
#define MPDPFT_ABI_VERSION 5

#define MPDPFT_MAX_SPECIES 8
#define MPDPFT_MAX_LIST 16
#define MPDPFT_MAX_MPP 64
#define MPDPFT_MAX_CENTRES 256

/* status codes */
#define MPDPFT_OK 0
#define MPDPFT_EINVAL (-1)     /* invalid parameter or unsupported feature       */
#define MPDPFT_ENOMEM (-2)     /* host or device allocation failed               */
#define MPDPFT_EHIP (-3)       /* a HIP runtime call failed                       */
#define MPDPFT_ENOTCONV (-4)   /* SC loop hit maxSCcount (results still valid)    */
#define MPDPFT_ENOBRACKET (-5) /* mu bracket not found (mpDPFT.cpp:4657-4661)     */
#define MPDPFT_EPULAY (-6)     /* Pulay Gram singular, fell back to linear mixing */
#define MPDPFT_EIO (-7)        /* file could not be read or written               */
#define MPDPFT_ENODEV (-8)     /* no HIP device                                   */

/* Extension environment types (the reference's GetEnv knows 0-25, 30, 61;
 * mpDPFT.cpp:1385-1698).  40: softened Coulomb -mpp[s]/sqrt(r^2+mpp[S+s]^2)
 * centred at the origin.  41: sum over the centres list of softened Coulomb
 * wells -Z_c/sqrt(|r-R_c|^2+a_c^2). */
#define MPDPFT_ENV_SOFT_COULOMB 40
#define MPDPFT_ENV_SOFT_COULOMB_CENTRES 41

/* Interaction ids follow mpDPFT.input line 27 (mpDPFT.cpp:2240-2318):
 * 0 none, 1 contact, 8 Hartree, 10 single-species contact, 11 Dirac exchange,
 * -1 LDA exchange, -12 PW92 LDA correlation (native, replaces libxc). */

/*
 * All parameters of one solve.  Fields 'System'..'Print' mirror the 47
 * positional lines of mpDPFT.input as parsed by GetInputParameters
 * (mpDPFT.cpp:621-721); the rest are the hard-coded defaults of
 * InitializeHardCodedParameters (mpDPFT.cpp:858-1383, mpDPFT.h:79) exposed as
 * overridable values, and the extension keys.  mpdpft_params_default() fills
 * every field; mpdpft_parse_input() overwrites what the file states.
 */
typedef struct mpdpft_params {
    int abi_version;
    /* mpDPFT.input, lines 1-47 */
    int System;
    int Symmetry;
    int TaskType;
    int TaskParameterCount;
    int DIM;
    int Units;
    double edge;
    int steps;
    int S;
    double Abundances[MPDPFT_MAX_SPECIES];
    double RelAcc;
    double InternalAcc;
    int K;
    int DensityExpression;
    int n_mpp;
    double mpp[MPDPFT_MAX_MPP];
    int n_Mpp;
    double Mpp[MPDPFT_MAX_LIST];
    double tauVec[MPDPFT_MAX_SPECIES];
    double tVec[MPDPFT_MAX_SPECIES];
    double TVec[MPDPFT_MAX_SPECIES];
    int Environments[MPDPFT_MAX_SPECIES];
    double stretchfactor;
    double Wall;
    double Noise;
    int InterpolVQ;
    int n_interactions;
    int Interactions[MPDPFT_MAX_LIST];
    double incrementalV;
    double muVec[MPDPFT_MAX_SPECIES];
    double DeltamuModifier;
    double Mixer[3];
    double thetaVec[MPDPFT_MAX_SPECIES];
    double SCcriterion;
    int maxSCcount;
    int method;
    int DelFieldMethod;
    int MovieQ;
    int n_schedule;
    int Schedule[MPDPFT_MAX_LIST];
    int n_ekin;
    int EkinTypes[MPDPFT_MAX_LIST];
    double alpha;
    double beta;
    double gamma;
    double gammaH;
    double degeneracy;
    int n_regularize;
    int regularize[MPDPFT_MAX_LIST];
    double RegularizationThreshold;
    int Print;
    /* hard-coded defaults of the reference, overridable */
    int TaperOff;                    /* mpDPFT.h:79, default 1 */
    int RegularizeCoulombSingularity;/* mpDPFT.h:79, default 1 */
    double SmearingExtentDx;         /* mpDPFT.cpp:1105, default 6 (x Deltax) */
    int OutPrecision;                /* mpDPFT.cpp:1363, default 16 */
    int MaxIteration;                /* mpDPFT.cpp:1101, default 4 (EkinAiT03D) */
    /* extension keys ("value //Key" lines after line 47) */
    int HartreeMode;       /* 0: periodic + Gaussian compensation (reference), 1: isolated, zero-padded 2s grid */
    double SmoothingKWidth;/* >0: DelField smoothing exp(-k^2/(2 w^2)); <=0: reference exp(-(6dx)^2 k^2/2) */
    int ConvergenceMode;   /* 0: reference overlap rule CC<SCcriterion; 1: max|V-V_old| < SCcriterion */
    int Precision;         /* 0: FP64 throughout; 1: FP32 transforms, multipliers and pointwise interaction
                              potentials (FP64 fields, reductions, mu, Pulay, density quadrature, energies)
                              until the FP32 floor, then FP64 polish iterations; DIM 3, isolated Hartree */
    int WriteFiles;        /* 1: mpdpft_run_dir writes Den/V/Control files (default 1) */
    int n_centres;         /* environment 41 */
    double centres[MPDPFT_MAX_CENTRES][5]; /* x, y, z, Z, a */
    /* mpDPFT.input line 5 (TaskHyperBox): (lower, upper) pairs of the task
     * variables, the ranges of TaskType 2/3's starting point and scale */
    int n_hyperbox;        /* number of values (twice the pairs) */
    double TaskHyperBox[2 * MPDPFT_MAX_SPECIES];
    unsigned long long NoiseSeed; /* extension key: seed of the Noise generator (mt19937_64, uniform on
                              [-1, 1)); 0 seeds it from the clock as the reference does; TaskType 10's
                              repetition k uses NoiseSeed + k */
    int CubeFile;          /* extension key: 1 writes mpDPFT_Den_Cube.dat (DIM 3) in place of mpDPFT_Den.dat,
                              the environment-41 centres as nuclei */
} mpdpft_params;

/* Results of one solve (the quantities the reference prints in its final
 * block, mpDPFT.cpp:12716-12800). */
typedef struct mpdpft_result {
    int S;
    int SCcount;          /* loop counter at exit; the file prints SCcount-1 */
    int converged;        /* 1 if ExitSCcount reached 3 */
    int LinMix, PulMix;
    int errorCount, warningCount;
    double CC;
    double muVec[MPDPFT_MAX_SPECIES];
    double FinalAbundances[MPDPFT_MAX_SPECIES];
    double DeltamuModifier;
    double TotalEnergy;
    double DispersalEnergies[MPDPFT_MAX_SPECIES];
    double EnvironmentalEnergies[MPDPFT_MAX_SPECIES];
    double InteractionEnergies[MPDPFT_MAX_LIST];
    double AuxEkin[MPDPFT_MAX_SPECIES][4];
    double solve_seconds;  /* wall time of the SC loop (device-synchronised) */
    int bisection_steps;   /* total density evaluations inside mu bisection */
    int fp32_iterations;   /* Precision 1: SC iterations run with the FP32 stages (the rest are FP64) */
    /* GetAuxEint (mpDPFT.cpp:5518-5549): for S = 1, DIM 3, Units 2 the Dirac exchange energy, its
       potential-functional form, the Gombas and the VWN correlation energies; 0 otherwise */
    double AuxEint[MPDPFT_MAX_SPECIES][4];
    /* recoverable events of the solve (neither is an error; the reference only prints them):
       EnforceConstraints calls whose mu bracket was not found, mu reset to its previous value
       (mpDPFT.cpp:4657-4661); PulayMixer solves rejected by the invertibility or 1e-8 validation
       test, the iteration then mixing linearly (:5089-5107) */
    int bracket_failures;
    int pulay_rejections;
} mpdpft_result;

typedef struct mpdpft_ctx mpdpft_ctx;

/* ---- parameters and files ------------------------------------------------ */
int mpdpft_abi_version(void);
/* sizeof(mpdpft_params) and sizeof(mpdpft_result), for FFI bindings to check layouts. */
size_t mpdpft_params_sizeof(void);
size_t mpdpft_result_sizeof(void);
/* Defaults for every field (the reference's hard-coded values). */
void mpdpft_params_default(mpdpft_params *p);
/* Parse an mpDPFT.input file (47 positional lines + optional extension keys);
 * replaces GetInputParameters, mpDPFT.cpp:621-721. */
int mpdpft_parse_input(const char *path, mpdpft_params *p);

/* ---- context ---------------------------------------------------------------- */
/* Validates p, allocates device memory, builds grids and environments
 * (InitializeHardCodedParameters + InitializeEnvironments, mpDPFT.cpp:858,
 * 6362).  device < 0 picks the current HIP device. */
int mpdpft_create(const mpdpft_params *p, int device, mpdpft_ctx **out);
void mpdpft_destroy(mpdpft_ctx *ctx);
const char *mpdpft_last_error(const mpdpft_ctx *ctx);
size_t mpdpft_grid_size(const mpdpft_ctx *ctx);

/* Whole self-consistent loop on the device (GetData, mpDPFT.cpp:135-244,
 * loop body :210-225), then energies (ExportData's GetEnergy, :12403). */
int mpdpft_solve(mpdpft_ctx *ctx, mpdpft_result *res);

/* Copy-out of [S][G] fields (caller-owned host arrays of S*G doubles). */
int mpdpft_get_density(mpdpft_ctx *ctx, double *host);
int mpdpft_get_potential(mpdpft_ctx *ctx, double *host);
int mpdpft_get_environment(mpdpft_ctx *ctx, double *host);
/* Replace V on the device (input of the stage entry points below; a solve
 * starts from Env, or from the initial potential set next). */
int mpdpft_set_potential(mpdpft_ctx *ctx, const double *host);

/* ---- warm start and loop control ----------------------------------------- */
/* LoadV/ImportV (mpDPFT.cpp:6430-6502): read a potential file in the
 * mpDPFT_V.dat format ("x [y [z]] V_0 .. V_{S-1}" per line, n^DIM lines) and
 * interpolate it multilinearly (linear extrapolation outside its grid) onto
 * the lattice of p; out: S*G doubles.  Host code only, no GPU.  *warnings
 * (may be NULL) counts the reference's cross-check warnings: column count,
 * NaN entries, integrals of the two fields differing by more than 5%. */
int mpdpft_import_v(const char *path, const mpdpft_params *p, double *out, int *warnings);
/* Potential the next mpdpft_solve starts from instead of V = Env (the
 * InterpolVQ = 1 path, mpDPFT.cpp:184); host [S][G]; NULL restores V = Env. */
int mpdpft_set_initial_potential(mpdpft_ctx *ctx, const double *host);
/* ManualSCloopBreakQ (mpDPFT.cpp:11959-11976): file polled once per SC
 * iteration while Print >= 0; a first integer 1 (or 2) ends the loop after
 * that iteration and the file is rewritten as 0.  NULL or "" turns it off. */
int mpdpft_set_break_file(mpdpft_ctx *ctx, const char *path);

/* ---- stage entry points (one call = one reference routine) ---------------- */
/* Centred continuous-FT approximation on a complex [(s+1)^D][2] field,
 * sign -1 forward / +1 backward (fftParallel, mpDPFT.cpp:8173-8335). */
int mpdpft_fft(mpdpft_ctx *ctx, int sign, double *host_complex);
/* Spectral first/second derivatives of field f (host, [G]) with method 4/5
 * (DelField, mpDPFT.cpp:7243-7582); outputs [G] each, any may be NULL. */
int mpdpft_delfield(mpdpft_ctx *ctx, const double *f, double *dx, double *dy, double *dz,
                    double *gradsq, double *laplacian);
/* Interaction potential of one type for species s given densities [S][G]
 * (getVint, mpDPFT.cpp:2240-2318). */
int mpdpft_interaction_potential(mpdpft_ctx *ctx, int type, int s, const double *den, double *vint);
/* Density of species s at chemical potential mu for potential v [G]
 * (GetDensity, mpDPFT.cpp:2142-2238); returns the trapezoid integral in *N. */
int mpdpft_density(mpdpft_ctx *ctx, int s, double mu, const double *v, double *den, double *N);

/* ---- one iteration at a time, and the loop's own stages ------------------- */
/* mpdpft_loop_begin starts a loop as mpdpft_solve does (V = Env or the
 * initial potential, the initial densities :194-201) without iterating;
 * each mpdpft_step then runs one SC iteration of GetData (:213-225:
 * StoreCurrentIteration, UpdateV, GetDensities, CheckConvergence, Intercept,
 * AdmixDensities, InjectSchedule, PrepareNextIteration) and reports the loop
 * state in r (SCcount, CC, converged, muVec, FinalAbundances, mixing and
 * event counters; energies stay 0) and *exit_loop (ExitSCloop, may be NULL).
 * A step after the loop has exited changes nothing.  Stepping until
 * exit_loop and calling mpdpft_energy gives mpdpft_solve's result. */
int mpdpft_loop_begin(mpdpft_ctx *ctx);
int mpdpft_step(mpdpft_ctx *ctx, mpdpft_result *r, int *exit_loop);
/* GetEnergy (mpDPFT.cpp:5414-5436) of the context's current density,
 * potential and mu; r receives the whole result (energies and loop state). */
int mpdpft_energy(mpdpft_ctx *ctx, mpdpft_result *r);
/* GetDensities + EnforceConstraints (mpDPFT.cpp:3306-3316, :4569-4743) on
 * the context's current potential (mpdpft_set_potential or the loop's):
 * N_s at mu[s], then the mu bracketing and bisection to |N - N_0| / 2N_0 <=
 * RelAcc with the context's DeltamuModifier and RelAcc state (carried from
 * call to call as across SC iterations).  mu (in/out, S values), N (out, may
 * be NULL); the density of the last evaluated mu becomes the context's
 * density (mpdpft_get_density).  Returns MPDPFT_ENOBRACKET when a species'
 * bracket was not found (its mu reset to the value it entered with). */
int mpdpft_enforce_constraints(mpdpft_ctx *ctx, double *mu, double *N);
/* AdmixDensities with PulayMixer (mpDPFT.cpp:4921-4992, :5005-5167):
 * n_out (host [S][G]) is the newest density, the context's current density
 * the n_in of this step; the mixed density replaces the context's density.
 * sc_count is the loop's SCcount (0: linear mixing, as the first iteration
 * is); the Pulay history is the context's (mpdpft_loop_begin resets it).
 * Returns MPDPFT_EPULAY when the Pulay solve was rejected and the step mixed
 * linearly (the mixed density is valid). */
int mpdpft_mix(mpdpft_ctx *ctx, int sc_count, const double *n_out);
/* Forcing hooks for the loop's recovery paths (no counterpart in the
 * reference, whose tests have none): at SC iteration `at` of the next solves,
 * MPDPFT_FAULT_DUPLICATE_PULAY pushes the newest Pulay entry twice (a
 * singular Gram matrix: the solve is rejected and the iteration mixes
 * linearly, :5089-5107); MPDPFT_FAULT_NAN_POTENTIAL writes NaN into species
 * 0's potential at the grid centre after UpdateV (the NaN reaches N(mu), the
 * bracket is not found and the loop runs on without converging, as the
 * reference's loop would for these density expressions).  kind 0 or at < 0
 * clears the hook. */
#define MPDPFT_FAULT_DUPLICATE_PULAY 1
#define MPDPFT_FAULT_NAN_POTENTIAL 2
int mpdpft_inject_fault(mpdpft_ctx *ctx, int kind, int at);

/* ---- per-stage GPU timing (the reference's StartTimer/EndTimer, :11997) --- */
/* Kernel time per stage, measured with HIP events on the context's stream.
 * 'units' counts the stage's work items (grid points, integrand evaluations,
 * transforms); 'bytes' and 'flops' are the algorithmic HBM traffic and
 * arithmetic of those launches (the roofline numerators). */
typedef struct mpdpft_stage_stats {
    char name[32];
    long long launches;
    double total_ms;
    double units;
    double bytes;
    double flops;
} mpdpft_stage_stats;
int mpdpft_enable_timing(mpdpft_ctx *ctx, int on);
int mpdpft_get_stage_stats(mpdpft_ctx *ctx, mpdpft_stage_stats *out, int max, int *n);

/* ---- derivatives of the converged solution -------------------------------- */
/* No counterpart in the reference, which differentiates by finite differences
 * over whole solves (its task loops, mpDPFTmain.cpp:77-214, mpDPFT.cpp:8365-8584).
 *
 * mpdpft_adjoint: for a loss L(n) of the density of the last converged solve,
 * given dL/dn ([G], one species, plain sum over grid points), returns
 * lambda[i] = dL/dEnv_i, the derivative with respect to an additive change of
 * the external potential at grid point i, with the particle number held
 * fixed (or mu held, when RelAcc = 0 switches the mu search off).  It linearises the fixed point V = Env + K[n], n = F(V, mu),
 * sum w n = N (implicit differentiation, no unrolled iterations) and solves
 * (I - F_V^T P K'^T) lambda = F_V^T P dL/dn by restarted GMRES on the device.
 * Supported: DIM 3, S = 1, DensityExpression 7 (TF) or 5 (nAiT, DelFieldMethod
 * 4/5 with its taper), interactions 8 (HartreeMode 1), 10, 11, -1, -12.
 * info (may be NULL): tol / max_iterations in (0: 1e-12 / 400), the rest out. */
typedef struct mpdpft_adjoint_info {
    double tol;               /* in: relative residual target of the linear solve */
    int max_iterations;       /* in: GMRES iteration cap */
    int iterations;           /* out */
    double relative_residual; /* out */
    double dN_dmu;            /* out: -sum w dn/dU, the particle-number response to mu */
} mpdpft_adjoint_info;
int mpdpft_adjoint(mpdpft_ctx *ctx, const double *dL_dn, double *lambda, mpdpft_adjoint_info *info);
/* Environment 41: out[3c + d] = sum_i lambda[i] dEnv_i/dR_{c,d} (chain rule of
 * mpdpft_adjoint's lambda to the centre coordinates). */
int mpdpft_centre_gradient(mpdpft_ctx *ctx, const double *lambda, double *out);
/* Environment 41: derivative of the last solve's total energy (GetEnergy,
 * mpDPFT.cpp:5414-5436) with respect to the centre coordinates through the
 * fixed point: out[3c + d] = sum_i (w_i dx^3 n_i + lambda_E,i) dEnv_i/dR_{c,d},
 * lambda_E from the adjoint system with the energy's own sources (dE/dn of the
 * interaction and environment energies, dE_disp/dV and dE_disp/dmu of the
 * kinetic energy's quadrature).  EkinTypes 0 only. */
int mpdpft_centre_forces(mpdpft_ctx *ctx, double *out);
/* The task variable's gradient for one species (replaces the finite-difference
 * gradient that ConjugateGradientDescent's mincg takes over whole solves of
 * functionToBeMinimized, mpDPFT.cpp:8365-8584): out[0] = dE/dN of the last
 * solve's total energy through the fixed point when the loop holds N
 * (RelAcc > 0; TaskType 3, GetGlobalMinimum_Abundances :256), or dE/dmu when it
 * holds mu (RelAcc = 0; TaskType 2, GetGlobalMinimum_muVec :246).  Same
 * support as mpdpft_centre_forces, any environment. */
int mpdpft_task_gradient(mpdpft_ctx *ctx, double *out);
/* Replace the centres list (n x {x, y, z, Z, a}) and rebuild Env of the
 * environment-41 species; the next solve starts from the new Env. */
int mpdpft_set_centres(mpdpft_ctx *ctx, int n, const double *centres);

/* ---- batched solves: many independent instances in one launch ------------ */
/* The jobs the reference runs as separate runY directories (mpDPFT.sh,
 * mpDPFTmain.cpp:46-86 once per job), here in one persistent launch: one
 * workgroup per running instance executes the whole SC loop of GetData
 * (:210-225) and takes the next instance from the queue when it retires one.
 * Instances: DensityExpression 7 (TF), DIM 2 or 3, S <= 2, interactions 0/1
 * (contact), environments 0/1/3, Pulay mixer 0/1 with scope <= 16.  Results are
 * what mpdpft_solve reports for the same instance (muVec, FinalAbundances,
 * energies, SCcount, mixing counts). */
#define MPDPFT_BATCH_FP32 1 /* FP32 fields; FP64 integrals, mu, Pulay solve, energies; then an FP64 polish */
typedef struct mpdpft_batch mpdpft_batch;
int mpdpft_batch_create(const mpdpft_params *instances, int n, int device, int flags, mpdpft_batch **out);
int mpdpft_batch_solve(mpdpft_batch *b, mpdpft_result *results);
/* Converged density and potential of instance i ([S][G] host arrays; either may be NULL). */
int mpdpft_batch_get_fields(mpdpft_batch *b, int i, double *den, double *V);
double mpdpft_batch_seconds(const mpdpft_batch *b);
/* Wall seconds, summed iteration counts and summed mu-bisection density
 * evaluations of the last solve's two phases (each argument may be NULL):
 * [0] the FP32 loop (zero without MPDPFT_BATCH_FP32), [1] the FP64 loop or polish.
 * The FP32 loop stops at max|dV| < max(criterion, 1e-6 max|V|) (density mode:
 * 1e-12) or after 16 iterations without a new smallest CC, and its mu search
 * at |N - N0| / 2N0 <= max(RelAcc, 1e-8); the polish then
 * runs the instance's own criterion, so results[i].SCcount counts polish
 * iterations only. */
int mpdpft_batch_phase_stats(const mpdpft_batch *b, double *seconds, long long *iterations,
                             long long *bisection_steps);
/* The last solve's k_batch_solve launches (the persistent kernel), each argument may be NULL:
 * kernel_ms[phase] is the launch's duration from HIP events recorded on the batch's stream around
 * it; stage_bytes[phase] the loop's algorithmic traffic counted per reference stage (inputs +
 * outputs of UpdateV, every N(mu) evaluation, the density, CheckConvergence, the Pulay Gram row
 * (I + 1 fields) and combination (2I + 1) or linear mixing (3), energies, fields out), at the
 * fields' storage width, summed over the instances; kernel_bytes[phase] what the fused sweeps of
 * this kernel access (fewer: the density is never stored, one sweep serves several stages);
 * phase [0] the FP32 loop and [1] the FP64 loop or polish; instance_us[i] (n values) the time
 * instance i spent in the last phase between leaving the queue and retiring (device clock,
 * 100 MHz). */
int mpdpft_batch_kernel_stats(const mpdpft_batch *b, double *kernel_ms, double *stage_bytes, double *kernel_bytes,
                              double *instance_us);
const char *mpdpft_batch_last_error(const mpdpft_batch *b);
void mpdpft_batch_destroy(mpdpft_batch *b);

/* ---- batched solve without device allocation ----------------------------- */
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  mpdpft_batch_create + mpdpft_batch_solve +
 * mpdpft_batch_get_fields in one call, same kernels and same results, with
 * no hipMalloc, hipFree, stream or event creation: the instance descriptors,
 * results, queue, team areas, the FP32 phase's start copy and each running
 * workgroup's field slot live in one static device array of
 * MPDPFT_NOALLOC_SCRATCH_BYTES per GPU, which the library holds until the
 * process ends.  The solve runs as many workgroups at once as slots of the
 * largest instance fit beside the descriptors (a 256^2 two-species instance
 * with Pulay scope 8 needs 22 MB per slot, a 64^2 one 1.4 MB); instances that
 * leave no room for one slot are rejected with MPDPFT_EINVAL.
 *   instances  n host parameter sets (n <= MPDPFT_NOALLOC_MAX_INSTANCES), as
 *              for mpdpft_batch_create; flags 0 or MPDPFT_BATCH_FP32
 *   fields     device memory of `device` (fields_len doubles) that receives
 *              each instance's converged [S][G] density then [S][G] potential,
 *              instance after instance (instance i at sum_{j<i} 2 S_j G_j)
 *   results    n host results, as mpdpft_batch_solve reports them
 *   err        optional (may be NULL) buffer of err_len bytes for the message
 *              of a failure
 * Runs on the device's null stream and returns after the solve finished.
 * Every call on one GPU uses that GPU's one array: calls that overlapped on a
 * GPU would race, so the library runs a process's calls one at a time; other
 * processes hold their own arrays. */
#define MPDPFT_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
#define MPDPFT_NOALLOC_MAX_INSTANCES 4096
int mpdpft_batch_solve_noalloc(const mpdpft_params *instances, int n, int device, int flags, double *fields,
                               size_t fields_len, mpdpft_result *results, char *err, size_t err_len);

/* ---- drop-in: run in a directory like ./mpDPFT --------------------------- */
/* Reads dir/mpDPFT.input (and dir/mpDPFT_V.dat when InterpolVQ = 1), polls
 * dir/mpDPFTmanualSCloopBreakQ.dat, solves, writes dir/mpDPFT_Den.dat,
 * mpDPFT_V.dat and mpDPFT_Control.dat in the reference formats (ExportData,
 * :12399-12520).  Sets the data directory (below) to dir. */
int mpdpft_run_dir(const char *dir, int device, mpdpft_result *res);

/* Directory of the TabFunc_*.dat tables that contexts read at mpdpft_create
 * (process-wide; "" or NULL = the current directory, where the reference reads
 * them): TabFunc_2D3D_eta.dat for interaction 24 (Load2D3Deta, :2927-2950). */
int mpdpft_set_data_dir(const char *dir);

#ifdef __cplusplus
}
#endif
#endif
