// solver.h -- the context behind the C ABI (internal).
#pragma once
#include <hip/hip_runtime.h>

#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "kernels.h"
#include "mpdpft.h"

struct NaitTables;    // nait.hip
struct AdjointState;  // adjoint.hip

struct mpdpft_ctx {
    mpdpft_params p;
    mk::Grid g;
    int S = 1, D = 3, s = 0, EL = 0;
    size_t G = 0, F = 0, SG = 0, CentreIndex = 0;
    double dx = 0, k0 = 0, dk = 0, kdel = 0, sigma = 0;
    int device = 0;
    hipStream_t st = nullptr;

    std::vector<double> lat_h;  // lattice x_i
    // device buffers
    double *d_lat = nullptr, *d_kq = nullptr, *d_kqh = nullptr;
    double *d_ek = nullptr;  // EXP(sigma kq^2) per standard index (DelField smoothing)
    double2 *d_kek = nullptr;  // (kq, ek) per standard index
    double *d_taperT = nullptr;   // DelField taper factor per grid point (built on first use)
    int *d_shell_idx = nullptr;   // grid points of the taper's shell |r - rt| < 2 dx
    int n_shell = -1;             // their number (-1: not built)
    unsigned *d_ticket = nullptr; // completion counter of the one-launch shell maximum
    // precision variant (Precision 1): FP32 transforms until the FP32 floor, then FP64 polish
    float2 *d_specf = nullptr, *d_workf = nullptr, *d_padf = nullptr, *d_twf = nullptr, *d_tw2f = nullptr,
           *d_kekf = nullptr;
    float *d_khatf = nullptr;
    int fp32_phase = 0, fp32_iterations = 0, fp32_stall = 0, fp32_hits = 0;
    double fp32_best = 0.;
    double2 *d_tw = nullptr, *d_tw2 = nullptr;
    double2 *d_spec = nullptr, *d_work = nullptr, *d_pad = nullptr;
    double *d_khat = nullptr;
    double *d_Env = nullptr, *d_V = nullptr, *d_OldV = nullptr, *d_Den = nullptr, *d_OldDen = nullptr;
    double *d_Dx = nullptr, *d_Dy = nullptr, *d_Dz = nullptr, *d_Grad2 = nullptr, *d_Lap = nullptr;
    double *d_tmp1 = nullptr, *d_tmp2 = nullptr, *d_tmp3 = nullptr;
    double *d_partials = nullptr, *d_scal = nullptr, *d_coeff = nullptr, *d_theta = nullptr;
    double *d_centres = nullptr, *d_mus = nullptr;
    double *h_scal = nullptr;  // pinned
    std::vector<void *> allocations;

    // Pulay history ring (scope+1 slots of S*G)
    int pulay_mixer = 0, scope = 0, pI = 0;
    std::vector<double *> nin, res;
    std::vector<double> ResMat;  // (scope+1)^2
    // Pulay momentum-space metric (Mixer[2] != 0, DIM > 1): each entry's residual
    // spectra [S][F] beside it and their k = 0 values on the host
    int pulay_momentum = 0;
    double pulay_weight = 0.;
    std::vector<double2 *> resF;
    std::vector<double> resF0;   // [scope+1][S][2]
    double2 *d_specx = nullptr;  // [S][F] Nyquist-weighted spectrum of the newest residual
    double *h_spec0 = nullptr;   // pinned [S][2]
    double coeff[64];

    // SC state (host scalars)
    double muVec[MPDPFT_MAX_SPECIES], TmpAb[MPDPFT_MAX_SPECIES], OldAb[MPDPFT_MAX_SPECIES];
    double OldmuVec[MPDPFT_MAX_SPECIES], PreviousmuVec[MPDPFT_MAX_SPECIES], thetaVec[MPDPFT_MAX_SPECIES];
    double TVec[MPDPFT_MAX_SPECIES];
    double DenMin[MPDPFT_MAX_SPECIES], DenMax[MPDPFT_MAX_SPECIES], VMin[MPDPFT_MAX_SPECIES],
        VMax[MPDPFT_MAX_SPECIES];
    double DeltamuModifier = 0, RelAcc = 0, RelAccOriginal = 0, incV = 1, incV0 = 1, CC = 0;
    int SCcount = 0, ExitSCcount = 0, ExitSCloop = 0, LinMix = 0, PulMix = 0, TrappedQ = 0;
    int warningCount = 0, errorCount = 0, bisection_steps = 0, PrintSC = 0, MaxIteration = 4;
    int bracket_failures = 0, pulay_rejections = 0;  // recoverable events of the current solve
    int fault_kind = 0, fault_at = -1;               // mpdpft_inject_fault
    int loop_ready = 0;                              // mpdpft_loop_begin has run (Pulay slots allocated)
    double EInt = 0, Etot = 0;
    double IntE[MPDPFT_MAX_LIST], DispE[MPDPFT_MAX_SPECIES], EnvE[MPDPFT_MAX_SPECIES];
    double AuxEkin[MPDPFT_MAX_SPECIES][4];
    double AuxEint[MPDPFT_MAX_SPECIES][4];
    double solve_seconds = 0;
    int has_hartree = 0;
    int verbose = 0;  // MPDPFT_VERBOSE: one line per SC iteration on stderr
    int nait_plain = 0;  // MPDPFT_NAIT_KERNEL=plain: per-node nAiT kernel without panel reuse
    int nait_generic = 0;  // MPDPFT_NAIT_KERNEL=reuse: panel reuse without the fixed lane layout (any BP)
    double *d_eta = nullptr;  // quasi-2D contact's eta table [eta_rows][3] (interaction 24)
    int eta_rows = 0;
    double *d_dip_r = nullptr, *d_dip_k = nullptr;  // [S][G] dipolar potentials of interactions 17 and 15
    std::vector<double> V0;   // initial potential of a solve (empty: V = Env)
    std::ostringstream notes; // lines for the next solve's control stream (LoadV)
    int importWarnings = 0;   // warnings of that import, added to the solve's count
    double fom = 0.;          // GetFigureOfMerit of the last solve (task types 5-7; 0 otherwise)
    // ImposeNoise: the generator of InitializeRandomNumberGenerator (mpDPFT.cpp:6521-6540), seeded per solve,
    // data.Noise (schedules 4, 7 and 11 change it during the loop), NoiseOriginal, and the factors' buffer
    std::mt19937_64 rng;
    double noise = 0., noise0 = 0.;
    double *d_noise = nullptr;
    // MovieQ: frames of the cut (StoreMovieData, mpDPFT.cpp:12317-12329); row i starts with the cut coordinate,
    // then per frame, per species the density, environment and potential at cut point i
    std::vector<std::vector<double>> movie;
    int frames = 0;
    std::string break_file;   // ManualSCloopBreakQ file ("" = not polled)
    double OldCC = 0.;        // CC of the previous iteration (InjectSchedule 5)
    double thetaVec0[MPDPFT_MAX_SPECIES] = {};  // thetaVecOriginal (InjectSchedule 9)
    int delfield_valid = -1;  // species whose Grad2/Lap hold DelField(V_s)

    NaitTables *nait = nullptr;
    AdjointState *adj = nullptr;  // buffers of the fixed-point adjoint (allocated on first use)

    // per-stage GPU timing (HIP events on st)
    int timing = 0;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<mpdpft_stage_stats> stats;

    std::string err;
    std::ostringstream control;
};

// internal helpers shared between translation units
const char *params_structure_error(const mpdpft_params *p);  // params.cpp
// output.cpp: one GetData call in directory d (LoadV, SC loop, ExportData when
// write_files); task_grad (may be null) receives mpdpft_task_gradient of the
// converged context, with its return code in *grad_rc.
// ref (may be null): the n x n TabFunc_ReferenceDensity.dat table; then the figure of merit against it goes into
// *fom (and the Control file).
// v_start (may be null): the initial potential [S][G] (InterpolVQ = 2: task.V); v_out (may be null) receives the
// final potential.
int run_one(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res, bool write_files,
            const std::string &notes, double *mpp_out, double *task_grad = nullptr, int *grad_rc = nullptr,
            const std::vector<double> *ref = nullptr, double *fom = nullptr,
            const std::vector<double> *v_start = nullptr, std::vector<double> *v_out = nullptr);
// output.cpp: the cut through the x-y plane (or the line in 1D): returns its length, fills
// fi[(s 3 + p) EL + i] with species s's density, environment, potential (p = 0, 1, 2) at cut point i
double cut_fields(const mpdpft_ctx *c, const std::vector<double> &den, const std::vector<double> &env,
                  const std::vector<double> &V, std::vector<double> &fi);
// tasks.cpp: TaskType 2 / 3 (the energy's minimum over muVec / the abundances)
int run_cgd_task(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res,
                 std::ostringstream &log);
// tasks.cpp: TaskType 9 (annealed gradient descent over the abundances)
int run_gd_task(const std::string &d, const mpdpft_params &p, int device, mpdpft_result *res,
                std::ostringstream &log);
int ctx_fail(mpdpft_ctx *c, int code, const std::string &msg);
int hip_fail(mpdpft_ctx *c, hipError_t e, const char *what);
int ctx_delfield(mpdpft_ctx *c, const double *f_dev, int s, int method);
int ctx_reduce(mpdpft_ctx *c, int nres, double scale, double *out_host);
void stage_begin(mpdpft_ctx *c);
void stage_end(mpdpft_ctx *c, const char *name, double units, double bytes, double flops);
void stage_add(mpdpft_ctx *c, const char *name, double units, double bytes, double flops);

// adjoint.hip
void adjoint_free(mpdpft_ctx *c);

// nait.hip
int nait_init(mpdpft_ctx *c);
void nait_free(mpdpft_ctx *c);
int nait_density(mpdpft_ctx *c, int s, double mu, double *den_out, double *N);
// n[V, mu] of species s and its partial derivatives d/dU (U = V - mu), d/d|grad V|^2
// and d/d lap V at every grid point (same quadrature tree as the density)
int nait_partials(mpdpft_ctx *c, int s, double mu, double *n, double *dU, double *dg2, double *dl);
int nait_dispersal_energy(mpdpft_ctx *c, int s, double *E);
// d/dU, d/d|grad V|^2, d/d lap V of the EkinAiT03D energy density per point at the
// final MaxIteration; the energy is prefactor * sum_i w_i dx^3 (density)_i
int nait_energy_partials(mpdpft_ctx *c, int s, double *eU, double *eg2, double *el, double *prefactor);
void nait_adapt(mpdpft_ctx *c);
void nait_state(const mpdpft_ctx *c, int s, double *BP, double *maxIter);
void nait_reset_state(mpdpft_ctx *c);
