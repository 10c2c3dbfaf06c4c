// solver.cpp -- the self-consistent DPFT loop on one MI355X.
//
// Host C++ keeps the reference's control flow (GetData, mpDPFT.cpp:135-244)
// and its small scalar decisions (mu bracketing/bisection, convergence,
// Pulay coefficients); every grid-sized operation runs as a gfx950 kernel on
// the context's stream.  Fields never leave HBM inside the loop: per step the
// host receives only the reduced scalars (particle numbers, dot products).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <limits>

#include "fft_core.h"
#include "solver.h"

static const double kPI = 3.141592653589793;
static const double kMP = 1.0e-12;  // mpDPFT.cpp:111

int ctx_fail(mpdpft_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}
int hip_fail(mpdpft_ctx *c, hipError_t e, const char *what) {
    return ctx_fail(c, MPDPFT_EHIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIPC(call)                                              \
    do {                                                        \
        hipError_t e_ = (call);                                 \
        if (e_ != hipSuccess) return hip_fail(c, e_, #call);    \
    } while (0)
#define RC(call)                     \
    do {                             \
        int r_ = (call);             \
        if (r_ != MPDPFT_OK) return r_; \
    } while (0)

static double hEXP(double x) {  // mpDPFT.h:266
    if (x < -700.) return 0.;
    if (x < 700.) return std::exp(x);
    return 1.0142320547350045e+304;
}
static double hSign(double x) { return x < 0. ? -1. : 1.; }

template <class T> static int dmalloc(mpdpft_ctx *c, T **p, size_t n) {
    void *q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) + 16);
    if (e != hipSuccess) return ctx_fail(c, MPDPFT_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocations.push_back(q);
    *p = (T *)q;
    return MPDPFT_OK;
}

void stage_begin(mpdpft_ctx *c) {
    if (c->timing) hipEventRecord(c->ev0, c->st);
}
void stage_end(mpdpft_ctx *c, const char *name, double units, double bytes, double flops) {
    if (!c->timing) return;
    hipEventRecord(c->ev1, c->st);
    hipEventSynchronize(c->ev1);
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0, c->ev1);
    for (auto &x : c->stats)
        if (std::strcmp(x.name, name) == 0) {
            x.launches++;
            x.total_ms += ms;
            x.units += units;
            x.bytes += bytes;
            x.flops += flops;
            return;
        }
    mpdpft_stage_stats x{};
    std::snprintf(x.name, sizeof x.name, "%s", name);
    x.launches = 1;
    x.total_ms = ms;
    x.units = units;
    x.bytes = bytes;
    x.flops = flops;
    c->stats.push_back(x);
}

void stage_add(mpdpft_ctx *c, const char *name, double units, double bytes, double flops) {
    if (!c->timing) return;
    for (auto &x : c->stats)
        if (std::strcmp(x.name, name) == 0) {
            x.units += units;
            x.bytes += bytes;
            x.flops += flops;
            return;
        }
}

extern "C" int mpdpft_enable_timing(mpdpft_ctx *c, int on) {
    if (!c || !c->st) return MPDPFT_EINVAL;
    if (on && !c->ev0) {
        HIPC(hipEventCreate(&c->ev0));
        HIPC(hipEventCreate(&c->ev1));
    }
    c->timing = on ? 1 : 0;
    c->stats.clear();
    return MPDPFT_OK;
}

extern "C" int mpdpft_get_stage_stats(mpdpft_ctx *c, mpdpft_stage_stats *out, int max, int *n) {
    if (!c || !n) return MPDPFT_EINVAL;
    int k = 0;
    for (auto &x : c->stats)
        if (out && k < max) out[k++] = x;
        else if (!out) k++;
    *n = out ? k : (int)c->stats.size();
    return MPDPFT_OK;
}

// Sum partial results and bring them to the host (blocks until ready).
int ctx_reduce(mpdpft_ctx *c, int nres, double scale, double *out_host) {
    HIPC(mk::finish_sums(c->d_partials, nres, c->d_scal, scale, c->st));
    HIPC(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * nres, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    for (int k = 0; k < nres; k++) out_host[k] = c->h_scal[k];
    return MPDPFT_OK;
}
static int ctx_reduce_min(mpdpft_ctx *c, int nres, double *out_host) {
    HIPC(mk::finish_minmax(c->d_partials, nres, c->d_scal, c->st));
    HIPC(hipMemcpyAsync(c->h_scal, c->d_scal, sizeof(double) * nres, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    for (int k = 0; k < nres; k++) out_host[k] = c->h_scal[k];
    return MPDPFT_OK;
}

// Trapezoid integral of a device field (Riemann, mpDPFT.cpp:7981-8026): the
// weighted sum times ((x_max - x_min)/(EL-1))^D applied factor by factor.
static double trap_scale(const mpdpft_ctx *c) { return (c->p.edge / 2. - (-c->p.edge / 2.)) / ((double)c->EL - 1.); }
static int integral_from_partials(mpdpft_ctx *c, double *out) {
    double v;
    RC(ctx_reduce(c, 1, 1., &v));
    double a = trap_scale(c);
    for (int d = 0; d < c->D; d++) v *= a;
    *out = v;
    return MPDPFT_OK;
}

// ---------------------------------------------------------------------------
// set-up: InitializeHardCodedParameters (grid part, :862-896, :1105-1106)
// ---------------------------------------------------------------------------
// Directory of the TabFunc_*.dat files the reference reads from its working
// directory (mpdpft_set_data_dir; "" = the current directory).
static std::string g_data_dir;
extern "C" int mpdpft_set_data_dir(const char *dir) {
    g_data_dir = dir ? dir : "";
    return MPDPFT_OK;
}

// TabFunc_2D3D_eta.dat (Load2D3Deta, mpDPFT.cpp:2927-2950): 31 x 31 triplets
// (eta1, eta2, eta3)(l1, l2) in row-major order, one per line "l1 l2 eta1 eta2 eta3".
static int load_eta_table(mpdpft_ctx *c) {
    const std::string path =
        (g_data_dir.empty() ? std::string() : g_data_dir + (g_data_dir.back() == '/' ? "" : "/")) +
        "TabFunc_2D3D_eta.dat";
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f) return ctx_fail(c, MPDPFT_EIO, "interaction 24 needs " + path);
    std::vector<double> eta;
    char line[512];
    while ((int)eta.size() < 3 * 31 * 31 && std::fgets(line, sizeof line, f)) {
        int a, b;
        double e0, e1, e2;
        if (std::sscanf(line, "%d %d %lf %lf %lf", &a, &b, &e0, &e1, &e2) != 5) continue;
        eta.push_back(e0);
        eta.push_back(e1);
        eta.push_back(e2);
    }
    std::fclose(f);
    if (eta.empty()) return ctx_fail(c, MPDPFT_EIO, "no eta triplets in " + path);
    c->eta_rows = (int)(eta.size() / 3);
    RC(dmalloc(c, &c->d_eta, eta.size()));
    HIPC(hipMemcpy(c->d_eta, eta.data(), sizeof(double) * eta.size(), hipMemcpyHostToDevice));
    return MPDPFT_OK;
}

static int validate(mpdpft_ctx *c, const mpdpft_params *p) {
    if (const char *why = params_structure_error(p)) return ctx_fail(c, MPDPFT_EINVAL, why);
    if (p->DIM != 2 && p->DIM != 3) return ctx_fail(c, MPDPFT_EINVAL, "DIM must be 2 or 3");
    static const int ok_steps[] = {16, 32, 64, 96, 128, 256};
    bool ok = false;
    for (int v : ok_steps) ok |= (p->steps == v);
    if (!ok) return ctx_fail(c, MPDPFT_EINVAL, "steps must be one of 16, 32, 64, 96, 128, 256");
    if (p->S < 1 || p->S > MPDPFT_MAX_SPECIES) return ctx_fail(c, MPDPFT_EINVAL, "S out of range");
    if (p->DensityExpression != 7 && p->DensityExpression != 5 && p->DensityExpression != 6 &&
        p->DensityExpression != 10)
        return ctx_fail(c, MPDPFT_EINVAL, "DensityExpression must be 7 (TF), 5 (nAiT), 6 or 10 (n3')");
    if (p->DensityExpression == 6 && p->Units != 1 && p->Units != 2)
        return ctx_fail(c, MPDPFT_EINVAL, "DensityExpression 6 needs Units 1 or 2");
    if (p->DensityExpression == 5 && p->DIM != 3) return ctx_fail(c, MPDPFT_EINVAL, "nAiT is built for DIM 3");
    if (p->method != 1) return ctx_fail(c, MPDPFT_EINVAL, "integration method must be 1 (Riemann)");
    if (p->TaskType < 0 || p->TaskType > 10 || p->TaskType == 4)  // the task loop is run_dir's
        return ctx_fail(c, MPDPFT_EINVAL, "TaskType must be 0, 1, 2, 3, 5, 6, 7, 8, 9 or 10");
    if (p->Symmetry != 0) return ctx_fail(c, MPDPFT_EINVAL, "Symmetry must be 0");
    if (!(p->Noise >= 0.)) return ctx_fail(c, MPDPFT_EINVAL, "Noise must be >= 0");
    if ((int)(p->Mixer[0] + 0.5) > 1) return ctx_fail(c, MPDPFT_EINVAL, "Mixer[0] must be 0 or 1");
    if ((int)p->Mixer[1] > 60) return ctx_fail(c, MPDPFT_EINVAL, "Pulay scope must be <= 60");
    for (int a = 0; a < p->n_interactions; a++) {
        int t = p->Interactions[a];
        if (!(t == 0 || t == 1 || t == 8 || t == 9 || t == 10 || t == 11 || t == 12 || t == 13 || t == 14 || t == 15 ||
              t == 17 || t == 24 || t == -1 || t == -12))
            return ctx_fail(c, MPDPFT_EINVAL, "unsupported interaction type " + std::to_string(t));
        if (t == 9 && !(p->DIM == 2 && p->S == 2))
            return ctx_fail(c, MPDPFT_EINVAL, "interaction 9 (renormalised contact) needs DIM 2 and two species");
        if (t == 14 && p->DIM != 2)
            return ctx_fail(c, MPDPFT_EINVAL, "interaction 14 (2D-3D crossover) needs DIM 2");
        if (t == 24 && !(p->DIM == 2 && p->S == 2))
            return ctx_fail(c, MPDPFT_EINVAL, "interaction 24 (quasi-2D contact) needs DIM 2 and two species");
        if (t == 15 && !(p->DIM == 3 && p->S == 1))
            return ctx_fail(c, MPDPFT_EINVAL, "interaction 15 (dipolar, momentum space) needs DIM 3 and one species");
        if (t == 17 && p->DIM != 3)
            return ctx_fail(c, MPDPFT_EINVAL, "interaction 17 (dipolar, position space) needs DIM 3");
        if ((t == 14 || t == 15 || t == 24) && p->alpha == 0.)
            return ctx_fail(c, MPDPFT_EINVAL, "interactions 14, 15, 24 need alpha != 0");
    }
    for (int r = 0; r < p->n_regularize; r++)
        if (p->regularize[r] != 0) return ctx_fail(c, MPDPFT_EINVAL, "regularize must be 0");
    for (int s = 0; s < p->S; s++) {
        int t = p->Environments[s];
        if (!(t == 0 || t == 1 || t == 3 || t == 40 || t == 41))
            return ctx_fail(c, MPDPFT_EINVAL, "unsupported environment type " + std::to_string(t));
        if ((t == 3 || t == 40) && p->n_mpp < 2 * p->S) return ctx_fail(c, MPDPFT_EINVAL, "mpp needs 2S values");
        if (t == 1 && p->n_mpp < p->S) return ctx_fail(c, MPDPFT_EINVAL, "mpp needs S values");
    }
    if (p->Precision != 0 && p->Precision != 1) return ctx_fail(c, MPDPFT_EINVAL, "Precision must be 0 or 1");
    if (p->Precision == 1) {
        if (p->DIM != 3) return ctx_fail(c, MPDPFT_EINVAL, "Precision 1: DIM 3");
        if (p->DensityExpression != 5 && p->DensityExpression != 7)
            return ctx_fail(c, MPDPFT_EINVAL, "Precision 1: DensityExpression 5 or 7");
        for (int a = 0; a < p->n_interactions; a++)
            if (p->Interactions[a] == 8 && p->HartreeMode != 1)
                return ctx_fail(c, MPDPFT_EINVAL, "Precision 1: isolated Hartree (HartreeMode 1)");
        if (p->DensityExpression == 5 && p->DelFieldMethod != 4 && p->DelFieldMethod != 5)
            return ctx_fail(c, MPDPFT_EINVAL, "Precision 1: DelFieldMethod 4 or 5");
    }
    if (p->InterpolVQ != 0 && p->InterpolVQ != 1)
        return ctx_fail(c, MPDPFT_EINVAL, "InterpolVQ must be 0 or 1 (2, 3 come from task loops)");
    for (int i = 0; i < p->n_schedule; i++) {
        int sc = p->Schedule[i];
        if (!(sc == 0 || sc == 3 || sc == 4 || sc == 5 || sc == 7 || sc == 9 || sc == 11))
            return ctx_fail(c, MPDPFT_EINVAL,
                            "Schedule " + std::to_string(sc) + " is not built (0, 3, 4, 5, 7, 9, 11 are)");
    }
    if (p->DensityExpression == 5)
        for (int s = 0; s < p->S; s++)
            if (p->TVec[s] == 0.) return ctx_fail(c, MPDPFT_EINVAL, "nAiT needs TVec != 0 (T < 0 picks it automatically)");
    if (p->DensityExpression == 5 && (p->n_Mpp < 3 || p->Mpp[0] < 1. || p->Mpp[1] > 60.))
        return ctx_fail(c, MPDPFT_EINVAL, "nAiT needs Mpp = (BoolePairs >= 1, maxIter <= 60, a_r threshold)");
    return MPDPFT_OK;
}

static void twiddles(int n, std::vector<double2> &w) {
    w.resize(n);
    for (int t = 0; t < n; t++) {
        double a = 2. * kPI * (double)t / (double)n;
        w[t].x = std::cos(a);
        w[t].y = -std::sin(a);
    }
}

static int taper_geometry(mpdpft_ctx *c);
static void reset_loop_scalars(mpdpft_ctx *c);

extern "C" int mpdpft_create(const mpdpft_params *p, int device, mpdpft_ctx **out) {
    if (!p || !out) return MPDPFT_EINVAL;
    *out = nullptr;
    mpdpft_ctx *c = new mpdpft_ctx();
    int rc = validate(c, p);
    if (rc) {
        *out = c;
        return rc;
    }
    *out = c;
    c->p = *p;
    if (c->p.Print < 0) c->p.Print = 1;  // promoted on the last task, :716
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return ctx_fail(c, MPDPFT_ENODEV, "no HIP device");
    if (device < 0) HIPC(hipGetDevice(&device));
    c->device = device;
    HIPC(hipSetDevice(device));
    HIPC(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));

    c->D = p->DIM;
    c->s = p->steps;
    c->EL = p->steps + 1;
    c->S = p->S;
    c->G = 1;
    c->F = 1;
    for (int d = 0; d < c->D; d++) {
        c->G *= c->EL;
        c->F *= c->s;
    }
    c->SG = c->G * c->S;
    c->CentreIndex = (c->G - 1) / 2;
    double x1 = -p->edge / 2., x2 = p->edge / 2.;
    c->dx = p->edge / (double)p->steps;
    c->k0 = -kPI / c->dx;
    c->dk = 2. * kPI / p->edge;
    c->kdel = (-c->k0 - c->k0) / (double)p->steps;
    c->sigma = p->SmoothingKWidth > 0. ? -0.5 / (p->SmoothingKWidth * p->SmoothingKWidth)
                                       : -0.5 * std::pow(p->SmearingExtentDx * c->dx, 2.);
    c->lat_h.resize(c->EL);
    double del = (x2 - x1) / (double)p->steps;
    for (int i = 0; i <= p->steps; i++) c->lat_h[i] = x1 + del * (double)i;
    // k per standard DFT index q: centred index p = (q + s/2) mod s
    std::vector<double> kq(c->s), kqh(c->s);
    for (int q = 0; q < c->s; q++) {
        int pidx = (q + c->s / 2) % c->s;
        kq[q] = c->k0 + c->kdel * (double)pidx;
        kqh[q] = c->k0 + (double)pidx * c->dk;
    }
    std::vector<double2> tw, tw2;
    twiddles(c->s, tw);

    size_t SG = c->SG, G = c->G;
    RC(dmalloc(c, &c->d_lat, c->EL));
    RC(dmalloc(c, &c->d_kq, c->s));
    RC(dmalloc(c, &c->d_ek, c->s));
    RC(dmalloc(c, &c->d_kqh, c->s));
    RC(dmalloc(c, &c->d_tw, c->s));
    RC(dmalloc(c, &c->d_spec, c->F));
    RC(dmalloc(c, &c->d_work, c->F));
    RC(dmalloc(c, &c->d_Env, SG));
    RC(dmalloc(c, &c->d_V, SG));
    RC(dmalloc(c, &c->d_OldV, SG));
    RC(dmalloc(c, &c->d_Den, SG));
    RC(dmalloc(c, &c->d_OldDen, SG));
    RC(dmalloc(c, &c->d_Dx, SG));
    RC(dmalloc(c, &c->d_Dy, SG));
    RC(dmalloc(c, &c->d_Dz, SG));
    RC(dmalloc(c, &c->d_Grad2, SG));
    RC(dmalloc(c, &c->d_Lap, SG));
    RC(dmalloc(c, &c->d_tmp1, G));
    RC(dmalloc(c, &c->d_tmp2, G));
    RC(dmalloc(c, &c->d_tmp3, G));
    RC(dmalloc(c, &c->d_partials, (size_t)mk::kRedBlocks * 64));
    RC(dmalloc(c, &c->d_scal, 256));
    RC(dmalloc(c, &c->d_coeff, 64));
    RC(dmalloc(c, &c->d_theta, MPDPFT_MAX_SPECIES));
    RC(dmalloc(c, &c->d_mus, 64));
    RC(dmalloc(c, &c->d_centres, (size_t)MPDPFT_MAX_CENTRES * 5));
    HIPC(hipHostMalloc((void **)&c->h_scal, 256 * sizeof(double), hipHostMallocDefault));
    HIPC(hipMemcpy(c->d_lat, c->lat_h.data(), sizeof(double) * c->EL, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_kq, kq.data(), sizeof(double) * c->s, hipMemcpyHostToDevice));
    {  // smoothing factor per axis index: EXP(sigma k^2) with the reference's EXP cut-offs (mpDPFT.h:266)
        std::vector<double> ek(c->s);
        for (int q = 0; q < c->s; q++) {
            double x = c->sigma * kq[q] * kq[q];
            ek[q] = x < -700. ? 0. : (x < 700. ? std::exp(x) : 1.0142320547350045e+304);
        }
        HIPC(hipMemcpy(c->d_ek, ek.data(), sizeof(double) * c->s, hipMemcpyHostToDevice));
        std::vector<double2> kek(c->s);
        for (int q = 0; q < c->s; q++) kek[q] = double2{kq[q], ek[q]};
        RC(dmalloc(c, &c->d_kek, (size_t)c->s));
        HIPC(hipMemcpy(c->d_kek, kek.data(), sizeof(double2) * c->s, hipMemcpyHostToDevice));
    }
    HIPC(hipMemcpy(c->d_kqh, kqh.data(), sizeof(double) * c->s, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(c->d_tw, tw.data(), sizeof(double2) * c->s, hipMemcpyHostToDevice));
    if (p->n_centres > 0)
        HIPC(hipMemcpy(c->d_centres, p->centres, sizeof(double) * 5 * p->n_centres, hipMemcpyHostToDevice));

    c->g = mk::Grid{c->D, c->s, c->EL, c->G, c->F, p->edge, c->dx, c->d_lat};

    for (int a = 0; a < p->n_interactions; a++) {
        const int t = p->Interactions[a];
        if (t == 8) c->has_hartree = 1;
        if (t == 15 && !c->d_dip_k) RC(dmalloc(c, &c->d_dip_k, SG));
        if (t == 17 && !c->d_dip_r) RC(dmalloc(c, &c->d_dip_r, SG));
        if (t == 24 && !c->d_eta) RC(load_eta_table(c));
    }
    if (c->has_hartree && p->HartreeMode == 1) {
        int M = 2 * c->s;
        size_t tot = (size_t)M * M * (c->D == 3 ? M : 1);
        twiddles(M, tw2);
        RC(dmalloc(c, &c->d_tw2, M));
        RC(dmalloc(c, &c->d_pad, tot));
        RC(dmalloc(c, &c->d_khat, tot));
        HIPC(hipMemcpy(c->d_tw2, tw2.data(), sizeof(double2) * M, hipMemcpyHostToDevice));
        // cell average of 1/r at the origin: cube (3D) or square (2D) of side dx
        double g0 = (c->D == 3 ? 3. * std::log((std::sqrt(3.) + 1.) / (std::sqrt(3.) - 1.)) - kPI / 2.
                               : 4. * std::log(1. + std::sqrt(2.))) /
                    c->dx;
        HIPC(mk::build_isolated_kernel(c->g, c->d_pad, c->d_tw2, g0, c->d_khat, c->st));
    }
    if (p->Precision == 1) {  // FP32 spectra, twiddles, multiplier and kernel tables
        RC(dmalloc(c, &c->d_specf, c->F));
        RC(dmalloc(c, &c->d_workf, c->F));
        RC(dmalloc(c, &c->d_twf, (size_t)c->s));
        RC(dmalloc(c, &c->d_kekf, (size_t)c->s));
        HIPC(mk::to_float2((size_t)c->s, c->d_tw, c->d_twf, c->st));
        std::vector<double2> kek_h(c->s);
        HIPC(hipMemcpy(kek_h.data(), c->d_kek, sizeof(double2) * c->s, hipMemcpyDeviceToHost));
        std::vector<float2> kekf(c->s);
        for (int q = 0; q < c->s; q++)
            kekf[q] = float2{(float)kek_h[q].x, p->DelFieldMethod == 5 ? (float)kek_h[q].y : 1.f};
        HIPC(hipMemcpy(c->d_kekf, kekf.data(), sizeof(float2) * c->s, hipMemcpyHostToDevice));
        if (c->has_hartree) {
            int M = 2 * c->s;
            size_t tot = (size_t)M * M * M;
            RC(dmalloc(c, &c->d_padf, tot));
            RC(dmalloc(c, &c->d_khatf, tot));
            RC(dmalloc(c, &c->d_tw2f, (size_t)M));
            HIPC(mk::to_float2((size_t)M, c->d_tw2, c->d_tw2f, c->st));
            HIPC(mk::to_float(tot, c->d_khat, c->d_khatf, c->st));
        }
    }

    // environments and V <- Env (InitializeEnvironments, :6362-6390)
    for (int s = 0; s < c->S; s++) {
        mk::EnvParams e{p->Environments[s], p->n_mpp > s ? p->mpp[s] : 0.,
                        p->n_mpp > p->S + s ? p->mpp[p->S + s] : 0., p->n_centres, c->d_centres};
        HIPC(mk::environment(c->g, e, c->d_Env + s * G, c->st));
    }
    HIPC(hipMemcpyAsync(c->d_V, c->d_Env, sizeof(double) * SG, hipMemcpyDeviceToDevice, c->st));
    if (p->TaperOff && (p->DelFieldMethod == 4 || p->DelFieldMethod == 5) &&
        (p->DensityExpression == 5 || p->DIM == 3))  // DelField will run (nAiT, or the 3D energies)
        RC(taper_geometry(c));
    {  // FFT block shapes: time the candidates once per length and precision (MPDPFT_FFT_TUNE=0: defaults;
       // MPDPFT_FFT_B=0|1|2: half, default or twice the default lines per block for every length)
        const char *tune = std::getenv("MPDPFT_FFT_TUNE");
        const char *force = std::getenv("MPDPFT_FFT_B");
        if (force && force[0] >= '0' && force[0] <= '2' && force[1] == 0) {
            fftk::bsel_force(force[0] - '0');
        } else if (!(tune && tune[0] == '0')) {
            const bool iso = c->has_hartree && p->HartreeMode == 1;
            HIPC(mk::fft_autotune(c->g, c->d_Env, c->d_spec, c->d_tw, iso ? c->d_pad : nullptr, c->d_khat, c->d_tw2,
                                  c->d_tmp1, c->st));
            if (p->Precision == 1)
                HIPC(mk::fft32_autotune(c->g, c->d_Env, c->d_specf, c->d_twf, iso ? c->d_padf : nullptr, c->d_khatf,
                                        c->d_tw2f, c->d_tmp1, c->st));
        }
    }
    HIPC(hipStreamSynchronize(c->st));
    for (int s = 0; s < c->S; s++) {  // scalar state used by the stage entry points
        c->muVec[s] = p->muVec[s];
        c->thetaVec[s] = p->thetaVec[s];
        c->TVec[s] = p->TVec[s];
    }
    c->MaxIteration = p->MaxIteration;
    const char *vb = getenv("MPDPFT_VERBOSE");
    c->verbose = vb && *vb && *vb != '0';
    const char *nk = std::getenv("MPDPFT_NAIT_KERNEL");
    c->nait_plain = nk && std::strcmp(nk, "plain") == 0;
    c->nait_generic = nk && std::strcmp(nk, "reuse") == 0;
    if (p->DensityExpression == 5) RC(nait_init(c));
    reset_loop_scalars(c);  // the stage entry points may run before any loop
    return MPDPFT_OK;
}

extern "C" void mpdpft_destroy(mpdpft_ctx *c) {
    if (!c) return;
    if (c->st) hipStreamSynchronize(c->st);
    nait_free(c);
    adjoint_free(c);
    for (void *q : c->allocations) hipFree(q);
    for (double *q : c->nin) hipFree(q);
    for (double *q : c->res) hipFree(q);
    for (double2 *q : c->resF) hipFree(q);
    if (c->h_spec0) hipHostFree(c->h_spec0);
    if (c->h_scal) hipHostFree(c->h_scal);
    if (c->ev0) hipEventDestroy(c->ev0);
    if (c->ev1) hipEventDestroy(c->ev1);
    if (c->st) hipStreamDestroy(c->st);
    delete c;
}

extern "C" const char *mpdpft_last_error(const mpdpft_ctx *c) { return c ? c->err.c_str() : "null context"; }
extern "C" size_t mpdpft_grid_size(const mpdpft_ctx *c) { return c ? c->G : 0; }

// ---------------------------------------------------------------------------
// stages
// ---------------------------------------------------------------------------

// GetDensity + Integrate for species s at chemical potential mu, density into
// den_out (device, [G]); returns the particle number.
static int density(mpdpft_ctx *c, int s, double mu, double *den_out, double *N) {
    const mpdpft_params &p = c->p;
    if (p.DensityExpression == 7) {  // GetTFdensity (:4415-4456)
        double g = p.degeneracy, pref;
        if (c->D == 2) pref = g / (2. * kPI);
        else pref = (p.Units == 2 ? 0.0475412512 : 1.) * std::sqrt(2.) * g / (3. * kPI * kPI);
        stage_begin(c);
        HIPC(mk::tf_density(c->g, c->d_V + s * c->G, mu, pref, den_out, c->d_partials, c->st));
        stage_end(c, "tf_density", (double)c->G, 16. * c->G, 0.);
        return integral_from_partials(c, N);
    }
    if (p.DensityExpression == 6) {  // Getn3pFFT (:3625-3684): O(G F) momentum sum, backward transform
        const double DD = (double)c->D, Uf = p.Units == 2 ? 0.131163 : 1.;
        const double pref = std::pow(c->dx, DD) * p.degeneracy * (c->D == 2 ? 2. * kPI : 4. * kPI) *
                            std::pow(8., DD / 2.) / (std::pow(2. * kPI, DD) * DD * std::pow(2., DD));
        stage_begin(c);
        HIPC(mk::n3p_spec(c->g, c->d_V + s * c->G, mu, Uf, pref, c->d_spec, c->st));
        mk::MulParams m{};
        m.kind = mk::MUL_NONE;
        m.kq = c->d_kq;
        HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_spec, m, den_out, nullptr,
                                   std::pow(1. / p.edge, DD), 0, 0., c->d_tw, c->st));
        HIPC(mk::trap_sum(c->g, den_out, c->d_partials, c->st));
        stage_end(c, "n3p_fft", (double)c->F * (double)c->G, 8. * (double)c->F * (double)c->G, 0.);
        return integral_from_partials(c, N);
    }
    if (p.DensityExpression == 10) {  // Getn3p (:3702-3729): O(G^2) Bessel-kernel sum
        const double mass = p.Units == 2 ? 0.131163 : 1.;
        double scale = p.degeneracy, a = trap_scale(c);
        for (int d = 0; d < c->D; d++) scale *= a;
        stage_begin(c);
        HIPC(mk::n3p_bessel(c->g, c->d_V + s * c->G, mu, mass, scale, den_out, c->st));
        HIPC(mk::trap_sum(c->g, den_out, c->d_partials, c->st));
        stage_end(c, "n3p_bessel", (double)c->G * (double)c->G, 8. * (double)c->G * (double)c->G, 0.);
        return integral_from_partials(c, N);
    }
    return nait_density(c, s, mu, den_out, N);
}

// EnforceConstraints (mpDPFT.cpp:4569-4743).
static int enforce_constraints(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    int S = c->S;
    c->TrappedQ = 0;
    bool abortQ = false;
    if (c->RelAcc > kMP && p.TaskType < 100) {
        double nrm = 0.;
        for (int s = 0; s < S; s++) {
            c->PreviousmuVec[s] = c->muVec[s];
            nrm += c->muVec[s] * c->muVec[s];
        }
        if (std::sqrt(nrm) < kMP)
            for (int s = 0; s < S; s++) c->PreviousmuVec[s] = hSign(c->PreviousmuVec[s]) * kMP;
        double minDelta = std::fabs(c->muVec[0]);
        double lo[MPDPFT_MAX_SPECIES], hi[MPDPFT_MAX_SPECIES], prev[MPDPFT_MAX_SPECIES], step[MPDPFT_MAX_SPECIES];
        bool loFound[MPDPFT_MAX_SPECIES], hiFound[MPDPFT_MAX_SPECIES];
        for (int s = 0; s < S; s++) {
            lo[s] = hi[s] = c->muVec[s];
            prev[s] = 0.;
            step[s] = c->muVec[s];
            loFound[s] = hiFound[s] = false;
        }
        // bracketing: walk mu down while N is too large, else up
        for (int s = 0; s < S && !abortQ; s++) {
            double target = p.Abundances[s], d0 = std::fabs(c->TmpAb[s] - target);
            if (!(d0 / std::fabs(2. * target) > c->RelAcc)) continue;
            int n = 0;
            while (c->TmpAb[s] > target && n < 20) {
                hiFound[s] = true;
                hi[s] = std::min(hi[s], c->muVec[s]);
                step[s] = hEXP(c->DeltamuModifier) * (1. + (double)n) * std::fabs(c->PreviousmuVec[s]);
                prev[s] = c->muVec[s];
                c->muVec[s] -= step[s];
                if (std::fabs(c->muVec[s]) > 1.0e+20) {
                    c->warningCount++;
                    c->muVec[s] = c->PreviousmuVec[s];
                    break;
                }
                RC(density(c, s, c->muVec[s], c->d_Den + s * c->G, &c->TmpAb[s]));
                c->bisection_steps++;
                n++;
                if (n == 1)
                    c->DeltamuModifier -= hSign(std::fabs(c->TmpAb[s] - target) - d0) * 0.1 * std::fabs(c->DeltamuModifier);
                lo[s] = c->muVec[s];
                if (c->RelAcc < kMP) c->RelAcc = c->RelAccOriginal;
            }
            if (c->TmpAb[s] <= target) loFound[s] = true;
            n = 0;
            if (!hiFound[s]) {
                while (c->TmpAb[s] < target && n < 20) {
                    loFound[s] = true;
                    if (std::fabs(c->muVec[s]) > 1.0e+100) {
                        c->errorCount++;
                        c->muVec[s] = c->PreviousmuVec[s];
                        break;
                    }
                    lo[s] = std::max(lo[s], c->muVec[s]);
                    step[s] = hEXP(c->DeltamuModifier) * (1. + (double)n) * std::fabs(c->PreviousmuVec[s]);
                    prev[s] = c->muVec[s];
                    c->muVec[s] += step[s];
                    if (std::fabs(c->muVec[s]) > 1.0e+20) {
                        c->warningCount++;
                        c->muVec[s] = c->PreviousmuVec[s];
                        break;
                    }
                    RC(density(c, s, c->muVec[s], c->d_Den + s * c->G, &c->TmpAb[s]));
                    c->bisection_steps++;
                    n++;
                    if (n == 1)
                        c->DeltamuModifier -=
                            hSign(std::fabs(c->TmpAb[s] - target) - d0) * 0.1 * std::fabs(c->DeltamuModifier);
                    hi[s] = c->muVec[s];
                    if (c->RelAcc < kMP) c->RelAcc = c->RelAccOriginal;
                }
            }
            if (c->TmpAb[s] >= target) hiFound[s] = true;
            if (!hiFound[s] || !loFound[s]) {  // :4657-4661 (printed, not counted as a warning)
                c->muVec[s] = c->PreviousmuVec[s];
                abortQ = true;
                c->bracket_failures++;
                c->control << "mu-ranges not found for s=" << s << " --- reset muVec to PreviousmuVec\n";
            }
        }
        if (!abortQ) {
            for (int s = 0; s < S; s++) {
                step[s] = (hi[s] - lo[s]) * 0.5;
                c->muVec[s] = (lo[s] + hi[s]) * 0.5;
            }
            for (int s = 0; s < S; s++) {
                double target = p.Abundances[s];
                if (!(std::fabs(c->TmpAb[s] - target) / std::fabs(2. * target) > c->RelAccOriginal)) continue;
                std::vector<double> hist;
                bool success = false, tooSmall = false;
                while (!success && !tooSmall) {
                    success = true;
                    RC(density(c, s, c->muVec[s], c->d_Den + s * c->G, &c->TmpAb[s]));
                    c->bisection_steps++;
                    if (c->TmpAb[s] > 0.) hist.push_back(c->TmpAb[s]);
                    bool trapped = false;
                    int nh = (int)hist.size();
                    if (nh > 5)
                        for (int h = nh - 5; h < nh - 1 && !trapped; h++)
                            for (int hp = h + 1; hp < nh - 1; hp++)
                                if (std::fabs(hist[h] - hist[hp]) < c->RelAccOriginal * target) {
                                    trapped = true;
                                    c->TrappedQ = 1;
                                    break;
                                }
                    if (trapped) break;
                    for (int sg = -1; sg < 2; sg += 2) {
                        if ((double)sg * (c->TmpAb[s] - target) / std::fabs(2. * target) > c->RelAccOriginal) {
                            prev[s] = c->muVec[s];
                            c->muVec[s] -= (double)sg * 0.5 * step[s];
                            if (std::fabs(c->muVec[s]) > 1.0e+20) {
                                c->warningCount++;
                                c->muVec[s] = c->PreviousmuVec[s];
                                break;
                            }
                            step[s] = std::fabs(c->muVec[s] - prev[s]);
                            minDelta = std::min(minDelta, step[s]);
                            if (std::fabs(step[s]) < kMP) {
                                c->warningCount++;
                                tooSmall = true;
                                break;
                            }
                            success = false;
                        }
                    }
                }
            }
            c->RelAcc = c->RelAccOriginal;
        }
    }
    for (int s = 0; s < S; s++)
        if (!std::isfinite(c->muVec[s])) c->muVec[s] = c->PreviousmuVec[s];
    return MPDPFT_OK;
}

static int get_densities(mpdpft_ctx *c) {  // GetDensities (:3306-3316)
    for (int s = 0; s < c->S; s++) RC(density(c, s, c->muVec[s], c->d_Den + s * c->G, &c->TmpAb[s]));
    return enforce_constraints(c);
}

static int field_stats(mpdpft_ctx *c) {  // GetFieldStats (:12096), min/max only
    double r[4 * MPDPFT_MAX_SPECIES];
    HIPC(mk::field_minmax_S(c->g, c->S, c->d_Den, c->d_V, c->d_partials, c->st));
    RC(ctx_reduce_min(c, 4 * c->S, r));
    for (int s = 0; s < c->S; s++) {
        c->DenMin[s] = r[4 * s];
        c->DenMax[s] = -r[4 * s + 1];
        c->VMin[s] = r[4 * s + 2];
        c->VMax[s] = -r[4 * s + 3];
    }
    return MPDPFT_OK;
}

// Hartree potential of density n (device [G]) into out (device [G]).
// Algorithmic HBM bytes of the pruned zero-padded isolated solve (FP64): every
// pass reads and writes only the lines that can hold nonzeros / reach [0,s]^D.
static double iso_bytes(const mpdpft_ctx *c) {
    double s = c->s, M = 2. * s, e = s + 1.;
    if (c->D == 3)
        return s * s * (8. * s + 16. * M) + s * M * (16. * s + 16. * M) + M * M * (16. * s + 16. * M + 8. * M) +
               32. * M * M * M + 32. * M * M * e + e * e * (16. * M + 8. * e);
    return s * (8. * s + 16. * M) + M * (16. * s + 16. * M + 8. * M) + 32. * M * M + e * (16. * M + 8. * e);
}

static int hartree(mpdpft_ctx *c, const double *n, double *out) {
    const mpdpft_params &p = c->p;
    double dxD = std::pow(c->dx, (double)c->D);
    if (p.HartreeMode == 1 && c->fp32_phase) {
        double M = 2. * c->s, tot = std::pow(M, (double)c->D);
        stage_begin(c);
        HIPC(mk::hartree_isolated32(c->g, n, c->d_padf, c->d_khatf, c->d_tw2f, (float)(p.gammaH * dxD / tot), out,
                                    c->st));
        stage_end(c, "hartree_isolated_fp32", 1., 0.5 * iso_bytes(c) + 4. * std::pow(M, 3.), 0.);
        return MPDPFT_OK;
    }
    if (p.HartreeMode == 1) {
        double M = 2. * c->s, tot = std::pow(M, (double)c->D);
        stage_begin(c);
        HIPC(mk::hartree_isolated(c->g, n, c->d_pad, c->d_khat, c->d_tw2, p.gammaH * dxD / tot, out, 0, c->st));
        stage_end(c, "hartree_isolated", 1., iso_bytes(c), 0.);
        return MPDPFT_OK;
    }
    stage_begin(c);
    if (c->D == 2) {  // periodic + Gaussian compensation, 2D (HartreeRepulsion :2627-2659)
        HIPC(mk::fft_forward_real(c->g, n, nullptr, c->d_spec, c->d_tw, c->d_scal + 200, c->st));
        mk::MulParams m{};
        m.kind = mk::MUL_HARTREE_2D;
        m.kq = c->d_kq;
        m.kq_h = c->d_kqh;
        m.f0slot = c->d_scal + 200;
        m.dxD = dxD;
        m.dkzero = 0.00001 * c->dk;
        HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_spec, m, out, nullptr, std::pow(1. / p.edge, 2.), 3,
                                   2. * kPI * p.gammaH, c->d_tw, c->st));
        stage_end(c, "hartree_periodic", 2., 2. * 32. * c->D * (double)c->F, 0.);
        return MPDPFT_OK;
    }
    // periodic + Gaussian compensation (HartreeRepulsion 3D, :2624-2712)
    HIPC(mk::fft_forward_real(c->g, n, nullptr, c->d_spec, c->d_tw, c->d_scal + 200, c->st));
    mk::MulParams m{};
    m.kind = mk::MUL_HARTREE_PER;
    m.kq = c->d_kq;
    m.kq_h = c->d_kqh;
    m.f0slot = c->d_scal + 200;
    m.dxD = dxD;
    m.dkzero = 0.000001 * c->dk * c->dk;
    m.regularize = p.RegularizeCoulombSingularity;
    double LD = std::pow(1. / p.edge, (double)c->D);
    HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_spec, m, out, nullptr, LD, 2, 4. * kPI * p.gammaH, c->d_tw,
                               c->st));
    stage_end(c, "hartree_periodic", 2., 2. * 32. * c->D * (double)c->F, 0.);
    return MPDPFT_OK;
}

// UpdateV + GetVint (:4815-4850): V_s = inc * sum_a Vint_a + Env_s from the
// mixed density (held in OldDen after the pointer swap of StoreCurrentIteration).
// The dipolar potentials of interactions 17 (DipolarInteractionPos, one
// transform pair per species) and 15 (DipolarInteraction, the pair sum), each
// into its [S][G] field for k_update_v to add in list order.
static int dipolar_fields(mpdpft_ctx *c, const double *den) {
    const mpdpft_params &p = c->p;
    const size_t G = c->G;
    if (c->d_dip_r) {
        for (int s = 0; s < c->S; s++) {
            stage_begin(c);
            HIPC(mk::fft_forward_real(c->g, den + s * G, nullptr, c->d_spec, c->d_tw, c->d_scal + 200, c->st));
            mk::MulParams m{};
            m.kind = mk::MUL_DIPOLAR;
            m.kq = c->d_kq;
            m.kq_h = c->d_kqh;
            m.dxD = std::pow(c->dx, 3.);
            m.dkzero = 0.01 * c->dk * c->dk;
            m.n0 = p.Abundances[s];
            HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_spec, m, c->d_dip_r + s * G, nullptr,
                                       std::pow(1. / p.edge, 3.), 0, 0., c->d_tw, c->st));
            HIPC(mk::dipolar_comp(c->g, p.Abundances[s], p.alpha, c->d_dip_r + s * G, c->st));
            stage_end(c, "dipolar_pos", 2., 2. * 32. * 3. * (double)c->F + 16. * (double)G, 0.);
        }
    }
    if (c->d_dip_k) {
        if (p.Units == 1) {
            double scale = p.alpha / (2. * std::pow(2. * kPI, 3.));
            const double a = (p.edge / 2. - (-p.edge / 2.)) / ((double)c->EL - 1.);
            for (int d = 0; d < 3; d++) scale *= a;
            stage_begin(c);
            HIPC(mk::dipolar_k(c->g, den, scale, c->d_dip_k, c->st));
            stage_end(c, "dipolar_k", 1., 8. * (double)G, 0.);
        } else {
            HIPC(mk::fill(G, 0., c->d_dip_k, c->st));
        }
    }
    return MPDPFT_OK;
}

static void vint_extra(const mpdpft_ctx *c, mk::VintParams &vp) {
    vp.alpha = c->p.alpha;
    vp.eta = c->d_eta;
    vp.eta_rows = c->eta_rows;
    vp.dip_r = c->d_dip_r;
    vp.dip_k = c->d_dip_k;
}

static int update_v(mpdpft_ctx *c, const double *den) {
    const mpdpft_params &p = c->p;
    if (c->has_hartree)
        for (int s = 0; s < c->S; s++) RC(hartree(c, den + s * c->G, c->d_V + s * c->G));
    RC(dipolar_fields(c, den));
    mk::VintParams vp{};
    vint_extra(c, vp);
    vp.n_inter = p.n_interactions;
    for (int a = 0; a < p.n_interactions; a++) vp.inter[a] = p.Interactions[a];
    vp.S = c->S;
    vp.s = c->s;
    vp.beta = p.beta;
    vp.gammaH = p.gammaH;
    vp.inc = c->incV;
    vp.units = p.Units;
    vp.hartree_in_V = c->has_hartree;
    vp.fp32 = c->fp32_phase;
    stage_begin(c);
    HIPC(mk::update_v(c->g, vp, den, c->d_Env, c->d_V, c->st));
    stage_end(c, "update_v", (double)c->SG, 8. * c->SG * (2. + c->S + c->has_hartree), 0.);
    if (c->SCcount > 0) c->incV += c->incV0;
    c->incV = std::min(1., c->incV);
    c->delfield_valid = -1;
    return MPDPFT_OK;
}

// DelField methods 4/5 (:7243-7582) as 5 fused transforms (3D):
// fwd(taper(V)); bwd(i kx e, i ky e) -> Dx,Dy; bwd(i kz e, -k^2 e^2) -> Dz,Lap;
// fwd(sum D^2); bwd(e) -> POS -> Grad2.  The taper and the sum of squares are
// formed in the first pass of their forward transforms.
// The taper's geometry (DelField TaperOff, :7251-7264), built once per context:
// the per-point factor table and the list of the shell's points.
static int taper_geometry(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    const size_t G = c->G;
    const double rt = p.edge * p.RegularizationThreshold;
    RC(dmalloc(c, &c->d_taperT, G));
    HIPC(mk::taper_factor(c->g, rt, c->d_taperT, c->st));
    int *flag = nullptr;
    HIPC(hipMalloc((void **)&flag, sizeof(int) * G));
    HIPC(mk::shell_mark(c->g, rt, flag, c->st));
    std::vector<int> hf(G), idx;
    hipError_t e = hipMemcpyAsync(hf.data(), flag, sizeof(int) * G, hipMemcpyDeviceToHost, c->st);
    if (e == hipSuccess) e = hipStreamSynchronize(c->st);
    (void)hipFree(flag);
    if (e != hipSuccess) return hip_fail(c, e, "shell_mark");
    for (size_t i = 0; i < G; i++)
        if (hf[i]) idx.push_back((int)i);
    c->n_shell = (int)idx.size();
    int *di = nullptr;
    RC(dmalloc(c, &di, std::max<size_t>(1, idx.size())));
    c->d_shell_idx = di;
    if (!idx.empty()) HIPC(hipMemcpy(c->d_shell_idx, idx.data(), sizeof(int) * idx.size(), hipMemcpyHostToDevice));
    RC(dmalloc(c, &c->d_ticket, 1));
    HIPC(hipMemset(c->d_ticket, 0, sizeof(unsigned)));
    return MPDPFT_OK;
}

static int delfield_impl(mpdpft_ctx *c, const double *f_dev, int s, int method);
int ctx_delfield(mpdpft_ctx *c, const double *f_dev, int s, int method) {
    stage_begin(c);
    int rc = delfield_impl(c, f_dev, s, method);
    if (rc) return rc;
    // HBM bytes the kernels move: a forward transform from k real grid fields
    // reads 8kF and writes 16F in its first pass, then 32F per pass; a backward
    // one to k real grids moves 32F per pass and 16F + 8kG in its last; the
    // taper reads its factor table (8G) and the shell maximum 12 bytes per shell
    // point; method 4 forms the square sum in its own kernel (8(D+1)G), method 5
    // in the first pass of the forward transform of it
    const double F = (double)c->F, Gd = (double)c->G, D = (double)c->D;
    const double cb = c->fp32_phase ? 8. : 16.;  // bytes of a complex spectrum value
    auto fwd = [&](double k) { return 8. * k * F + cb * F + 2. * cb * F * (D - 1.); };
    auto bwd = [&](double k) { return 2. * cb * F * (D - 1.) + cb * F + 8. * k * Gd; };
    double bytes = fwd(1.) + bwd(2.) + bwd(c->D == 3 ? 2. : 1.);
    if (method == 5) bytes += fwd(D) + bwd(1.);
    else bytes += 8. * (D + 1.) * Gd;
    if (c->p.TaperOff && (method == 4 || method == 5)) bytes += 8. * Gd + 12. * std::max(0, c->n_shell);
    stage_end(c, c->fp32_phase ? "delfield_fp32" : "delfield", method == 5 ? 5. : 3., bytes, 0.);
    return MPDPFT_OK;
}
static int delfield_impl(mpdpft_ctx *c, const double *f_dev, int s, int method) {
    const mpdpft_params &p = c->p;
    c->delfield_valid = -1;
    size_t G = c->G;
    if (p.TaperOff && (method == 4 || method == 5)) {  // shift by the shell maximum, taper (fused)
        if (c->n_shell < 0) RC(taper_geometry(c));
        HIPC(mk::shell_max_list(f_dev, c->d_shell_idx, c->n_shell, c->d_partials, c->d_ticket, c->d_scal + 210,
                                c->st));
    }
    const bool taper = p.TaperOff && (method == 4 || method == 5);
    if (c->fp32_phase) {  // the same transforms in FP32 (Precision 1)
        const float sc = (float)(1. / (double)c->F);
        double *Dx = c->d_Dx + s * G, *Dy = c->d_Dy + s * G, *Dz = c->d_Dz + s * G;
        double *g2 = c->d_Grad2 + s * G, *lap = c->d_Lap + s * G;
        HIPC(mk::fft32_forward(c->g, f_dev, taper ? c->d_taperT : nullptr, c->d_scal + 210, c->d_specf, c->d_twf,
                               c->st));
        HIPC(mk::fft32_backward(c->g, c->d_specf, c->d_workf, mk::MUL_DERIV_PAIR, 0, 1, c->d_kekf, Dx, Dy, sc, 0,
                                c->d_twf, c->st));
        HIPC(mk::fft32_backward(c->g, c->d_specf, c->d_workf, mk::MUL_DERIV_LAP, 2, -1, c->d_kekf, Dz, lap, sc, 0,
                                c->d_twf, c->st));
        if (method == 5) {
            HIPC(mk::fft32_forward_sumsq(c->g, Dx, Dy, Dz, c->d_specf, c->d_twf, c->st));
            HIPC(mk::fft32_backward(c->g, c->d_specf, c->d_specf, mk::MUL_SMOOTH, -1, -1, c->d_kekf, g2, nullptr, sc,
                                    1, c->d_twf, c->st));
        } else {
            HIPC(mk::sum_squares3(c->g, Dx, Dy, Dz, g2, c->st));
        }
        return MPDPFT_OK;
    }
    if (taper)
        HIPC(mk::fft_forward_taper(c->g, f_dev, c->d_taperT, c->d_scal + 210, c->d_spec, c->d_tw, c->st));
    else
        HIPC(mk::fft_forward_real(c->g, f_dev, nullptr, c->d_spec, c->d_tw, nullptr, c->st));
    double sig = method == 5 ? c->sigma : 0.;
    double scale = 1. / (double)c->F;
    double *Dx = c->d_Dx + s * G, *Dy = c->d_Dy + s * G, *Dz = c->d_Dz + s * G;
    double *g2 = c->d_Grad2 + s * G, *lap = c->d_Lap + s * G;
    mk::MulParams m{};
    m.kq = c->d_kq;
    m.sigma = sig;
    m.ek = method == 5 ? c->d_ek : nullptr;
    m.kek = method == 5 ? c->d_kek : nullptr;
    m.kind = mk::MUL_DERIV_PAIR;
    m.axis_a = 0;
    m.axis_b = 1;
    HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_work, m, Dx, Dy, scale, 0, 0., c->d_tw, c->st));
    m.kind = mk::MUL_DERIV_LAP;
    m.axis_a = c->D == 3 ? 2 : -1;
    HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_work, m, c->D == 3 ? Dz : nullptr, lap, scale, 0, 0., c->d_tw,
                               c->st));
    if (method != 5) HIPC(mk::sum_squares3(c->g, Dx, Dy, c->D == 3 ? Dz : nullptr, g2, c->st));
    if (method == 5) {
        HIPC(mk::fft_forward_sumsq(c->g, Dx, Dy, c->D == 3 ? Dz : nullptr, c->d_spec, c->d_tw, c->st));
        mk::MulParams ms{};
        ms.kind = mk::MUL_SMOOTH;
        ms.kq = c->d_kq;
        ms.sigma = sig;
        ms.ek = c->d_ek;
        ms.kek = c->d_kek;
        HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_spec, ms, g2, nullptr, scale, 1, 0., c->d_tw, c->st));
    }
    return MPDPFT_OK;
}

// CheckConvergence (:4759-4782) with RelDev (:4784-4802).
static int check_convergence(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    const double mp = 100. * 1.0e-16;
    if (c->SCcount > 0) {
        RC(field_stats(c));
        c->CC = 0.;
        double tot = 0.;
        for (int s = 0; s < c->S; s++) tot += c->TmpAb[s];
        if (p.ConvergenceMode == 1) {
            double r;
            HIPC(mk::max_abs_diff(c->SG, c->d_V, c->d_OldV, c->d_partials, c->st));
            RC(ctx_reduce_min(c, 1, &r));
            c->CC = -r;
        } else {
            double d[3 * MPDPFT_MAX_SPECIES];
            HIPC(mk::conv_dots_S(c->g, c->S, c->d_Den, c->d_OldDen, c->d_partials, c->st));
            RC(ctx_reduce(c, 3 * c->S, 1., d));
            double maxtest = 0.;
            for (int s = 0; s < c->S; s++) {
                double aa = d[3 * s], bb = d[3 * s + 1], ab = d[3 * s + 2];
                double rd = (aa + bb > kMP) ? 1. - 2. * ab / (aa + bb) : 0.;
                double test = rd * c->TmpAb[s] / tot;
                if (test > maxtest) {
                    maxtest = test;
                    c->CC = maxtest;
                }
            }
        }
        double dab = 0., dmu = 0., th = 0.;
        for (int s = 0; s < c->S; s++) {
            dab += (c->TmpAb[s] - c->OldAb[s]) * (c->TmpAb[s] - c->OldAb[s]);
            dmu += (c->muVec[s] - c->OldmuVec[s]) * (c->muVec[s] - c->OldmuVec[s]);
            th += c->thetaVec[s];
        }
        th /= (double)c->S;
        bool steady = c->CC < mp && std::sqrt(dab) < mp && std::sqrt(dmu) < mp;
        if ((tot > mp || (tot <= mp && (double)c->SCcount * th > 1.)) && p.SCcriterion >= 0. &&
            (c->CC < p.SCcriterion || steady))
            c->ExitSCcount++;
        else
            c->ExitSCcount = 0;
        if (c->ExitSCcount >= 3) c->ExitSCloop = 1;
    } else
        c->CC = 1.2345678901234567890;
    return MPDPFT_OK;
}

// Full-pivoting Householder QR inverse (the decomposition Eigen's
// FullPivHouseholderQR computes, with its rank rule), host FP64.
static bool fullpiv_inverse(int n, const double *A, double *inv) {
    std::vector<double> qr(A, A + n * n), hc(n);
    std::vector<int> rt(n), ct(n);
    const double eps = 2.220446049250313e-16;
    double prec = eps * n, biggest = 0., maxpivot = 0.;
    int nonzero = n;
    for (int k = 0; k < n; k++) {
        int rb = k, cb = k;
        double best = -1.;
        for (int j = k; j < n; j++)
            for (int i = k; i < n; i++)
                if (std::fabs(qr[i * n + j]) > best) {
                    best = std::fabs(qr[i * n + j]);
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
            for (int j = k; j < n; j++) std::swap(qr[k * n + j], qr[rb * n + j]);
        if (cb != k)
            for (int i = 0; i < n; i++) std::swap(qr[i * n + k], qr[i * n + cb]);
        // Householder on column k below the diagonal
        double c0 = qr[k * n + k], tail = 0.;
        for (int i = k + 1; i < n; i++) tail += qr[i * n + k] * qr[i * n + k];
        double tau, beta;
        if (tail <= 2.2250738585072014e-308) {
            tau = 0.;
            beta = c0;
            for (int i = k + 1; i < n; i++) qr[i * n + k] = 0.;
        } else {
            beta = std::sqrt(c0 * c0 + tail);
            if (c0 >= 0.) beta = -beta;
            for (int i = k + 1; i < n; i++) qr[i * n + k] /= (c0 - beta);
            tau = (beta - c0) / beta;
        }
        qr[k * n + k] = beta;
        hc[k] = tau;
        maxpivot = std::max(maxpivot, std::fabs(beta));
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
    for (int i = 0; i < nonzero; i++) rank += std::fabs(qr[i * n + i]) > thr;
    if (rank != n) return false;
    std::vector<int> perm(n);
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int k = 0; k < n; k++) std::swap(perm[k], perm[ct[k]]);
    std::vector<double> b(n), y(n);
    for (int col = 0; col < n; col++) {
        for (int i = 0; i < n; i++) b[i] = i == col ? 1. : 0.;
        for (int k = 0; k < n; k++) {
            if (rt[k] != k) std::swap(b[k], b[rt[k]]);
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
    return true;
}

// PulayMixer, mixer 1, Euclidean metric (:5005-5167).  History slots hold
// n_in and the residual R = n_out - n_in; only the newest Gram row/column is
// computed (one fused multi-dot pass), older entries come from the cache.
// |k| at grid index i (kVecAt, x slowest): the momentum metric's q of history position i.
static double kvec_norm(const mpdpft_ctx *c, int i) {
    double q2 = 0.;
    int rem = i;
    for (int d = c->D - 1; d >= 0; d--) {
        const double k = c->k0 + (double)(rem % c->EL) * c->dk;
        q2 += k * k;
        rem /= c->EL;
    }
    return std::sqrt(q2);
}

// PulayScalarProduct with the momentum-space metric (mpDPFT.cpp:5212-5227)
// from the full-spectrum dot D = sum over the (s+1)^D wrap-filled centred
// entries and the k = 0 product A0 (both summed over species): dk^D [(1 + w
// AngularFactor(D) sqrt(pi)/2) A0 + (1 + w/q^{D-1})(D - A0) - (G - 1) w/q^{D-1}
// EXP(-q^2) A0], q = |kVecAt[i]| at the first argument's history position i.
static double pulay_sp_momentum(const mpdpft_ctx *c, int i, double D, double A0) {
    const double w = c->pulay_weight, DD = (double)c->D;
    const double af = c->D == 2 ? 2. * kPI : 4. * kPI;  // AngularFactor, mpDPFT.h:291
    const double q = kvec_norm(c, i), wq = w / std::pow(q, DD - 1.), x = -q * q;
    const double ex = x < -700. ? 0. : std::exp(x);
    const double res = (1. + w * af * std::sqrt(kPI) / 2.) * A0 + (1. + wq) * (D - A0) -
                       ((double)c->G - 1.) * wq * ex * A0;
    return std::pow(c->dk, DD) * res;
}

static int pulay_push_row(mpdpft_ctx *c) {
    int scope = c->scope, ld = scope + 1;
    int I = c->pI;
    const int S = c->S;
    HIPC(mk::pulay_push(c->SG, c->d_OldDen, c->d_Den, c->nin[I], c->res[I], c->st));
    if (c->pulay_momentum) {  // the residuals in momentum space (:5028-5034)
        for (int s = 0; s < S; s++) {
            HIPC(mk::fft_forward_real(c->g, c->res[I] + s * c->G, nullptr, c->resF[I] + s * c->F, c->d_tw,
                                      c->d_scal + 200, c->st));
            HIPC(hipMemcpyAsync(c->h_spec0 + 2 * s, c->resF[I] + s * c->F, sizeof(double2), hipMemcpyDeviceToHost,
                                c->st));
        }
        HIPC(hipStreamSynchronize(c->st));
        for (int k = 0; k < 2 * S; k++) c->resF0[(size_t)I * 2 * S + k] = c->h_spec0[k];
    }
    I++;
    int shift = 0;
    if (I > scope) {  // drop the oldest entry (rotate slot pointers)
        double *a = c->nin[0], *b = c->res[0];
        double2 *f = c->pulay_momentum ? c->resF[0] : nullptr;
        std::vector<double> f0(c->resF0.begin(), c->resF0.begin() + (c->pulay_momentum ? 2 * S : 0));
        for (int k = 0; k < I - 1; k++) {
            c->nin[k] = c->nin[k + 1];
            c->res[k] = c->res[k + 1];
            if (f) {
                c->resF[k] = c->resF[k + 1];
                for (int q = 0; q < 2 * S; q++) c->resF0[(size_t)k * 2 * S + q] = c->resF0[(size_t)(k + 1) * 2 * S + q];
            }
        }
        c->nin[I - 1] = a;
        c->res[I - 1] = b;
        if (f) {
            c->resF[I - 1] = f;
            for (int q = 0; q < 2 * S; q++) c->resF0[(size_t)(I - 1) * 2 * S + q] = f0[q];
        }
        I--;
        shift = 1;
    }
    c->pI = I;
    // the new row (I-1, j) and column (i, I-1) of the Gram matrix; Euclidean: one symmetric dot per entry
    std::vector<double> row(I), col(I);
    if (c->pulay_momentum) {
        HIPC(mk::weight_nyquist(c->g, S, c->resF[I - 1], c->d_specx, c->st));
        std::vector<const double *> ys(I);
        for (int i = 0; i < I; i++) ys[i] = (const double *)c->resF[i];
        HIPC(mk::multi_dot(2 * (size_t)S * c->F, (const double *)c->d_specx, ys.data(), I, c->d_partials, c->st));
        std::vector<double> dots(I);
        RC(ctx_reduce(c, I, 1., dots.data()));
        // centred spectrum = dx^D (-1)^{sum q} DFT: products of two carry dx^{2D} (the signs cancel)
        const double sc = std::pow(c->dx, 2. * c->D);
        const double *n0 = &c->resF0[(size_t)(I - 1) * 2 * S];
        for (int i = 0; i < I; i++) {
            const double *f0 = &c->resF0[(size_t)i * 2 * S];
            double A0 = 0.;
            for (int s = 0; s < S; s++) A0 += f0[2 * s] * n0[2 * s] + f0[2 * s + 1] * n0[2 * s + 1];
            const double D = sc * dots[i];
            A0 *= sc;
            row[i] = pulay_sp_momentum(c, I - 1, D, A0);
            col[i] = pulay_sp_momentum(c, i, D, A0);
        }
    } else {
        HIPC(mk::multi_dot(c->SG, c->res[I - 1], c->res.data(), I, c->d_partials, c->st));
        RC(ctx_reduce(c, I, 1., row.data()));
        col = row;
    }
    std::vector<double> old(c->ResMat);
    for (int i = 0; i < I; i++)
        for (int j = 0; j < I; j++) {
            double v;
            if (i == I - 1) v = row[j];
            else if (j == I - 1) v = col[i];
            else v = old[(i + shift) * ld + (j + shift)];
            c->ResMat[i * ld + j] = v;
        }
    return MPDPFT_OK;
}

static int pulay_mixer(mpdpft_ctx *c, int *ok) {
    RC(pulay_push_row(c));
    if (c->fault_kind == MPDPFT_FAULT_DUPLICATE_PULAY && c->SCcount == c->fault_at)
        RC(pulay_push_row(c));  // forcing hook: the same (n_in, R) twice, a singular Gram matrix
    const int I = c->pI, ld = c->scope + 1;
    std::vector<double> M((size_t)I * I);
    for (int i = 0; i < I; i++)
        for (int j = 0; j < I; j++) M[i * I + j] = c->ResMat[i * ld + j];
    std::vector<double> inv((size_t)I * I);
    *ok = 0;
    if (fullpiv_inverse(I, M.data(), inv.data())) {
        double tv = 0.;
        for (int i = 0; i < I; i++)
            for (int j = 0; j < I; j++) {
                double acc = 0.;
                for (int k = 0; k < I; k++) acc += inv[i * I + k] * M[k * I + j];
                acc -= i == j ? 1. : 0.;
                tv += acc * acc;
            }
        if (tv <= 1.0e-8) {
            std::vector<double> x(I);
            double A = 0.;
            for (int i = 0; i < I; i++) {
                double acc = 0.;
                for (int k = 0; k < I; k++) acc += inv[i * I + k];
                x[i] = acc;
            }
            for (int i = 0; i < I; i++) A += x[i];
            for (int i = 0; i < I; i++) c->coeff[i] = x[i] / A;
            *ok = 1;
        } else {
            c->control << "PulayMixer validation of ResMat_inv failed: TestVal = " << tv << " (not 0)\n";
        }
    }
    if (!*ok) {
        c->pulay_rejections++;
        c->control << "PulayMixer: success = 0 ResidualMatrix is ill-conditioned\n";
    }
    return MPDPFT_OK;
}

// AdmixDensities (:4921-4992).
static int admix(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    int S = c->S;
    double nA = 0., nD = 0.;
    for (int s = 0; s < S; s++) {
        nD += (p.Abundances[s] - c->TmpAb[s]) * (p.Abundances[s] - c->TmpAb[s]);
        nA += p.Abundances[s] * p.Abundances[s];
    }
    double rel = std::sqrt(nD) / std::sqrt(nA), mod = 1.;
    if (rel > kMP) mod = std::min(1., std::pow(std::fabs(1. - rel), 30.));
    if (c->TrappedQ)
        for (int s = 0; s < S; s++) c->thetaVec[s] *= mod;
    int ok = 0;
    if (c->pulay_mixer > 0 && c->SCcount > 0 && c->CC < 1.0e+300) RC(pulay_mixer(c, &ok));
    double h[MPDPFT_MAX_SPECIES + 64];
    for (int s = 0; s < S; s++) h[s] = c->thetaVec[s];
    if (ok) {
        c->PulMix++;
        for (int i = 0; i < c->pI; i++) h[MPDPFT_MAX_SPECIES + i] = c->coeff[i];
        HIPC(hipMemcpyAsync(c->d_theta, h, sizeof(double) * S, hipMemcpyHostToDevice, c->st));
        HIPC(hipMemcpyAsync(c->d_coeff, h + MPDPFT_MAX_SPECIES, sizeof(double) * c->pI, hipMemcpyHostToDevice, c->st));
        HIPC(mk::pulay_mix(c->g, S, c->pI, c->nin.data(), c->res.data(), c->d_coeff, c->d_theta, c->d_Den, c->st));
    } else {
        c->LinMix++;
        HIPC(hipMemcpyAsync(c->d_theta, h, sizeof(double) * S, hipMemcpyHostToDevice, c->st));
        HIPC(mk::linear_mix(c->G, S, c->d_theta, c->d_OldDen, c->d_Den, c->st));
    }
    HIPC(hipStreamSynchronize(c->st));  // h[] is a stack buffer
    if (c->TrappedQ)
        for (int s = 0; s < S; s++) c->thetaVec[s] *= 1. / mod;
    return MPDPFT_OK;
}

// ---------------------------------------------------------------------------
// energies: GetEnergy (:5414-5436) and the routines it calls
// ---------------------------------------------------------------------------
static int energy_integral(mpdpft_ctx *c, const mk::EnergyParams &ep, const double *a, const double *b,
                           const double *cc, const double *d, double *out) {
    HIPC(mk::energy_integrand(c->g, ep, a, b, cc, d, c->d_partials, c->st));
    return integral_from_partials(c, out);
}

static int interaction_energy(mpdpft_ctx *c, int type, double *E) {  // :5805-5918
    const mpdpft_params &p = c->p;
    size_t G = c->G;
    *E = 0.;
    mk::EnergyParams ep{};
    ep.units = p.Units;
    if (type == 0) return MPDPFT_OK;
    if (type < 0) {
        double acc = 0.;
        ep.op = mk::E_XC;
        ep.xc_id = type;
        // the reference sums the species integrands first, then integrates;
        // the integral is linear, so summing per-species integrals is equivalent
        for (int s = 0; s < c->S; s++) {
            double v;
            RC(energy_integral(c, ep, c->d_Den + s * G, nullptr, nullptr, nullptr, &v));
            acc += v;
        }
        *E = acc;
        return MPDPFT_OK;
    }
    if (type == 1) {
        double acc = 0.;
        ep.op = mk::E_PROD;
        for (int s = 0; s < c->S; s++)
            for (int sp = s + 1; sp < c->S; sp++) {
                double v;
                RC(energy_integral(c, ep, c->d_Den + s * G, c->d_Den + sp * G, nullptr, nullptr, &v));
                acc += v;
            }
        *E = p.beta * acc;
        return MPDPFT_OK;
    }
    if (type == 8) {
        double acc = 0.;
        ep.op = mk::E_HALF_PROD;
        for (int s = 0; s < c->S; s++) {
            RC(hartree(c, c->d_Den + s * G, c->d_tmp2));
            double v;
            RC(energy_integral(c, ep, c->d_Den + s * G, c->d_tmp2, nullptr, nullptr, &v));
            acc += v;
        }
        *E = acc;
        return MPDPFT_OK;
    }
    if (type == 10) {
        ep.op = mk::E_PROD;
        double v;
        RC(energy_integral(c, ep, c->d_Den, c->d_Den, nullptr, nullptr, &v));
        *E = p.beta * v;
        return MPDPFT_OK;
    }
    if (c->D == 2 && type == 9) {  // RenormalizedContact energy (:5859-5862)
        ep.op = mk::E_RENORM;
        ep.p0 = p.beta;
        return energy_integral(c, ep, c->d_Den, c->d_Den + G, nullptr, nullptr, E);
    }
    if (c->D == 3 && c->S == 1 && (type == 12 || type == 13)) {  // Gombas / VWN energy (:5871-5878)
        ep.op = type == 12 ? mk::E_GOMBAS : mk::E_VWN;
        ep.p0 = p.gammaH;
        return energy_integral(c, ep, c->d_Den, nullptr, nullptr, nullptr, E);
    }
    if (c->D == 2 && type == 14) {  // DeltaEkin2D3Dcrossover energy, summed over species (:5879-5886)
        if (p.Units != 1) return MPDPFT_OK;
        ep.op = mk::E_CROSSOVER;
        ep.p0 = p.alpha;
        double acc = 0.;
        for (int s = 0; s < c->S; s++) {
            double v;
            RC(energy_integral(c, ep, c->d_Den + s * G, nullptr, nullptr, nullptr, &v));
            acc += v;
        }
        *E = acc;
        return MPDPFT_OK;
    }
    if (c->D == 2 && c->S == 2 && type == 24) {  // Quasi2Dcontact energy, counted once (:5890-5898)
        if (!(p.Units == 1 && std::fabs(p.beta) > kMP)) return MPDPFT_OK;
        ep.op = mk::E_QUASI2D;
        ep.p0 = p.alpha;
        ep.p1 = p.beta;
        ep.eta = c->d_eta;
        ep.eta_rows = c->eta_rows;
        return energy_integral(c, ep, c->d_Den, c->d_Den + G, nullptr, nullptr, E);
    }
    if (c->D == 3 && c->S == 1 && type == 11) {  // DiracExchange energy (:2762-2764)
        double DD = 3.;
        double cXD = 2. * std::sqrt(4. * kPI) / ((DD - 1.) * kPI) * std::pow(0.5 * std::pow(0.75 * std::sqrt(kPI), DD - 2.), 1. / DD);
        ep.op = mk::E_DIRAC;
        ep.p0 = -p.gammaH * cXD * DD / (1. + DD);
        return energy_integral(c, ep, c->d_Den, nullptr, nullptr, nullptr, E);
    }
    return MPDPFT_OK;
}

// GetAuxEint (:5518-5549): for S = 1, DIM 3, Units 2 the Dirac exchange energy,
// its potential-functional form -gammaH g^2/(16 pi^5) int (2*0.131163*(mu - V)_+)^2,
// and the Gombas and VWN correlation energies; zeros otherwise.
static int aux_eint(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    for (int s = 0; s < c->S; s++)
        for (int k = 0; k < 4; k++) c->AuxEint[s][k] = 0.;
    if (!(c->S == 1 && c->D == 3 && p.Units == 2)) return MPDPFT_OK;
    const double *n = c->d_Den;
    mk::EnergyParams ep{};
    ep.units = p.Units;
    double DD = 3.;
    double cXD = 2. * std::sqrt(4. * kPI) / ((DD - 1.) * kPI) * std::pow(0.5 * std::pow(0.75 * std::sqrt(kPI), DD - 2.), 1. / DD);
    ep.op = mk::E_DIRAC;
    ep.p0 = -p.gammaH * cXD * DD / (1. + DD);
    RC(energy_integral(c, ep, n, nullptr, nullptr, nullptr, &c->AuxEint[0][0]));
    ep.op = mk::E_QT2;
    ep.p0 = c->muVec[0];
    double v;
    RC(energy_integral(c, ep, c->d_V, nullptr, nullptr, nullptr, &v));
    c->AuxEint[0][1] = -p.gammaH * std::pow(p.degeneracy, 2.) / (16. * std::pow(kPI, 5.)) * v;
    ep.op = mk::E_GOMBAS;
    ep.p0 = p.gammaH;
    RC(energy_integral(c, ep, n, nullptr, nullptr, nullptr, &c->AuxEint[0][2]));
    ep.op = mk::E_VWN;
    RC(energy_integral(c, ep, n, nullptr, nullptr, nullptr, &c->AuxEint[0][3]));
    return MPDPFT_OK;
}

static int aux_ekin(mpdpft_ctx *c) {  // GetAuxEkin (:5461-5516)
    const mpdpft_params &p = c->p;
    size_t G = c->G;
    double DD = (double)c->D, g = p.degeneracy;
    for (int s = 0; s < c->S; s++) {
        const double *n = c->d_Den + s * G;
        mk::EnergyParams ep{};
        if (p.Units > 0) {
            double pref = c->D == 2 ? kPI / g : std::pow(243. * std::pow(kPI, 4.) / (250. * g * g), 1. / 3.);
            if (p.Units == 2) pref *= 1. / 0.131163;
            ep.op = mk::E_POW_POS;
            ep.p0 = (DD + 2.) / DD;
            double v;
            RC(energy_integral(c, ep, n, nullptr, nullptr, nullptr, &v));
            c->AuxEkin[s][0] = pref * v;
        } else
            c->AuxEkin[s][0] = 0.;
        if (c->D == 3) {
            RC(ctx_delfield(c, n, s, p.DelFieldMethod));
            ep.op = mk::E_GRADCORR;
            ep.p0 = 1.0e-6 * c->DenMax[s];
            double v;
            RC(energy_integral(c, ep, n, c->d_Grad2 + s * G, nullptr, nullptr, &v));
            v /= 72.;
            if (p.Units == 2) v *= 1. / 0.131163;
            c->AuxEkin[s][1] = v;
            c->AuxEkin[s][2] = 9. * v;
        } else
            c->AuxEkin[s][1] = c->AuxEkin[s][2] = 0.;
        if (c->S == 1 && c->D == 3) {
            RC(ctx_delfield(c, c->d_Env + s * G, s, p.DelFieldMethod));
            ep.op = mk::E_VIRIAL;
            double v;
            RC(energy_integral(c, ep, n, c->d_Dx + s * G, c->d_Dy + s * G, c->d_Dz + s * G, &v));
            c->AuxEkin[s][3] = 0.5 * (v - c->EInt);
        } else
            c->AuxEkin[s][3] = 0.;
        c->delfield_valid = -1;
    }
    return MPDPFT_OK;
}

static int dispersal_energy(mpdpft_ctx *c, int s, double *E) {  // :5551-5762
    const mpdpft_params &p = c->p;
    double DD = (double)c->D;
    *E = 0.;
    for (int e = 0; e < p.n_ekin; e++) {
        int t = p.EkinTypes[e];
        if (t == 0) {
            if (p.DensityExpression >= 6 && p.DensityExpression <= 10) {
                mk::EnergyParams ep{};
                ep.op = mk::E_TF_DISP;
                ep.p0 = c->muVec[s];
                double v;
                RC(energy_integral(c, ep, c->d_V + s * c->G, nullptr, nullptr, nullptr, &v));
                double U = p.Units == 2 ? 7.619964158443410 : 1.;
                double af = c->D == 2 ? 2. * kPI : 4. * kPI;
                *E += p.degeneracy * af * std::pow(2., (DD + 2.) / 2.) /
                      (std::pow(2. * kPI * std::sqrt(U), DD) * (2. * DD + 4.)) * v;
            } else if (p.DensityExpression == 5 && c->D == 3) {
                double v;
                RC(nait_dispersal_energy(c, s, &v));
                *E += v;
            }
        } else if (t > 0 && t <= 4)
            *E += c->AuxEkin[s][t - 1];
    }
    return MPDPFT_OK;
}

static int get_energy(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    c->EInt = 0.;
    c->Etot = 0.;
    for (int a = 0; a < p.n_interactions; a++) {
        RC(interaction_energy(c, p.Interactions[a], &c->IntE[a]));
        c->EInt += c->IntE[a];
    }
    c->Etot += c->EInt;
    RC(aux_ekin(c));
    RC(aux_eint(c));
    for (int s = 0; s < c->S; s++) {
        mk::EnergyParams ep{};
        ep.op = mk::E_PROD;
        RC(energy_integral(c, ep, c->d_Env + s * c->G, c->d_Den + s * c->G, nullptr, nullptr, &c->EnvE[s]));
        RC(dispersal_energy(c, s, &c->DispE[s]));
        c->Etot += c->DispE[s] + c->EnvE[s];
    }
    return MPDPFT_OK;
}

// ---------------------------------------------------------------------------
// the loop
// ---------------------------------------------------------------------------
static std::string vec_str(const double *v, int n) {
    std::ostringstream o;
    o.precision(16);
    for (int i = 0; i < n; i++) o << v[i] << " ";
    return o.str();
}

// The loop's scalar state at its start (mu, theta, T, DeltamuModifier,
// RelAcc, counters), InitializeHardCodedParameters / GetData (:858-1383,
// :135-201).
static void reset_loop_scalars(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    for (int s = 0; s < c->S; s++) {
        c->muVec[s] = p.muVec[s];
        c->thetaVec[s] = p.thetaVec[s];
        c->TVec[s] = p.TVec[s];
        c->TmpAb[s] = c->OldAb[s] = 0.;
    }
    c->DeltamuModifier = p.DeltamuModifier;
    c->RelAcc = c->RelAccOriginal = p.RelAcc;
    c->incV = c->incV0 = p.incrementalV;
    c->SCcount = c->ExitSCcount = c->ExitSCloop = 0;
    c->LinMix = c->PulMix = c->TrappedQ = 0;
    c->warningCount = c->importWarnings;
    c->importWarnings = 0;
    c->errorCount = c->bisection_steps = 0;
    c->bracket_failures = c->pulay_rejections = 0;
    c->MaxIteration = p.MaxIteration;
    c->CC = c->OldCC = 0.;
    for (int s = 0; s < c->S; s++) c->thetaVec0[s] = p.thetaVec[s];
    c->delfield_valid = -1;
    nait_reset_state(c);
}

static int store_movie_frame(mpdpft_ctx *c);
static void seed_noise(mpdpft_ctx *c);
static int impose_noise(mpdpft_ctx *c);

static int init_loop(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    reset_loop_scalars(c);
    c->pulay_mixer = (int)(p.Mixer[0] + 0.5);
    c->scope = std::max(1, (int)p.Mixer[1]);
    if ((int)c->nin.size() != c->scope + 1) {
        for (double *q : c->nin) hipFree(q);
        for (double *q : c->res) hipFree(q);
        c->nin.assign(c->scope + 1, nullptr);
        c->res.assign(c->scope + 1, nullptr);
        if (c->pulay_mixer > 0)
            for (int i = 0; i <= c->scope; i++) {
                HIPC(hipMalloc((void **)&c->nin[i], sizeof(double) * c->SG));
                HIPC(hipMalloc((void **)&c->res[i], sizeof(double) * c->SG));
            }
    }
    // InitializePulayMixer (:7029-7062): the metric weight (k0^2 + dk^2 when Mixer[2] < 0; the
    // reference leaves it unset when Mixer[2] = 0, taken as 0 here: Euclidean)
    c->pulay_weight = 0.;
    if (std::fabs(p.Mixer[2]) > kMP) c->pulay_weight = p.Mixer[2] < 0. ? c->k0 * c->k0 + c->dk * c->dk : p.Mixer[2];
    c->pulay_momentum = c->pulay_mixer > 0 && c->D > 1 && c->pulay_weight > 0.;
    if (c->pulay_momentum && (int)c->resF.size() != c->scope + 1) {
        for (double2 *q : c->resF) hipFree(q);
        c->resF.assign(c->scope + 1, nullptr);
        for (int i = 0; i <= c->scope; i++) HIPC(hipMalloc((void **)&c->resF[i], sizeof(double2) * c->S * c->F));
        if (!c->d_specx) RC(dmalloc(c, &c->d_specx, (size_t)c->S * c->F));
        if (!c->h_spec0) HIPC(hipHostMalloc((void **)&c->h_spec0, 2 * MPDPFT_MAX_SPECIES * sizeof(double),
                                            hipHostMallocDefault));
    }
    c->resF0.assign((size_t)(c->scope + 1) * 2 * c->S, 0.);
    c->pI = 0;
    c->ResMat.assign((size_t)(c->scope + 1) * (c->scope + 1), 0.);
    // initial densities (:194-201)
    RC(get_densities(c));
    seed_noise(c);
    c->noise = c->noise0 = p.Noise;
    RC(impose_noise(c));
    RC(field_stats(c));
    c->movie.clear();
    c->frames = 0;
    if (p.MovieQ) RC(store_movie_frame(c));  // :199
    c->loop_ready = 1;
    return MPDPFT_OK;
}

// StoreMovieData (mpDPFT.cpp:12317-12329) with SetupMovieStorage (:6422-6427):
// one frame of the cut interpolations of the current density, environment and
// potential per species.
static int store_movie_frame(mpdpft_ctx *c) {
    std::vector<double> den(c->SG), env(c->SG), V(c->SG), fi;
    HIPC(hipMemcpyAsync(den.data(), c->d_Den, sizeof(double) * c->SG, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(env.data(), c->d_Env, sizeof(double) * c->SG, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(V.data(), c->d_V, sizeof(double) * c->SG, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    const double CutLine = cut_fields(c, den, env, V, fi);
    const int EL = c->EL;
    if (c->movie.empty()) {
        c->movie.resize(EL);
        for (int i = 0; i < EL; i++) c->movie[i].push_back(-CutLine / 2. + (double)i * CutLine / ((double)c->s));
    }
    for (int i = 0; i < EL; i++)
        for (int s = 0; s < c->S; s++)
            for (int p = 0; p < 3; p++) c->movie[i].push_back(fi[((size_t)s * 3 + p) * EL + i]);
    c->frames++;
    return MPDPFT_OK;
}

// ManualSCloopBreakQ (mpDPFT.cpp:11959-11976): the first integer of the break
// file; a nonzero value is reported, the file reset to 0 and Print set to 2.
static int manual_break(mpdpft_ctx *c) {
    if (c->break_file.empty()) return 0;
    FILE *f = std::fopen(c->break_file.c_str(), "r");
    if (!f) return 0;
    int Q = 0;
    if (std::fscanf(f, "%d", &Q) != 1) Q = 0;
    std::fclose(f);
    if (Q) {
        if (Q == 1)
            c->control << "---------------------------------------SCloop--Manually--Terminated----------------------"
                          "-----------------\n";
        if (Q == 2)
            c->control << "---------------------------------------TASKloop--Manually--Terminated--------------------"
                          "-------------------\n";
        FILE *w = std::fopen(c->break_file.c_str(), "w");
        if (w) {
            std::fprintf(w, "0\n");
            std::fclose(w);
        }
        c->p.Print = 2;
    }
    return Q;
}

// InitializeRandomNumberGenerator (mpDPFT.cpp:6521-6540): the seed is the
// extension key NoiseSeed, or, when that is 0, the reference's own from the
// clock (the fractional part of ns/1e10 times 1e10, folded into [1e6, INT_MAX)).
static void seed_noise(mpdpft_ctx *c) {
    unsigned long long seed = c->p.NoiseSeed;
    if (seed == 0) {
        const double MaxInt = (double)std::numeric_limits<int>::max();
        const double Now =
            std::fabs((double)std::chrono::high_resolution_clock::now().time_since_epoch().count() / 1.0e10);
        double sd = 1.0e10 * (Now - (double)((int)Now));
        if (sd < 1.0e-6) sd = (double)std::rand();
        while (sd >= MaxInt) sd /= 2.;
        while (sd < 1.0e6) sd *= 2.;
        seed = (unsigned long long)(int)sd;
    }
    c->rng.seed(seed);
}

// ImposeNoise(1) (mpDPFT.cpp:4490-4504): Den_s[i] *= 1 + Noise RN with RN
// uniform on [-1, 1) (std::uniform_real_distribution on std::mt19937_64, as
// the reference draws them), species by species in grid order (the
// reference's order with one thread); only with Noise > 0 and InterpolVQ = 0
// or a schedule > 0.  The factors are drawn on the host and applied on the
// device.
static int impose_noise(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    bool scheduleQ = false;
    for (int i = 0; i < p.n_schedule; i++)
        if (p.Schedule[i] > 0) scheduleQ = true;
    if (!(c->noise > 0. && (p.InterpolVQ == 0 || scheduleQ))) return MPDPFT_OK;
    if (!c->d_noise) RC(dmalloc(c, &c->d_noise, c->SG));
    std::uniform_real_distribution<double> RN(-1.0, 1.0);
    std::vector<double> f(c->SG);
    for (size_t k = 0; k < c->SG; k++) f[k] = 1. + c->noise * RN(c->rng);
    HIPC(hipMemcpyAsync(c->d_noise, f.data(), sizeof(double) * c->SG, hipMemcpyHostToDevice, c->st));
    HIPC(mk::mul_fields(c->SG, c->d_noise, c->d_Den, c->st));
    HIPC(hipStreamSynchronize(c->st));  // f lives on this stack frame
    return MPDPFT_OK;
}

// InjectSchedule (mpDPFT.cpp:5229-5315): 5 (theta up by 1.1 while CC falls,
// else halved, clipped to [10/maxSCcount, 0.8]), 9 (theta ramped linearly
// towards 0.2) and the noise schedules on the mixed density: 3 (at SCcount =
// i 10^e, i = 1..9; at SCcount 0 all nine i match), 4 and 7 (Noise =
// 0.1 exp(-40 or -10 SCcount/maxSCcount) every iteration), 11 (Noise =
// NoiseOriginal exp(-40 SCcount/(0.1 maxSCcount)) during the first tenth of
// maxSCcount).
static int inject_schedule(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    const double sc_frac = (double)c->SCcount / ((double)p.maxSCcount);
    for (int i = 0; i < p.n_schedule; i++) {
        int sc = p.Schedule[i];
        if (sc == 3) {
            const double expon = std::floor(std::log10((double)c->SCcount));
            for (int k = 1; k < 10; k++)
                if (k * (int)std::pow(10., expon) == c->SCcount) {
                    c->control << "InjectSchedule: ImposeNoise on densities\n";
                    RC(impose_noise(c));
                }
        } else if (sc == 4 || sc == 7) {
            c->noise = 0.1 * std::exp((sc == 4 ? -40. : -10.) * sc_frac);
            RC(impose_noise(c));
            if (c->PrintSC) c->control << "InjectSchedule: ImposeNoise of magnitude " << c->noise << "\n";
        } else if (sc == 11) {
            const bool early = (double)c->SCcount < 0.1 * (double)p.maxSCcount;
            if (c->noise0 > kMP && early) {
                c->noise = c->noise0 * std::exp(-40. * (double)c->SCcount / (0.1 * (double)p.maxSCcount));
                RC(impose_noise(c));
            }
            if (c->PrintSC && early) c->control << "InjectSchedule: ImposeNoise of magnitude " << c->noise << "\n";
        } else if (sc == 5) {
            for (int s = 0; s < c->S; s++) {
                if (c->CC < c->OldCC) c->thetaVec[s] *= 1.1;
                else c->thetaVec[s] *= 0.5;
                c->thetaVec[s] = std::max(10. / ((double)p.maxSCcount), std::min(0.8, c->thetaVec[s]));
            }
            if (p.Print >= 0 && c->PrintSC)
                c->control << "InjectSchedule: Update thetaVec -> " << vec_str(c->thetaVec, c->S) << "\n";
        } else if (sc == 9) {
            const double maxTheta = 0.2;
            for (int s = 0; s < c->S; s++)
                if (c->thetaVec[s] < maxTheta)
                    c->thetaVec[s] = c->thetaVec0[s] + (maxTheta - c->thetaVec0[s]) * (double)c->SCcount /
                                                           ((double)p.maxSCcount);
        }
    }
    return MPDPFT_OK;
}

static int iteration(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    // StoreCurrentIteration (:4804): swap instead of copying; UpdateV then
    // reads the mixed density from OldDen and GetDensities overwrites Den.
    std::swap(c->d_Den, c->d_OldDen);
    std::swap(c->d_V, c->d_OldV);
    c->OldCC = c->CC;
    for (int s = 0; s < c->S; s++) {
        c->OldmuVec[s] = c->muVec[s];
        c->OldAb[s] = c->TmpAb[s];
    }
    RC(update_v(c, c->d_OldDen));
    if (c->fault_kind == MPDPFT_FAULT_NAN_POTENTIAL && c->SCcount == c->fault_at) {  // forcing hook
        const double nan = std::nan("");
        HIPC(hipMemcpyAsync(c->d_V + c->CentreIndex, &nan, sizeof(double), hipMemcpyHostToDevice, c->st));
        HIPC(hipStreamSynchronize(c->st));  // nan is a stack value
    }
    for (int s = 0; s < c->S; s++)  // PreparationsForSCdensity (:4506-4529)
        if (c->TVec[s] < 0.) c->TVec[s] = std::max(std::fabs(c->muVec[s] - c->VMin[s]), std::fabs(c->muVec[s] - c->VMax[s])) / 50.;
    RC(get_densities(c));
    RC(check_convergence(c));
    if (c->fp32_phase) {
        // Precision 1: the FP32 stages end at max|dV| < max(criterion, 1e-6 max|V|) for three
        // iterations (overlap mode: CC < 1e-12), or after 16 iterations without a new smallest CC;
        // the loop then continues in FP64 (polish) with the instance's own criterion
        double vmax = 0.;
        for (int s = 0; s < c->S; s++) vmax = std::max(vmax, std::max(std::fabs(c->VMin[s]), std::fabs(c->VMax[s])));
        double floorA = p.ConvergenceMode == 1 ? std::max(p.SCcriterion, 1e-6 * vmax) : std::max(p.SCcriterion, 1e-12);
        bool stop = c->ExitSCloop != 0;
        if (c->SCcount > 0) {
            c->fp32_hits = c->CC < floorA ? c->fp32_hits + 1 : 0;
            if (c->CC < c->fp32_best) {
                c->fp32_best = c->CC;
                c->fp32_stall = 0;
            } else {
                c->fp32_stall++;
            }
            if (c->fp32_hits >= 3 || c->fp32_stall >= 16) stop = true;
        }
        if (stop) {
            c->fp32_phase = 0;
            c->fp32_iterations = c->SCcount + 1;
            c->ExitSCcount = 0;
            c->ExitSCloop = 0;
            c->delfield_valid = -1;
            c->control << " ***** FP32 stages done after " << c->fp32_iterations << " iterations (CC = " << c->CC
                       << "); FP64 polish *****\n";
        }
    }
    // Intercept (:4852-4879)
    if (c->PrintSC == 1)
        c->control << " ***** SCcount = " << c->SCcount << "(" << c->ExitSCcount << "/3);  CC = "
                   << std::setprecision(8) << c->CC << std::setprecision(16) << " *****\n";
    if (p.Print >= 0 && manual_break(c) > 0) c->ExitSCloop = 1;  // ManualSCloopBreakQ (:4864)
    if (c->SCcount == p.maxSCcount) c->ExitSCloop = 1;
    if (c->verbose) {  // the reference prints every iteration to stdout (PRINT, mpDPFT.cpp:12016)
        double bp = 0., mi = 0.;
        if (p.DensityExpression == 5) nait_state(c, 0, &bp, &mi);
        std::fprintf(stderr,
                     " ***** SCcount = %d(%d/3);  CC = %.8g;  mu0 = %.12g;  density evaluations = %d;  nAiT BP %g "
                     "maxIter %g *****\n",
                     c->SCcount, c->ExitSCcount, c->CC, c->muVec[0], c->bisection_steps, bp, mi);
        std::fflush(stderr);
    }
    if (p.DensityExpression == 5) {
        if (p.Mpp[2] > 0.) nait_adapt(c);  // AdaptnAiT (:4859)
        // CalculateEnergyQ turns true at the exit iteration (:4870) and the
        // energy computed there raises MaxIteration for the final one
        if (c->ExitSCcount == 3 && c->PrintSC == 1) RC(get_energy(c));
    }
    if (c->ExitSCcount == 3 && c->PrintSC == 1) {
        c->control << "\n ***** exit SCloop;         muVec = " << vec_str(c->muVec, c->S) << " *****\n";
        c->control << " *****               NVec(Target) = " << vec_str(p.Abundances, c->S) << " *****\n";
        c->control << " *****                       NVec = " << vec_str(c->TmpAb, c->S) << " *****\n";
    }
    RC(admix(c));
    RC(inject_schedule(c));
    // PrepareNextIteration (:5317-5389): counter and print cadence
    c->SCcount++;
    if (p.Print > 0) c->PrintSC = 1;
    else {
        int e = (int)std::floor(std::log10((double)c->SCcount));
        int base = (int)std::pow(10., e);
        bool pr = c->SCcount % base == 0 && c->SCcount / base < 10;
        c->PrintSC = (pr || c->SCcount % 1000 == 0 || c->SCcount == p.maxSCcount || c->ExitSCcount > 0) ? 1 : 0;
    }
    if (p.MovieQ && c->PrintSC == 1 && (c->ExitSCcount == 0 || c->ExitSCloop)) RC(store_movie_frame(c));  // :5342
    return MPDPFT_OK;
}

// The loop's start as in GetData (:135-201): V = Env (or the imported
// potential), the initial densities.
static int loop_begin(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    HIPC(hipSetDevice(c->device));
    c->control.str("");
    c->control.precision(16);
    c->control << " ***** mpdpft_hip: self-consistent DPFT loop on " << "gfx950 *****\n";
    c->control << c->notes.str();
    c->notes.str("");
    if (c->V0.empty())  // InitializeEnvironments: V <- Env (:6388)
        HIPC(hipMemcpyAsync(c->d_V, c->d_Env, sizeof(double) * c->SG, hipMemcpyDeviceToDevice, c->st));
    else  // LoadV (InterpolVQ = 1, :184)
        HIPC(hipMemcpyAsync(c->d_V, c->V0.data(), sizeof(double) * c->SG, hipMemcpyHostToDevice, c->st));
    c->fp32_phase = p.Precision == 1 ? 1 : 0;
    c->fp32_iterations = 0;
    c->fp32_stall = 0;
    c->fp32_hits = 0;
    c->fp32_best = 1.0e300;
    RC(init_loop(c));
    c->PrintSC = p.Print > 0 ? 1 : 0;
    if (p.maxSCcount > 0) c->control << "\n S E L F C O N S I S T E N T   L O O P ...\n\n";
    else c->ExitSCloop = 1;
    return MPDPFT_OK;
}

static void fill_result(const mpdpft_ctx *c, mpdpft_result *r) {
    const mpdpft_params &p = c->p;
    std::memset(r, 0, sizeof(*r));
    r->S = c->S;
    r->SCcount = c->SCcount;
    r->converged = c->ExitSCcount >= 3;
    r->LinMix = c->LinMix;
    r->PulMix = c->PulMix;
    r->errorCount = c->errorCount;
    r->warningCount = c->warningCount;
    r->CC = c->CC;
    r->DeltamuModifier = c->DeltamuModifier;
    r->TotalEnergy = c->Etot;
    r->solve_seconds = c->solve_seconds;
    r->bisection_steps = c->bisection_steps;
    r->fp32_iterations = c->fp32_iterations;
    r->bracket_failures = c->bracket_failures;
    r->pulay_rejections = c->pulay_rejections;
    for (int s = 0; s < c->S; s++) {
        r->muVec[s] = c->muVec[s];
        r->FinalAbundances[s] = c->TmpAb[s];
        r->DispersalEnergies[s] = c->DispE[s];
        r->EnvironmentalEnergies[s] = c->EnvE[s];
        for (int k = 0; k < 4; k++) r->AuxEkin[s][k] = c->AuxEkin[s][k];
        for (int k = 0; k < 4; k++) r->AuxEint[s][k] = c->AuxEint[s][k];
    }
    for (int a = 0; a < p.n_interactions; a++) r->InteractionEnergies[a] = c->IntE[a];
}

static void clear_energies(mpdpft_ctx *c) {
    c->Etot = c->EInt = 0.;
    for (int s = 0; s < MPDPFT_MAX_SPECIES; s++) {
        c->DispE[s] = c->EnvE[s] = 0.;
        for (int k = 0; k < 4; k++) c->AuxEkin[s][k] = c->AuxEint[s][k] = 0.;
    }
    for (int a = 0; a < MPDPFT_MAX_LIST; a++) c->IntE[a] = 0.;
}

extern "C" int mpdpft_solve(mpdpft_ctx *c, mpdpft_result *r) {
    if (!c || !r) return MPDPFT_EINVAL;
    if (!c->st) return ctx_fail(c, MPDPFT_EINVAL, "context was not created");
    const mpdpft_params &p = c->p;
    auto t0 = std::chrono::steady_clock::now();
    RC(loop_begin(c));
    while (!c->ExitSCloop) RC(iteration(c));
    if (p.maxSCcount > 0 && c->SCcount >= p.maxSCcount) {
        c->warningCount++;
        c->control << "\n Warning!!! SC loop not converged (CC=" << c->CC << ") \n";
    }
    HIPC(hipStreamSynchronize(c->st));
    auto t1 = std::chrono::steady_clock::now();
    RC(get_energy(c));
    HIPC(hipStreamSynchronize(c->st));
    c->solve_seconds = std::chrono::duration<double>(t1 - t0).count();
    fill_result(c, r);
    return (r->converged || p.maxSCcount == 0) ? MPDPFT_OK : MPDPFT_ENOTCONV;
}

extern "C" int mpdpft_loop_begin(mpdpft_ctx *c) {
    if (!c) return MPDPFT_EINVAL;
    if (!c->st) return ctx_fail(c, MPDPFT_EINVAL, "context was not created");
    clear_energies(c);
    return loop_begin(c);
}

extern "C" int mpdpft_step(mpdpft_ctx *c, mpdpft_result *r, int *exit_loop) {
    if (!c) return MPDPFT_EINVAL;
    if (!c->loop_ready) return ctx_fail(c, MPDPFT_EINVAL, "mpdpft_step: call mpdpft_loop_begin first");
    HIPC(hipSetDevice(c->device));
    if (!c->ExitSCloop) {
        RC(iteration(c));
        if (c->ExitSCloop && c->p.maxSCcount > 0 && c->SCcount >= c->p.maxSCcount) {
            c->warningCount++;
            c->control << "\n Warning!!! SC loop not converged (CC=" << c->CC << ") \n";
        }
    }
    HIPC(hipStreamSynchronize(c->st));
    if (r) fill_result(c, r);
    if (exit_loop) *exit_loop = c->ExitSCloop;
    return MPDPFT_OK;
}

extern "C" int mpdpft_energy(mpdpft_ctx *c, mpdpft_result *r) {
    if (!c || !r) return MPDPFT_EINVAL;
    if (!c->st) return ctx_fail(c, MPDPFT_EINVAL, "context was not created");
    HIPC(hipSetDevice(c->device));
    RC(get_energy(c));
    HIPC(hipStreamSynchronize(c->st));
    fill_result(c, r);
    return MPDPFT_OK;
}

extern "C" int mpdpft_enforce_constraints(mpdpft_ctx *c, double *mu, double *N) {
    if (!c || !mu) return MPDPFT_EINVAL;
    if (!c->st) return ctx_fail(c, MPDPFT_EINVAL, "context was not created");
    HIPC(hipSetDevice(c->device));
    for (int s = 0; s < c->S; s++) c->muVec[s] = mu[s];
    const int before = c->bracket_failures;
    c->delfield_valid = -1;  // the potential may have been replaced
    RC(get_densities(c));
    HIPC(hipStreamSynchronize(c->st));
    for (int s = 0; s < c->S; s++) {
        mu[s] = c->muVec[s];
        if (N) N[s] = c->TmpAb[s];
    }
    return c->bracket_failures > before ? MPDPFT_ENOBRACKET : MPDPFT_OK;
}

extern "C" int mpdpft_mix(mpdpft_ctx *c, int sc_count, const double *n_out) {
    if (!c || !n_out || sc_count < 0) return MPDPFT_EINVAL;
    if (!c->loop_ready) return ctx_fail(c, MPDPFT_EINVAL, "mpdpft_mix: call mpdpft_loop_begin first (Pulay history)");
    HIPC(hipSetDevice(c->device));
    std::swap(c->d_Den, c->d_OldDen);  // the current density is this step's n_in
    HIPC(hipMemcpyAsync(c->d_Den, n_out, sizeof(double) * c->SG, hipMemcpyHostToDevice, c->st));
    const int keep = c->SCcount, before = c->pulay_rejections;
    c->SCcount = sc_count;
    const int rc = admix(c);
    c->SCcount = keep;
    RC(rc);
    HIPC(hipStreamSynchronize(c->st));
    return c->pulay_rejections > before ? MPDPFT_EPULAY : MPDPFT_OK;
}

extern "C" int mpdpft_inject_fault(mpdpft_ctx *c, int kind, int at) {
    if (!c) return MPDPFT_EINVAL;
    if (kind != 0 && kind != MPDPFT_FAULT_DUPLICATE_PULAY && kind != MPDPFT_FAULT_NAN_POTENTIAL)
        return ctx_fail(c, MPDPFT_EINVAL, "mpdpft_inject_fault: unknown kind");
    c->fault_kind = at < 0 ? 0 : kind;
    c->fault_at = at;
    return MPDPFT_OK;
}

// ---------------------------------------------------------------------------
// field access
// ---------------------------------------------------------------------------
static int copy_out(mpdpft_ctx *c, const double *d, double *h) {
    if (!c || !h) return MPDPFT_EINVAL;
    HIPC(hipMemcpyAsync(h, d, sizeof(double) * c->SG, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return MPDPFT_OK;
}
extern "C" int mpdpft_get_density(mpdpft_ctx *c, double *h) { return c ? copy_out(c, c->d_Den, h) : MPDPFT_EINVAL; }
extern "C" int mpdpft_get_potential(mpdpft_ctx *c, double *h) { return c ? copy_out(c, c->d_V, h) : MPDPFT_EINVAL; }
extern "C" int mpdpft_get_environment(mpdpft_ctx *c, double *h) { return c ? copy_out(c, c->d_Env, h) : MPDPFT_EINVAL; }
extern "C" int mpdpft_set_initial_potential(mpdpft_ctx *c, const double *h) {
    if (!c || !c->st) return MPDPFT_EINVAL;
    if (h) c->V0.assign(h, h + c->SG);
    else c->V0.clear();
    return MPDPFT_OK;
}

extern "C" int mpdpft_set_break_file(mpdpft_ctx *c, const char *path) {
    if (!c) return MPDPFT_EINVAL;
    c->break_file = path ? path : "";
    return MPDPFT_OK;
}

extern "C" int mpdpft_set_potential(mpdpft_ctx *c, const double *h) {
    if (!c || !h) return MPDPFT_EINVAL;
    HIPC(hipMemcpyAsync(c->d_V, h, sizeof(double) * c->SG, hipMemcpyHostToDevice, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return MPDPFT_OK;
}

// ---------------------------------------------------------------------------
// stage entry points (use the context's buffers as scratch)
// ---------------------------------------------------------------------------
extern "C" int mpdpft_fft(mpdpft_ctx *c, int sign, double *hc) {
    if (!c || !hc || (sign != -1 && sign != 1)) return MPDPFT_EINVAL;
    int s = c->s, D = c->D, EL = c->EL;
    size_t G = c->G, F = c->F;
    std::vector<double> re(G), im(G);
    auto gidx = [&](int i, int j, int k) -> size_t {
        return D == 3 ? ((size_t)i * EL + j) * EL + k : (size_t)j * EL + k;
    };
    double dxD = std::pow(c->dx, (double)D), LD = std::pow(1. / c->p.edge, (double)D);
    double gph = (((long long)D * s / 2) & 1) ? -1. : 1.;
    if (sign < 0) {
        for (size_t i = 0; i < G; i++) {
            re[i] = hc[2 * i];
            im[i] = hc[2 * i + 1];
        }
        HIPC(hipMemcpy(c->d_tmp1, re.data(), sizeof(double) * G, hipMemcpyHostToDevice));
        HIPC(hipMemcpy(c->d_tmp2, im.data(), sizeof(double) * G, hipMemcpyHostToDevice));
        HIPC(mk::fft_forward_real(c->g, c->d_tmp1, c->d_tmp2, c->d_spec, c->d_tw, nullptr, c->st));
        std::vector<double2> sp(F);
        HIPC(hipMemcpyAsync(sp.data(), c->d_spec, sizeof(double2) * F, hipMemcpyDeviceToHost, c->st));
        HIPC(hipStreamSynchronize(c->st));
        // centred value at index p = dx^D (-1)^{sum q} e^{-i D pi s/2} spec[q], q = (p + s/2) mod s
        for (size_t gi = 0; gi < G; gi++) {
            int idx[3] = {0, 0, 0};
            size_t rem = gi;
            for (int d = 2; d >= 3 - D; d--) {
                idx[d] = (int)(rem % EL);
                rem /= EL;
            }
            int q[3] = {0, 0, 0}, par = 0;
            for (int d = 3 - D; d < 3; d++) {
                int pp = idx[d] % s;
                q[d] = (pp + s / 2) % s;
                par += q[d];
            }
            size_t f = ((size_t)q[0] * s + q[1]) * s + q[2];
            double sg = (par & 1) ? -gph : gph;
            hc[2 * gi] = dxD * sg * sp[f].x;
            hc[2 * gi + 1] = dxD * sg * sp[f].y;
        }
        return MPDPFT_OK;
    }
    std::vector<double2> sp(F);
    for (int a = 0; a < (D == 3 ? s : 1); a++)
        for (int b = 0; b < s; b++)
            for (int k = 0; k < s; k++) {
                int q[3] = {a, b, k}, par = 0, pc[3] = {0, 0, 0};
                for (int d = 3 - D; d < 3; d++) {
                    pc[d] = (q[d] + s / 2) % s;
                    par += q[d];
                }
                double sg = (par & 1) ? -gph : gph;
                size_t gi = gidx(pc[0], pc[1], pc[2]);
                size_t f = ((size_t)a * s + b) * s + k;
                sp[f].x = sg * hc[2 * gi];
                sp[f].y = sg * hc[2 * gi + 1];
            }
    HIPC(hipMemcpy(c->d_spec, sp.data(), sizeof(double2) * F, hipMemcpyHostToDevice));
    mk::MulParams m{};
    m.kind = mk::MUL_NONE;
    m.kq = c->d_kq;
    HIPC(mk::fft_backward_real(c->g, c->d_spec, c->d_spec, m, c->d_tmp1, c->d_tmp2, LD, 0, 0., c->d_tw, c->st));
    HIPC(hipMemcpyAsync(re.data(), c->d_tmp1, sizeof(double) * G, hipMemcpyDeviceToHost, c->st));
    HIPC(hipMemcpyAsync(im.data(), c->d_tmp2, sizeof(double) * G, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    for (size_t i = 0; i < G; i++) {
        hc[2 * i] = re[i];
        hc[2 * i + 1] = im[i];
    }
    return MPDPFT_OK;
}

extern "C" int mpdpft_delfield(mpdpft_ctx *c, const double *f, double *dx, double *dy, double *dz, double *gradsq,
                               double *laplacian) {
    if (!c || !f) return MPDPFT_EINVAL;
    size_t G = c->G;
    HIPC(hipMemcpy(c->d_tmp1, f, sizeof(double) * G, hipMemcpyHostToDevice));
    RC(ctx_delfield(c, c->d_tmp1, 0, c->p.DelFieldMethod));
    HIPC(hipStreamSynchronize(c->st));
    if (dx) HIPC(hipMemcpy(dx, c->d_Dx, sizeof(double) * G, hipMemcpyDeviceToHost));
    if (dy) HIPC(hipMemcpy(dy, c->d_Dy, sizeof(double) * G, hipMemcpyDeviceToHost));
    if (dz && c->D == 3) HIPC(hipMemcpy(dz, c->d_Dz, sizeof(double) * G, hipMemcpyDeviceToHost));
    if (gradsq) HIPC(hipMemcpy(gradsq, c->d_Grad2, sizeof(double) * G, hipMemcpyDeviceToHost));
    if (laplacian) HIPC(hipMemcpy(laplacian, c->d_Lap, sizeof(double) * G, hipMemcpyDeviceToHost));
    c->delfield_valid = -1;
    return MPDPFT_OK;
}

extern "C" int mpdpft_interaction_potential(mpdpft_ctx *c, int type, int s, const double *den, double *vint) {
    if (!c || !den || !vint || s < 0 || s >= c->S) return MPDPFT_EINVAL;
    size_t G = c->G;
    HIPC(hipMemcpy(c->d_OldDen, den, sizeof(double) * c->SG, hipMemcpyHostToDevice));
    if (type == 8) {
        RC(hartree(c, c->d_OldDen + s * G, c->d_tmp1));
    } else {
        if (type == 15 || type == 17) RC(dipolar_fields(c, c->d_OldDen));
        mk::VintParams vp{};
        vint_extra(c, vp);
        vp.n_inter = 1;
        vp.inter[0] = type;
        vp.S = c->S;
        vp.s = c->s;
        vp.beta = c->p.beta;
        vp.gammaH = c->p.gammaH;
        vp.inc = 1.;
        vp.units = c->p.Units;
        HIPC(mk::fill(c->SG, 0., c->d_OldV, c->st));
        HIPC(mk::update_v(c->g, vp, c->d_OldDen, c->d_OldV, c->d_Dx, c->st));
        HIPC(hipMemcpyAsync(c->d_tmp1, c->d_Dx + s * G, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
    }
    HIPC(hipMemcpyAsync(vint, c->d_tmp1, sizeof(double) * G, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return MPDPFT_OK;
}

extern "C" int mpdpft_density(mpdpft_ctx *c, int s, double mu, const double *v, double *den, double *N) {
    if (!c || !v || s < 0 || s >= c->S) return MPDPFT_EINVAL;
    size_t G = c->G;
    HIPC(hipMemcpy(c->d_V + s * G, v, sizeof(double) * G, hipMemcpyHostToDevice));
    c->delfield_valid = -1;
    double n;
    RC(density(c, s, mu, c->d_Den + s * G, &n));
    if (den) HIPC(hipMemcpy(den, c->d_Den + s * G, sizeof(double) * G, hipMemcpyDeviceToHost));
    if (N) *N = n;
    return MPDPFT_OK;
}
