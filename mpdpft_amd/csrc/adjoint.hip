// adjoint.hip -- implicit adjoint of the converged self-consistent fixed point.
//
// At convergence (one species, incrementalV ramp finished) the loop of GetData
// (mpDPFT.cpp:210-225) satisfies
//     V = Env(theta) + K[n],      n = F(V, mu),      sum_i w_i n_i = N,
// with K the interaction potentials of UpdateV/getVint (:4815-4850, :2240-2318;
// isolated Hartree, contact 10, Dirac 11, LDA -1/-12) and F the density of
// GetDensity (:2142-2238; TF 7 or nAiT 5, whose gradient terms come from
// DelField(V), :7243-7582).  Mixing does not move the fixed point.  For a loss
// L(n) the implicit function theorem gives dL/dEnv_i = lambda_i, where
//     lambda = F_V^T P (l + K'^T lambda),   l = dL/dn,
//     P x = x - w (F_mu . x) / (F_mu . w),  F_mu = -dF/dU,
// (the second row eliminates the chemical potential's multiplier).  The
// linear system (I - F_V^T P K'^T) lambda = F_V^T P l is solved by restarted
// GMRES on the device.  When the loop holds mu instead of N (RelAcc = 0: no
// mu search in EnforceConstraints, :4578) the constraint row is absent and P
// is the identity.  F_V^T is the pointwise dF/dU plus the transposed
// DelField chain (taper with its shell-maximum shift, derivative and
// Laplacian multipliers, the POS-clipped smoothed |grad V|^2); K'^T is the
// transposed isolated Hartree convolution plus the pointwise dV_xc/dn.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "solver.h"

#define HIPC(x)                                                           \
    do {                                                                  \
        hipError_t e_ = (x);                                              \
        if (e_ != hipSuccess) return hip_fail(c, e_, #x);                 \
    } while (0)
#define RC(x)                  \
    do {                       \
        int r_ = (x);          \
        if (r_) return r_;     \
    } while (0)

struct AdjointState {
    size_t G = 0, F = 0;
    int m = 0;  // GMRES restart length
    double *dfdU = nullptr, *dfdg2 = nullptr, *dfdl = nullptr, *xcd = nullptr, *nbuf = nullptr;
    double *Dx = nullptr, *Dy = nullptr, *Dz = nullptr, *g2 = nullptr;
    double *ell = nullptr, *h = nullptr, *x = nullptr, *direct = nullptr, *mug = nullptr, *mul = nullptr;
    double *t1 = nullptr, *t2 = nullptr, *t3 = nullptr, *fbar = nullptr, *w = nullptr, *rhs = nullptr, *res = nullptr;
    double *basis = nullptr;  // (m + 1) G
    double2 *spec[4] = {nullptr, nullptr, nullptr, nullptr};
    double *partials = nullptr, *sums = nullptr, *coef = nullptr;  // [3 MAX_CENTRES kRedBlocks], [..], [m + 1]
    int *istar = nullptr;
    double fmu_w = 0.;  // F_mu . w
    int nait = 0, taper = 0, method = 0, hartree = 0;
    int fixed_mu = 0;   // the loop holds mu (RelAcc = 0): no constraint row, P = I
    int verbose = 0;  // MPDPFT_VERBOSE at the adjoint call: GMRES residual per restart on stderr
    std::vector<void *> mem;
};

namespace {

using mk::Grid;
using mk::kRedBlocks;
constexpr int kT = 256;
constexpr double kPI = 3.141592653589793;
constexpr double kMPa = 1.0e-12;  // MP (mpDPFT.cpp:111): RelAcc at or below it holds mu

inline unsigned gblocks(size_t n) { return (unsigned)std::min<size_t>(4096, (n + kT - 1) / kT); }

__device__ __forceinline__ void dcoords(const Grid &g, size_t i, int c[3]) {
    const size_t EL = (size_t)g.EL;
    c[0] = (int)(i / (EL * EL));
    size_t rem = i - (size_t)c[0] * EL * EL;
    c[1] = (int)(rem / EL);
    c[2] = (int)(rem - (size_t)c[1] * EL);
}

__device__ __forceinline__ double trapw3(const Grid &g, const int c[3]) {
    double w = 1.;
    for (int d = 0; d < 3; d++)
        if (c[d] == 0 || c[d] == g.s) w *= 0.5;
    return w;
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[wid] = v;
    __syncthreads();
    double r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    __syncthreads();
    return r;
}

// out = W^T (mask_01 * in): at each base point of [0,s)^3 the sum over its
// periodic images (coordinates 0 duplicated at s); 0 at the images.
// mask (may be null): factor 1 where mask > 0, else 0.
__global__ void k_fold(Grid g, const double *in, const double *mask, double *out) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        if (c[0] == g.s || c[1] == g.s || c[2] == g.s) {
            out[i] = 0.;
            continue;
        }
        const int nx = c[0] == 0 ? 2 : 1, ny = c[1] == 0 ? 2 : 1, nz = c[2] == 0 ? 2 : 1;
        double acc = 0.;
        for (int a = 0; a < nx; a++)
            for (int b = 0; b < ny; b++)
                for (int d = 0; d < nz; d++) {
                    size_t j = ((size_t)(a ? g.s : c[0]) * g.EL + (b ? g.s : c[1])) * g.EL + (d ? g.s : c[2]);
                    double v = in[j];
                    if (mask && !(mask[j] > 0.)) v = 0.;
                    acc += v;
                }
        out[i] = acc;
    }
}

__global__ void k_zero_images(Grid g, double *f) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        if (c[0] == g.s || c[1] == g.s || c[2] == g.s) f[i] = 0.;
    }
}

__global__ void k_prod(size_t n, const double *a, const double *b, double s, double *out) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) out[i] = s * a[i] * b[i];
}

// The transposed derivative multipliers of DelField on the standard-order
// spectra: out = sum_a conj-free transpose of (i k_a f) X_a + (-k^2 f^2) Y with
// f = prod ek (1 without smoothing).  i k_a f is odd in k, so its transpose is
// -i k_a f; the Nyquist plane of axis a carries no first derivative.
__global__ void k_spec_combine(int s, const double *kq, const double *ek, const double2 *X0, const double2 *X1,
                               const double2 *X2, const double2 *Y, double2 *out) {
    const size_t F = (size_t)s * s * s;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < F; i += (size_t)gridDim.x * kT) {
        int q[3];
        q[0] = (int)(i / ((size_t)s * s));
        size_t rem = i - (size_t)q[0] * s * s;
        q[1] = (int)(rem / s);
        q[2] = (int)(rem - (size_t)q[1] * s);
        double f = ek ? ek[q[0]] * ek[q[1]] * ek[q[2]] : 1.;
        double k2 = 0., ore = 0., oim = 0.;
        const double2 *X[3] = {X0, X1, X2};
        for (int a = 0; a < 3; a++) {
            double k = kq[q[a]];
            k2 += k * k;
            if (q[a] != s / 2) {  // (-i k f)(xr + i xi) = k f xi - i k f xr
                double2 v = X[a][i];
                ore += k * f * v.y;
                oim -= k * f * v.x;
            }
        }
        double2 y = Y[i];
        double cl = -k2 * f * f;
        ore += cl * y.x;
        oim += cl * y.y;
        out[i] = double2{ore, oim};
    }
}

// the smallest grid index of the taper shell's point list (ascending) where f equals
// the shell maximum (-*shift); -1 when the shell is empty (the shift is then 0)
__global__ __launch_bounds__(256) void k_shell_argmax(const double *f, const int *idx, int n, const double *shift,
                                                      int *out) {
    __shared__ int sh[256];
    const double sft = -(*shift);
    int best = 0x7fffffff;
    if (!(sft < -1e307 || sft > 1e307))
        for (int j = threadIdx.x; j < n; j += 256)
            if (f[idx[j]] == sft && idx[j] < best) best = idx[j];
    sh[threadIdx.x] = best;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] = min(sh[threadIdx.x], sh[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0] == 0x7fffffff ? -1 : sh[0];
}

// out = T(r) fbar (+= when acc), partial sums of T fbar (the shift's transpose)
__global__ __launch_bounds__(256) void k_taper_adj(Grid g, const double *fbar, double rt, double *out,
                                                   double *partials) {
    __shared__ double sh[4];
    const double a = -2500. / (g.edge * g.edge);
    double acc = 0.;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        double r0 = g.lat[c[0]], r1 = g.lat[c[1]], r2 = g.lat[c[2]];
        double rn = sqrt(r0 * r0 + r1 * r1 + r2 * r2);
        double T = 1.;
        if (rn > rt) {
            double x = a * (rn - rt) * (rn - rt);
            T = x < -700. ? 0. : exp(x);
        }
        double v = T * fbar[i];
        out[i] = v;
        acc += v;
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

__global__ void k_sub_at(double *out, const int *istar, const double *sum) {
    if (threadIdx.x == 0 && *istar >= 0) out[*istar] -= *sum;
}

// x = ell + h + xcd lam (absent terms skipped); partials of F_mu . x = -dfdU . x
__global__ __launch_bounds__(256) void k_adj_x(size_t G, const double *ell, const double *h, const double *xcd,
                                               const double *lam, const double *dfdU, double *x, double *partials) {
    __shared__ double sh[4];
    double acc = 0.;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < G; i += (size_t)gridDim.x * kT) {
        double v = ell ? ell[i] : 0.;
        if (h) v += h[i];
        if (xcd) v += xcd[i] * lam[i];
        x[i] = v;
        acc -= dfdU[i] * v;
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// nu = x - alpha w (w: trapezoid weight dx^3); direct = dfdU nu, mug = dfdg2 nu, mul = dfdl nu
__global__ void k_adj_nu(Grid g, double dxD, const double *x, double alpha, const double *dfdU, const double *dfdg2,
                         const double *dfdl, double *direct, double *mug, double *mul) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        double nu = x[i] - alpha * trapw3(g, c) * dxD;
        direct[i] = dfdU[i] * nu;
        if (mug) mug[i] = dfdg2[i] * nu;
        if (mul) mul[i] = dfdl[i] * nu;
    }
}

// partials of F_mu . w
__global__ __launch_bounds__(256) void k_fmu_w(Grid g, double dxD, const double *dfdU, double *partials) {
    __shared__ double sh[4];
    double acc = 0.;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        acc -= dfdU[i] * trapw3(g, c) * dxD;
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = r;
}

// TF: n = pref (mu - V)_+^{3/2}, dn/dU = -1.5 pref (mu - V)_+^{1/2}
__global__ void k_tf_partials(size_t G, const double *V, double mu, double pref, double *dU) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < G; i += (size_t)gridDim.x * kT) {
        double t = mu - V[i];
        dU[i] = t > 0. ? -1.5 * pref * sqrt(t) : 0.;
    }
}

__global__ void k_axpby(size_t n, double a, const double *x, double b, double *y) {  // y = a x + b y
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) y[i] = a * x[i] + b * y[i];
}

// partials[j kRedBlocks + blk] = <w, basis_j>, j < nh
__global__ __launch_bounds__(256) void k_basis_dots(size_t n, const double *w, const double *basis, int nh,
                                                    double *partials) {
    __shared__ double sh[4];
    for (int j = 0; j < nh; j++) {
        const double *b = basis + (size_t)j * n;
        double acc = 0.;
        for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) acc += w[i] * b[i];
        double r = block_sum(acc, sh);
        if (threadIdx.x == 0) partials[(size_t)j * kRedBlocks + blockIdx.x] = r;
    }
}

// w -= sum_j coef[j] basis_j
__global__ void k_basis_axpy(size_t n, const double *basis, int nh, const double *coef, double *w) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
        double acc = w[i];
        for (int j = 0; j < nh; j++) acc -= coef[j] * basis[(size_t)j * n + i];
        w[i] = acc;
    }
}

// x += sum_j y[j] basis_j
__global__ void k_basis_combine(size_t n, const double *basis, int nh, const double *y, double *x) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < n; i += (size_t)gridDim.x * kT) {
        double acc = x[i];
        for (int j = 0; j < nh; j++) acc += y[j] * basis[(size_t)j * n + i];
        x[i] = acc;
    }
}

// dEnv/dR of the softened Coulomb centres, contracted with lam:
// out[3c + d] = sum_i lam_i (-Z_c)(r_d - R_d) / (|r - R_c|^2 + a_c^2)^{3/2}
__global__ __launch_bounds__(256) void k_centre_grad(Grid g, const double *centres, const double *lam,
                                                     double *partials) {
    __shared__ double sh[4];
    const int cidx = blockIdx.y;
    const double *C = centres + 5 * cidx;
    double acc[3] = {0., 0., 0.};
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        double dr[3], dd = 0.;
        for (int d = 0; d < 3; d++) {
            dr[d] = g.lat[c[d]] - C[d];
            dd += dr[d] * dr[d];
        }
        double q = dd + C[4] * C[4];
        double f = -C[3] * lam[i] / (q * sqrt(q));
        for (int d = 0; d < 3; d++) acc[d] += f * dr[d];
    }
    for (int d = 0; d < 3; d++) {
        double r = block_sum(acc[d], sh);
        if (threadIdx.x == 0) partials[(size_t)(3 * cidx + d) * kRedBlocks + blockIdx.x] = r;
    }
}

// out = w dx^3 n (the explicit part of dE/dEnv)
__global__ void k_weighted(Grid g, double dxD, const double *n, double *out) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        out[i] = trapw3(g, c) * dxD * n[i];
    }
}

// dE/dn of the loop's total energy (explicit n dependence):
// w dx^3 (Env + v_xc + K n / 2) + K^T (w dx^3 n) / 2
__global__ void k_energy_source(Grid g, double dxD, const double *envxc, const double *Kn, const double *KtWn,
                                double *out) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        double W = trapw3(g, c) * dxD;
        double v = W * envxc[i];
        if (Kn) v += 0.5 * (W * Kn[i] + KtWn[i]);
        out[i] = v;
    }
}

// TF dispersal energy pref sum w (mu - V)_+^{5/2}: d/dU = -(5/2) pref (mu - V)_+^{3/2}
__global__ void k_tf_energy_dU(size_t G, const double *V, double mu, double pref, double *out) {
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < G; i += (size_t)gridDim.x * kT) {
        double t = mu - V[i];
        out[i] = t > 0. ? -2.5 * pref * t * sqrt(t) : 0.;
    }
}

// out = s w dx^3 a (a may alias out); partial sums of out
__global__ __launch_bounds__(256) void k_weight_scale(Grid g, double dxD, double s, const double *a, double *out,
                                                      double *partials) {
    __shared__ double sh[4];
    double acc = 0.;
    for (size_t i = blockIdx.x * (size_t)kT + threadIdx.x; i < g.G; i += (size_t)gridDim.x * kT) {
        int c[3];
        dcoords(g, i, c);
        double v = s * trapw3(g, c) * dxD * a[i];
        out[i] = v;
        acc += v;
    }
    double r = block_sum(acc, sh);
    if (partials && threadIdx.x == 0) partials[blockIdx.x] = r;
}

}  // namespace

void adjoint_free(mpdpft_ctx *c) {
    if (!c->adj) return;
    for (void *q : c->adj->mem) (void)hipFree(q);
    delete c->adj;
    c->adj = nullptr;
}

namespace {

template <class T> int amalloc(mpdpft_ctx *c, T **p, size_t n) {
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e != hipSuccess) return hip_fail(c, e, "hipMalloc (adjoint)");
    c->adj->mem.push_back((void *)*p);
    return MPDPFT_OK;
}

int reduce_to_host(mpdpft_ctx *c, const double *partials, int nres, double *out) {
    AdjointState &A = *c->adj;
    HIPC(mk::finish_sums(partials, nres, A.sums, 1., c->st));
    HIPC(hipMemcpyAsync(c->h_scal, A.sums, sizeof(double) * nres, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    for (int k = 0; k < nres; k++) out[k] = c->h_scal[k];
    return MPDPFT_OK;
}

double dxD(const mpdpft_ctx *c) { return std::pow(c->dx, (double)c->D); }

// Vbar = (DelField chain)^T applied to (mug on |grad V|^2, mul on lap V), into out
int delfield_adjoint(mpdpft_ctx *c, const double *mug, const double *mul, double *out) {
    AdjointState &A = *c->adj;
    const Grid &g = c->g;
    const size_t G = c->G;
    const double scale = 1. / (double)c->F;
    const unsigned nb = gblocks(G);
    mk::MulParams m{};
    m.kind = mk::MUL_SMOOTH;
    m.kq = c->d_kq;
    m.ek = A.method == 5 ? c->d_ek : nullptr;
    m.sigma = A.method == 5 ? c->sigma : 0.;
    // qbar: the adjoint of |grad V|^2 = POS(W C_s R sum_a D_a^2) (method 5), or of sum_a D_a^2 (method 4)
    double *qbar = A.t2;
    if (A.method == 5) {
        hipLaunchKernelGGL(k_fold, dim3(nb), dim3(kT), 0, c->st, g, mug, (const double *)A.g2, A.t1);
        HIPC(mk::fft_forward_real(g, A.t1, nullptr, A.spec[3], c->d_tw, nullptr, c->st));
        HIPC(mk::fft_backward_real(g, A.spec[3], A.spec[3], m, qbar, nullptr, scale, 0, 0., c->d_tw, c->st));
        hipLaunchKernelGGL(k_zero_images, dim3(nb), dim3(kT), 0, c->st, g, qbar);
    } else {
        HIPC(hipMemcpyAsync(qbar, mug, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
    }
    // D_a = W C_a R f: Dbar_a = 2 D_a qbar, folded, transformed
    const double *Ds[3] = {A.Dx, A.Dy, A.Dz};
    for (int a = 0; a < 3; a++) {
        hipLaunchKernelGGL(k_prod, dim3(nb), dim3(kT), 0, c->st, G, Ds[a], (const double *)qbar, 2., A.t1);
        hipLaunchKernelGGL(k_fold, dim3(nb), dim3(kT), 0, c->st, g, (const double *)A.t1, (const double *)nullptr,
                           A.t3);
        HIPC(mk::fft_forward_real(g, A.t3, nullptr, A.spec[a], c->d_tw, nullptr, c->st));
    }
    hipLaunchKernelGGL(k_fold, dim3(nb), dim3(kT), 0, c->st, g, mul, (const double *)nullptr, A.t3);
    HIPC(mk::fft_forward_real(g, A.t3, nullptr, A.spec[3], c->d_tw, nullptr, c->st));
    hipLaunchKernelGGL(k_spec_combine, dim3(gblocks(c->F)), dim3(kT), 0, c->st, c->s, (const double *)c->d_kq,
                       (const double *)m.ek, (const double2 *)A.spec[0], (const double2 *)A.spec[1],
                       (const double2 *)A.spec[2], (const double2 *)A.spec[3], c->d_spec);
    mk::MulParams mn{};
    mn.kind = mk::MUL_NONE;
    mn.kq = c->d_kq;
    HIPC(mk::fft_backward_real(g, c->d_spec, c->d_work, mn, A.fbar, nullptr, scale, 0, 0., c->d_tw, c->st));
    hipLaunchKernelGGL(k_zero_images, dim3(nb), dim3(kT), 0, c->st, g, A.fbar);
    if (A.taper) {  // f = T (V - V[i*]): Vbar = T fbar - e_{i*} sum T fbar
        double rt = c->p.edge * c->p.RegularizationThreshold;
        hipLaunchKernelGGL(k_taper_adj, dim3(kRedBlocks), dim3(kT), 0, c->st, g, (const double *)A.fbar, rt, out,
                           A.partials);
        HIPC(mk::finish_sums(A.partials, 1, A.sums, 1., c->st));
        hipLaunchKernelGGL(k_sub_at, dim3(1), dim3(64), 0, c->st, out, (const int *)A.istar, (const double *)A.sums);
    } else {
        HIPC(hipMemcpyAsync(out, A.fbar, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
    }
    HIPC(hipGetLastError());
    return MPDPFT_OK;
}

// fx = F_mu . (ell + K'^T lam), leaving x = ell + K'^T lam in A.x; lam or ell may be null
int fmu_dot(mpdpft_ctx *c, const double *lam, const double *ell, double *fx) {
    AdjointState &A = *c->adj;
    const Grid &g = c->g;
    const size_t G = c->G;
    const double *h = nullptr;
    if (lam && A.hartree) {
        const int M = 2 * c->s;
        double tot = std::pow((double)M, (double)c->D);
        HIPC(mk::hartree_isolated(g, lam, c->d_pad, c->d_khat, c->d_tw2, c->p.gammaH * dxD(c) / tot, A.h, 0, c->st,
                                  1));
        h = A.h;
    }
    hipLaunchKernelGGL(k_adj_x, dim3(kRedBlocks), dim3(kT), 0, c->st, G, ell, h, lam ? (const double *)A.xcd : nullptr,
                       lam, (const double *)A.dfdU, A.x, A.partials);
    HIPC(hipGetLastError());
    return reduce_to_host(c, A.partials, 1, fx);
}

// out = F_V^T (P (ell + K'^T lam) - w alpha_add); lam or ell may be null (term dropped)
int apply_adjoint_map(mpdpft_ctx *c, const double *lam, const double *ell, double *out, double alpha_add = 0.) {
    AdjointState &A = *c->adj;
    const Grid &g = c->g;
    const size_t G = c->G;
    const unsigned nb = gblocks(G);
    double fx;
    RC(fmu_dot(c, lam, ell, &fx));
    const double alpha = A.fixed_mu ? 0. : fx / A.fmu_w + alpha_add;
    hipLaunchKernelGGL(k_adj_nu, dim3(nb), dim3(kT), 0, c->st, g, dxD(c), (const double *)A.x, alpha,
                       (const double *)A.dfdU, (const double *)A.dfdg2, (const double *)A.dfdl, A.direct,
                       A.nait ? A.mug : nullptr, A.nait ? A.mul : nullptr);
    if (A.nait) {
        RC(delfield_adjoint(c, A.mug, A.mul, out));
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, G, 1., (const double *)A.direct, 1., out);
    } else {
        HIPC(hipMemcpyAsync(out, A.direct, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
    }
    HIPC(hipGetLastError());
    return MPDPFT_OK;
}

// y = (I - F_V^T P K'^T) v
int apply_system(mpdpft_ctx *c, const double *v, double *y) {
    RC(apply_adjoint_map(c, v, nullptr, y));
    hipLaunchKernelGGL(k_axpby, dim3(gblocks(c->G)), dim3(kT), 0, c->st, c->G, 1., v, -1., y);
    HIPC(hipGetLastError());
    return MPDPFT_OK;
}

int dot(mpdpft_ctx *c, const double *a, const double *b, double *out) {
    AdjointState &A = *c->adj;
    hipLaunchKernelGGL(k_basis_dots, dim3(kRedBlocks), dim3(kT), 0, c->st, c->G, a, b, 1, A.partials);
    HIPC(hipGetLastError());
    return reduce_to_host(c, A.partials, 1, out);
}

// Restarted GMRES(m) for (I - F_V^T P K'^T) x = b from x = 0; classical
// Gram-Schmidt with one reorthogonalisation, Givens rotations on the host.
int gmres(mpdpft_ctx *c, const double *b, double *x, double tol, int maxit, int *its, double *relres) {
    AdjointState &A = *c->adj;
    const size_t G = c->G;
    const int m = A.m;
    const unsigned nb = gblocks(G);
    HIPC(hipMemsetAsync(x, 0, sizeof(double) * G, c->st));
    double bn;
    RC(dot(c, b, b, &bn));
    bn = std::sqrt(bn);
    *its = 0;
    *relres = 0.;
    if (bn == 0.) return MPDPFT_OK;
    std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), e(m + 1), y(m), hcol(m + 1), h2(m + 1);
    double *r = A.res;
    int total = 0;
    for (;;) {
        // r = b - (I - A) x
        RC(apply_system(c, x, r));
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, G, 1., b, -1., r);
        double rn;
        RC(dot(c, r, r, &rn));
        rn = std::sqrt(rn);
        *relres = rn / bn;
        if (A.verbose) {
            std::fprintf(stderr, " ***** adjoint GMRES: %d iterations, relative residual %.3e *****\n", total, *relres);
            std::fflush(stderr);
        }
        if (*relres <= tol || total >= maxit) break;
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, G, 1. / rn, (const double *)r, 0., A.basis);
        std::fill(e.begin(), e.end(), 0.);
        e[0] = rn;
        int k = 0;
        for (; k < m && total < maxit; k++, total++) {
            double *vk = A.basis + (size_t)k * G, *w = A.basis + (size_t)(k + 1) * G;
            RC(apply_system(c, vk, w));
            // two passes of classical Gram-Schmidt against v_0..v_k
            std::fill(hcol.begin(), hcol.end(), 0.);
            for (int pass = 0; pass < 2; pass++) {
                hipLaunchKernelGGL(k_basis_dots, dim3(kRedBlocks), dim3(kT), 0, c->st, G, (const double *)w,
                                   (const double *)A.basis, k + 1, A.partials);
                RC(reduce_to_host(c, A.partials, k + 1, h2.data()));
                HIPC(hipMemcpyAsync(A.coef, h2.data(), sizeof(double) * (k + 1), hipMemcpyHostToDevice, c->st));
                hipLaunchKernelGGL(k_basis_axpy, dim3(nb), dim3(kT), 0, c->st, G, (const double *)A.basis, k + 1,
                                   (const double *)A.coef, w);
                for (int j = 0; j <= k; j++) hcol[j] += h2[j];
            }
            double wn;
            RC(dot(c, w, w, &wn));
            wn = std::sqrt(wn);
            hcol[k + 1] = wn;
            if (wn > 0.) hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, G, 1. / wn, (const double *)w, 0., w);
            // apply the previous rotations, then a new one
            for (int j = 0; j < k; j++) {
                double t = cs[j] * hcol[j] + sn[j] * hcol[j + 1];
                hcol[j + 1] = -sn[j] * hcol[j] + cs[j] * hcol[j + 1];
                hcol[j] = t;
            }
            double den = std::hypot(hcol[k], hcol[k + 1]);
            cs[k] = den > 0. ? hcol[k] / den : 1.;
            sn[k] = den > 0. ? hcol[k + 1] / den : 0.;
            hcol[k] = den;
            hcol[k + 1] = 0.;
            e[k + 1] = -sn[k] * e[k];
            e[k] = cs[k] * e[k];
            for (int j = 0; j <= k; j++) H[(size_t)j * m + k] = hcol[j];
            if (std::fabs(e[k + 1]) / bn <= tol) {
                k++;
                total++;
                break;
            }
        }
        // x += V_k y, H y = e (upper triangular k x k)
        for (int i = k - 1; i >= 0; i--) {
            double acc = e[i];
            for (int j = i + 1; j < k; j++) acc -= H[(size_t)i * m + j] * y[j];
            y[i] = acc / H[(size_t)i * m + i];
        }
        HIPC(hipMemcpyAsync(A.coef, y.data(), sizeof(double) * k, hipMemcpyHostToDevice, c->st));
        hipLaunchKernelGGL(k_basis_combine, dim3(nb), dim3(kT), 0, c->st, G, (const double *)A.basis, k,
                           (const double *)A.coef, x);
        HIPC(hipGetLastError());
    }
    *its = total;
    return MPDPFT_OK;
}

int check_supported(mpdpft_ctx *c) {
    const mpdpft_params &p = c->p;
    if (c->D != 3 || c->S != 1) return ctx_fail(c, MPDPFT_EINVAL, "adjoint: DIM 3 with one species");
    if (p.DensityExpression != 5 && p.DensityExpression != 7)
        return ctx_fail(c, MPDPFT_EINVAL, "adjoint: DensityExpression 5 (nAiT) or 7 (TF)");
    if (p.DensityExpression == 5 && p.DelFieldMethod != 4 && p.DelFieldMethod != 5)
        return ctx_fail(c, MPDPFT_EINVAL, "adjoint: DelFieldMethod 4 or 5");
    for (int a = 0; a < p.n_interactions; a++) {
        int t = p.Interactions[a];
        if (t == 8 && p.HartreeMode != 1) return ctx_fail(c, MPDPFT_EINVAL, "adjoint: isolated Hartree only");
        if (t != 0 && t != 8 && t != 10 && t != 11 && t != -1 && t != -12)
            return ctx_fail(c, MPDPFT_EINVAL, "adjoint: interactions 8, 10, 11, -1, -12");
    }
    if (c->incV < 1.) return ctx_fail(c, MPDPFT_EINVAL, "adjoint: incrementalV ramp not finished");
    if (c->SCcount < 1) return ctx_fail(c, MPDPFT_EINVAL, "adjoint: no converged solve in this context");
    return MPDPFT_OK;
}

// Linearisation point: the converged V, mu; partial derivatives of the
// density, the interaction diagonal, the DelField fields and the shift index.
int setup(mpdpft_ctx *c, int restart) {
    if (!c->adj) {
        c->adj = new AdjointState();
        AdjointState &A = *c->adj;
        A.G = c->G;
        A.F = c->F;
        A.m = restart;
        const size_t G = c->G;
        double **fields[] = {&A.dfdU, &A.dfdg2, &A.dfdl, &A.xcd, &A.nbuf, &A.Dx, &A.Dy, &A.Dz, &A.g2, &A.ell,
                             &A.h, &A.x, &A.direct, &A.mug, &A.mul, &A.t1, &A.t2, &A.t3, &A.fbar, &A.w, &A.rhs,
                             &A.res};
        for (double **f : fields) RC(amalloc(c, f, G));
        RC(amalloc(c, &A.basis, (size_t)(restart + 1) * G));
        for (int k = 0; k < 4; k++) RC(amalloc(c, &A.spec[k], c->F));
        RC(amalloc(c, &A.partials, (size_t)3 * MPDPFT_MAX_CENTRES * kRedBlocks));
        RC(amalloc(c, &A.sums, (size_t)3 * MPDPFT_MAX_CENTRES));
        RC(amalloc(c, &A.coef, (size_t)restart + 1));
        RC(amalloc(c, &A.istar, 1));
    }
    AdjointState &A = *c->adj;
    const mpdpft_params &p = c->p;
    const size_t G = c->G;
    const unsigned nb = gblocks(G);
    const double mu = c->muVec[0];
    A.nait = p.DensityExpression == 5;
    A.method = p.DelFieldMethod;
    A.taper = A.nait && p.TaperOff && (A.method == 4 || A.method == 5);
    A.hartree = c->has_hartree;
    A.fixed_mu = !(p.RelAcc > kMPa && p.TaskType < 100);
    if (A.nait) {
        c->delfield_valid = -1;  // DelField(V) now: its fields and shell maximum are the linearisation's
        RC(nait_partials(c, 0, mu, A.nbuf, A.dfdU, A.dfdg2, A.dfdl));
        // the DelField fields of V (nait_partials leaves DelField(V) in the context)
        HIPC(hipMemcpyAsync(A.Dx, c->d_Dx, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
        HIPC(hipMemcpyAsync(A.Dy, c->d_Dy, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
        HIPC(hipMemcpyAsync(A.Dz, c->d_Dz, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
        HIPC(hipMemcpyAsync(A.g2, c->d_Grad2, sizeof(double) * G, hipMemcpyDeviceToDevice, c->st));
        if (A.taper) {  // the forward's shell points and maximum (left by nait_partials' DelField of V)
            if (c->n_shell < 0) return ctx_fail(c, MPDPFT_EINVAL, "adjoint: taper geometry missing");
            hipLaunchKernelGGL(k_shell_argmax, dim3(1), dim3(256), 0, c->st, (const double *)c->d_V,
                               (const int *)c->d_shell_idx, c->n_shell, (const double *)(c->d_scal + 210), A.istar);
        }
    } else {
        double g = p.degeneracy;
        double pref = (p.Units == 2 ? 0.0475412512 : 1.) * std::sqrt(2.) * g / (3. * kPI * kPI);
        hipLaunchKernelGGL(k_tf_partials, dim3(nb), dim3(kT), 0, c->st, G, (const double *)c->d_V, mu, pref, A.dfdU);
        HIPC(hipMemsetAsync(A.dfdg2, 0, sizeof(double) * G, c->st));
        HIPC(hipMemsetAsync(A.dfdl, 0, sizeof(double) * G, c->st));
    }
    mk::VintParams vp{};
    vp.n_inter = p.n_interactions;
    for (int a = 0; a < p.n_interactions; a++) vp.inter[a] = p.Interactions[a];
    vp.S = 1;
    vp.s = c->s;
    vp.beta = p.beta;
    vp.gammaH = p.gammaH;
    vp.inc = 1.;
    vp.units = p.Units;
    HIPC(mk::vint_dvdn(c->g, vp, c->d_Den, A.xcd, c->st));
    hipLaunchKernelGGL(k_fmu_w, dim3(kRedBlocks), dim3(kT), 0, c->st, c->g, dxD(c), (const double *)A.dfdU,
                       A.partials);
    HIPC(hipGetLastError());
    RC(reduce_to_host(c, A.partials, 1, &A.fmu_w));
    if (!(A.fmu_w > 0.)) return ctx_fail(c, MPDPFT_EINVAL, "adjoint: dN/dmu is not positive");
    return MPDPFT_OK;
}

int centre_contract(mpdpft_ctx *c, const double *lam_dev, double *out) {
    AdjointState &A = *c->adj;
    const int nc = c->p.n_centres;
    hipLaunchKernelGGL(k_centre_grad, dim3(kRedBlocks, nc), dim3(kT), 0, c->st, c->g, (const double *)c->d_centres,
                       lam_dev, A.partials);
    HIPC(hipGetLastError());
    HIPC(mk::finish_sums(A.partials, 3 * nc, A.sums, 1., c->st));
    HIPC(hipMemcpyAsync(out, A.sums, sizeof(double) * 3 * nc, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    return MPDPFT_OK;
}

int centres_ok(mpdpft_ctx *c) {
    if (c->D != 3 || c->S != 1 || c->p.Environments[0] != MPDPFT_ENV_SOFT_COULOMB_CENTRES || c->p.n_centres < 1)
        return ctx_fail(c, MPDPFT_EINVAL, "centre derivatives need environment 41 with centres (3D, one species)");
    return MPDPFT_OK;
}

}  // namespace

extern "C" int mpdpft_adjoint(mpdpft_ctx *c, const double *dL_dn, double *lambda, mpdpft_adjoint_info *info) {
    if (!c || !dL_dn || !lambda) return MPDPFT_EINVAL;
    HIPC(hipSetDevice(c->device));
    RC(check_supported(c));
    const int restart = 40;
    RC(setup(c, restart));
    AdjointState &A = *c->adj;
    const size_t G = c->G;
    const char *vb = std::getenv("MPDPFT_VERBOSE");
    A.verbose = vb && *vb && *vb != '0';
    double tol = 1e-12;
    int maxit = 400;
    if (info && info->tol > 0.) tol = info->tol;
    if (info && info->max_iterations > 0) maxit = info->max_iterations;
    HIPC(hipMemcpyAsync(A.ell, dL_dn, sizeof(double) * G, hipMemcpyHostToDevice, c->st));
    RC(apply_adjoint_map(c, nullptr, A.ell, A.rhs));  // b = F_V^T P l
    int its = 0;
    double rr = 0.;
    RC(gmres(c, A.rhs, A.w, tol, maxit, &its, &rr));
    HIPC(hipMemcpyAsync(lambda, A.w, sizeof(double) * G, hipMemcpyDeviceToHost, c->st));
    HIPC(hipStreamSynchronize(c->st));
    if (info) {
        info->iterations = its;
        info->relative_residual = rr;
        info->dN_dmu = A.fmu_w;
    }
    if (rr > tol) return ctx_fail(c, MPDPFT_ENOTCONV, "adjoint: GMRES did not reach the tolerance");
    return MPDPFT_OK;
}

extern "C" int mpdpft_centre_gradient(mpdpft_ctx *c, const double *lambda, double *out) {
    if (!c || !lambda || !out) return MPDPFT_EINVAL;
    HIPC(hipSetDevice(c->device));
    RC(centres_ok(c));
    if (!c->adj) RC(setup(c, 40));
    HIPC(hipMemcpyAsync(c->adj->t3, lambda, sizeof(double) * c->G, hipMemcpyHostToDevice, c->st));
    return centre_contract(c, c->adj->t3, out);
}

namespace {

// The total energy's derivative: E(V, mu, n = F(V, mu); theta) of GetEnergy
// (:5414-5436) differentiated through the fixed point.  Sources: dE/dn =
// w (Env + v_xc + K n / 2) + K^T (w n) / 2 (Hartree, Dirac, LDA energies are
// exact antiderivatives of their potentials); k_V = dE_disp/dV through U and
// the DelField terms (the kinetic energy's own quadrature, differentiated with
// its refinements held fixed); k_mu = dE_disp/dmu.  Leaves lambda_E in A.w and
// A.ell = dE/dn; *kmu_out = k_mu.  Then dE/dtheta = sum_i (w n + lambda_E)_i
// dEnv_i/dtheta, and the constraint's multiplier gives dE/dN (below).
int energy_adjoint(mpdpft_ctx *c, double *kmu_out) {
    HIPC(hipSetDevice(c->device));
    RC(check_supported(c));
    const mpdpft_params &p = c->p;
    for (int e = 0; e < p.n_ekin; e++)
        if (p.EkinTypes[e] != 0) return ctx_fail(c, MPDPFT_EINVAL, "energy derivatives: EkinTypes 0 only");
    RC(setup(c, 40));
    AdjointState &A = *c->adj;
    const Grid &g = c->g;
    const size_t G = c->G;
    const unsigned nb = gblocks(G);
    const double dV = dxD(c);
    const char *vb = std::getenv("MPDPFT_VERBOSE");
    A.verbose = vb && *vb && *vb != '0';
    // Env + v_xc (pointwise interactions; V of UpdateV without Hartree)
    mk::VintParams vp{};
    for (int a = 0; a < p.n_interactions; a++)
        if (p.Interactions[a] != 8) vp.inter[vp.n_inter++] = p.Interactions[a];
    vp.S = 1;
    vp.s = c->s;
    vp.beta = p.beta;
    vp.gammaH = p.gammaH;
    vp.inc = 1.;
    vp.units = p.Units;
    HIPC(mk::update_v(g, vp, c->d_Den, c->d_Env, A.t1, c->st));
    const double *Kn = nullptr, *KtWn = nullptr;
    if (A.hartree) {
        const int M = 2 * c->s;
        double sc = p.gammaH * dV / std::pow((double)M, (double)c->D);
        HIPC(mk::hartree_isolated(g, c->d_Den, c->d_pad, c->d_khat, c->d_tw2, sc, A.t2, 0, c->st, 0));
        hipLaunchKernelGGL(k_weighted, dim3(nb), dim3(kT), 0, c->st, g, dV, (const double *)c->d_Den, A.fbar);
        HIPC(mk::hartree_isolated(g, A.fbar, c->d_pad, c->d_khat, c->d_tw2, sc, A.t3, 0, c->st, 1));
        Kn = A.t2;
        KtWn = A.t3;
    }
    hipLaunchKernelGGL(k_energy_source, dim3(nb), dim3(kT), 0, c->st, g, dV, (const double *)A.t1, Kn, KtWn, A.ell);
    HIPC(hipGetLastError());
    // kinetic (dispersal) energy sources: kV into A.w, k_mu
    double kmu = 0.;
    if (A.nait) {
        double pref;
        RC(nait_energy_partials(c, 0, A.direct, A.mug, A.mul, &pref));
        hipLaunchKernelGGL(k_weight_scale, dim3(kRedBlocks), dim3(kT), 0, c->st, g, dV, pref, (const double *)A.direct,
                           A.direct, A.partials);
        double sU;
        RC(reduce_to_host(c, A.partials, 1, &sU));
        kmu = -sU;
        hipLaunchKernelGGL(k_weight_scale, dim3(kRedBlocks), dim3(kT), 0, c->st, g, dV, pref, (const double *)A.mug,
                           A.mug, (double *)nullptr);
        hipLaunchKernelGGL(k_weight_scale, dim3(kRedBlocks), dim3(kT), 0, c->st, g, dV, pref, (const double *)A.mul,
                           A.mul, (double *)nullptr);
        RC(delfield_adjoint(c, A.mug, A.mul, A.w));
        hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, G, 1., (const double *)A.direct, 1., A.w);
    } else {
        double U = p.Units == 2 ? 7.619964158443410 : 1.;
        double pref = p.degeneracy * 4. * kPI * std::pow(2., 2.5) / (std::pow(2. * kPI * std::sqrt(U), 3.) * 10.);
        hipLaunchKernelGGL(k_tf_energy_dU, dim3(nb), dim3(kT), 0, c->st, G, (const double *)c->d_V, c->muVec[0], pref,
                           A.direct);
        hipLaunchKernelGGL(k_weight_scale, dim3(kRedBlocks), dim3(kT), 0, c->st, g, dV, 1., (const double *)A.direct,
                           A.w, A.partials);
        double sU;
        RC(reduce_to_host(c, A.partials, 1, &sU));
        kmu = -sU;
    }
    HIPC(hipGetLastError());
    // rhs = F_V^T (P l_E - w k_mu / F_mu.w) + k_V
    RC(apply_adjoint_map(c, nullptr, A.ell, A.rhs, A.fixed_mu ? 0. : kmu / A.fmu_w));
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, G, 1., (const double *)A.w, 1., A.rhs);
    int its = 0;
    double rr = 0.;
    RC(gmres(c, A.rhs, A.w, 1e-12, 400, &its, &rr));
    if (rr > 1e-12) return ctx_fail(c, MPDPFT_ENOTCONV, "energy derivatives: adjoint GMRES did not reach 1e-12");
    *kmu_out = kmu;
    return MPDPFT_OK;
}

}  // namespace

extern "C" int mpdpft_centre_forces(mpdpft_ctx *c, double *out) {
    if (!c || !out) return MPDPFT_EINVAL;
    HIPC(hipSetDevice(c->device));
    RC(centres_ok(c));
    double kmu;
    RC(energy_adjoint(c, &kmu));
    AdjointState &A = *c->adj;
    const unsigned nb = gblocks(c->G);
    // lambda_E + w n
    hipLaunchKernelGGL(k_weighted, dim3(nb), dim3(kT), 0, c->st, c->g, dxD(c), (const double *)c->d_Den, A.t3);
    hipLaunchKernelGGL(k_axpby, dim3(nb), dim3(kT), 0, c->st, c->G, 1., (const double *)A.w, 1., A.t3);
    HIPC(hipGetLastError());
    return centre_contract(c, A.t3, out);
}

// With N held, the Lagrangian E + a.R_V + b (w.n - N) is stationary in V and mu
// at a = -lambda_E, b = -(k_mu + F_mu.(dE/dn + K'^T lambda_E)) / F_mu.w, and
// dE/dN = -b.  With mu held (RelAcc = 0) there is no b, and
// dE/dmu = k_mu + F_mu.(dE/dn + K'^T lambda_E).
extern "C" int mpdpft_task_gradient(mpdpft_ctx *c, double *out) {
    if (!c || !out) return MPDPFT_EINVAL;
    HIPC(hipSetDevice(c->device));
    double kmu, fx;
    RC(energy_adjoint(c, &kmu));
    AdjointState &A = *c->adj;
    RC(fmu_dot(c, A.w, A.ell, &fx));
    out[0] = A.fixed_mu ? kmu + fx : (kmu + fx) / A.fmu_w;
    return MPDPFT_OK;
}

extern "C" int mpdpft_set_centres(mpdpft_ctx *c, int n, const double *centres) {
    if (!c || !centres || n < 1 || n > MPDPFT_MAX_CENTRES) return MPDPFT_EINVAL;
    HIPC(hipSetDevice(c->device));
    c->p.n_centres = n;
    std::memcpy(c->p.centres, centres, sizeof(double) * 5 * n);
    HIPC(hipMemcpyAsync(c->d_centres, centres, sizeof(double) * 5 * n, hipMemcpyHostToDevice, c->st));
    for (int s = 0; s < c->S; s++) {
        if (c->p.Environments[s] != MPDPFT_ENV_SOFT_COULOMB_CENTRES) continue;
        mk::EnvParams e{c->p.Environments[s], 0., 0., n, c->d_centres};
        HIPC(mk::environment(c->g, e, c->d_Env + s * c->G, c->st));
    }
    HIPC(hipStreamSynchronize(c->st));
    return MPDPFT_OK;
}
