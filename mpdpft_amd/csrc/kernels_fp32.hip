// kernels_fp32.hip -- FP32 transforms of the precision variant (Precision 1).
//
// DelField's transforms and the isolated Hartree convolution with FP32
// spectra, FP32 twiddles, FP32 multipliers and FP32 accumulation in every
// radix stage; the fields they read and write stay FP64 in HBM (converted on
// load and store), so the loop's reductions, mu search, Pulay solve and
// energies keep their FP64 inputs.  Same pass structure as the FP64 path
// (fft_core.h with float2), 3D only.
#include <hip/hip_runtime.h>

#include "fft_core.h"
#include "kernels.h"

namespace mk {
namespace {

using fftk::launch_pass;

// DelField input: V, or (V - max) T with the taper table T (taper_factor)
struct F32LdTaper {
    const double *re, *T, *shift;
    int EL;
    __device__ float2 operator()(int x, int y, int z) const {
        int i = (x * EL + y) * EL + z;
        double v = re[i];
        if (T) {
            double sft = -(*shift);
            if (sft < -1e307 || sft > 1e307) sft = 0.;
            v = (v - sft) * T[i];
        }
        return float2{(float)v, 0.f};
    }
};
// |grad V|^2 = Dx^2 + Dy^2 + Dz^2 in FP32
struct F32LdSumSq {
    const double *a, *b, *c;
    int EL;
    __device__ float2 operator()(int x, int y, int z) const {
        int i = (x * EL + y) * EL + z;
        float u = (float)a[i], v = (float)b[i], w = (float)c[i];
        return float2{u * u + v * v + w * w, 0.f};
    }
};
struct F32LdBuf {
    const float2 *p;
    int n1, n2;
    __device__ float2 operator()(int x, int y, int z) const { return p[((size_t)x * n1 + y) * n2 + z]; }
};
struct F32StBuf {
    float2 *p;
    int n1, n2;
    __device__ void operator()(int x, int y, int z, float2 v) const { p[((size_t)x * n1 + y) * n2 + z] = v; }
};
// DelField's k-space multipliers (LdMulD of the FP64 path): kek[q] = (k, smoothing factor)
template <int KIND> struct F32LdMul {
    const float2 *p, *kek;
    int n1, n2, s, axis_a, axis_b;
    __device__ float2 operator()(int x, int y, int z) const {
        float2 v = p[((size_t)x * n1 + y) * n2 + z];
        const int q[3] = {x, y, z};
        float k[3], f = 1.f, k2 = 0.f;
#pragma unroll
        for (int d = 0; d < 3; d++) {
            float2 t = kek[q[d]];
            k[d] = t.x;
            k2 += t.x * t.x;
            f *= t.y;
        }
        if (KIND == MUL_SMOOTH) return float2{v.x * f, v.y * f};
        if (KIND == MUL_NONE) return v;
        float2 A = {0.f, 0.f}, B = {0.f, 0.f};
        if (axis_a >= 0 && q[axis_a] != s / 2) {
            float c = k[axis_a] * f;
            A = float2{-c * v.y, c * v.x};
        }
        if (KIND == MUL_DERIV_PAIR) {
            if (q[axis_b] != s / 2) {
                float c = k[axis_b] * f;
                B = float2{-c * v.y, c * v.x};
            }
        } else {  // MUL_DERIV_LAP
            float c = -k2 * f * f;
            B = float2{c * v.x, c * v.y};
        }
        return float2{A.x - B.y, A.y + B.x};
    }
};
// real and imaginary parts onto the (s+1)^3 grid (FP64 fields) with the periodic images
struct F32StGrid {
    double *re, *im;
    float scale;
    int post;  // 1: POS of re
    int EL, s;
    __device__ void put(int i, float2 v) const {
        if (re) {
            float a = scale * v.x;
            if (post == 1 && !(a > 0.f)) a = 0.f;
            re[i] = (double)a;
        }
        if (im) im[i] = (double)(scale * v.y);
    }
    __device__ void operator()(int x, int y, int z, float2 v) const {
        const int i0 = (x * EL + y) * EL + z, sy = s * EL;
        put(i0, v);
        if (z == 0) put(i0 + s, v);
        if (y == 0) {
            put(i0 + sy, v);
            if (z == 0) put(i0 + sy + s, v);
        }
        if (x == 0) {
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
// isolated Hartree: zero-padded loads, kernel multiply, [0,s]^3 store
struct F32LdPadZ {
    const double *n;
    int EL, lim;
    __device__ float2 operator()(int x, int y, int z) const {
        float2 v = {0.f, 0.f};
        if (z < lim) v.x = (float)n[(x * EL + y) * EL + z];
        return v;
    }
};
struct F32LdBufPad {
    const float2 *p;
    int n1, n2, axis, s;
    __device__ float2 operator()(int x, int y, int z) const {
        int c = axis == 0 ? x : (axis == 1 ? y : z);
        if (c >= s) return float2{0.f, 0.f};
        return p[((size_t)x * n1 + y) * n2 + z];
    }
};
struct F32StMulK {
    float2 *p;
    const float *khat;
    int n1, n2;
    __device__ void operator()(int x, int y, int z, float2 v) const {
        size_t i = ((size_t)x * n1 + y) * n2 + z;
        float k = khat[i];
        p[i] = float2{v.x * k, v.y * k};
    }
};
struct F32StIso {
    double *out;
    float scale;
    int EL, s;
    __device__ void operator()(int x, int y, int z, float2 v) const {
        if (z > s) return;
        out[(x * EL + y) * EL + z] = (double)(scale * v.x);
    }
};

template <int AXIS, class Ld, class St>
hipError_t pass32(int N, Ld ld, St st, int L0, int L1, int sign, const float2 *tw, hipStream_t stream) {
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

#define CHK(x)                           \
    do {                                 \
        hipError_t e_ = (x);             \
        if (e_ != hipSuccess) return e_; \
    } while (0)

template <class Ld>
hipError_t forward3(const Grid &g, Ld ld, float2 *spec, const float2 *tw, hipStream_t st) {
    int s = g.s;
    F32StBuf sb{spec, s, s};
    F32LdBuf lb{spec, s, s};
    CHK(pass32<2>(s, ld, sb, s, s, -1, tw, st));
    CHK(pass32<1>(s, lb, sb, s, s, -1, tw, st));
    CHK(pass32<0>(s, lb, sb, s, s, -1, tw, st));
    return hipSuccess;
}

__global__ void k_to_float(size_t n, const double *a, float *b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}
__global__ void k_to_float2(size_t n, const double2 *a, float2 *b) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = float2{(float)a[i].x, (float)a[i].y};
}

}  // namespace

hipError_t fft32_forward(const Grid &g, const double *V, const double *T, const double *shift_dev, float2 *spec,
                         const float2 *tw, hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    return forward3(g, F32LdTaper{V, T, shift_dev, g.EL}, spec, tw, st);
}

hipError_t fft32_forward_sumsq(const Grid &g, const double *a, const double *b, const double *c, float2 *spec,
                               const float2 *tw, hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    return forward3(g, F32LdSumSq{a, b, c, g.EL}, spec, tw, st);
}

hipError_t fft32_backward(const Grid &g, const float2 *spec, float2 *work, int kind, int axis_a, int axis_b,
                          const float2 *kek, double *re, double *im, float scale, int post, const float2 *tw,
                          hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    int s = g.s;
    F32StBuf sb{work, s, s};
    F32LdBuf lb{work, s, s};
    F32StGrid sg{re, im, scale, post, g.EL, s};
    switch (kind) {
    case MUL_DERIV_PAIR:
        CHK(pass32<0>(s, F32LdMul<MUL_DERIV_PAIR>{spec, kek, s, s, s, axis_a, axis_b}, sb, s, s, +1, tw, st));
        break;
    case MUL_DERIV_LAP:
        CHK(pass32<0>(s, F32LdMul<MUL_DERIV_LAP>{spec, kek, s, s, s, axis_a, axis_b}, sb, s, s, +1, tw, st));
        break;
    case MUL_SMOOTH:
        CHK(pass32<0>(s, F32LdMul<MUL_SMOOTH>{spec, kek, s, s, s, axis_a, axis_b}, sb, s, s, +1, tw, st));
        break;
    default:
        CHK(pass32<0>(s, F32LdMul<MUL_NONE>{spec, kek, s, s, s, axis_a, axis_b}, sb, s, s, +1, tw, st));
        break;
    }
    CHK(pass32<1>(s, lb, sb, s, s, +1, tw, st));
    CHK(pass32<2>(s, lb, sg, s, s, +1, tw, st));
    return hipSuccess;
}

hipError_t hartree_isolated32(const Grid &g, const double *n, float2 *pad, const float *khat, const float2 *tw2,
                              float scale, double *out, hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    int s = g.s, M = 2 * s;
    F32StBuf sb{pad, M, M};
    F32LdBuf lb{pad, M, M};
    F32StMulK sk{pad, khat, M, M};
    F32StIso so{out, scale, g.EL, s};
    CHK(pass32<2>(M, F32LdPadZ{n, g.EL, s}, sb, s, s, -1, tw2, st));
    CHK(pass32<1>(M, F32LdBufPad{pad, M, M, 1, s}, sb, s, M, -1, tw2, st));
    CHK(pass32<0>(M, F32LdBufPad{pad, M, M, 0, s}, sk, M, M, -1, tw2, st));
    CHK(pass32<0>(M, lb, sb, M, M, +1, tw2, st));
    CHK(pass32<1>(M, lb, sb, s + 1, M, +1, tw2, st));
    CHK(pass32<2>(M, lb, so, s + 1, s + 1, +1, tw2, st));
    return hipSuccess;
}

hipError_t fft32_autotune(const Grid &g, const double *field, float2 *spec, const float2 *tw, float2 *pad,
                          const float *khat, const float2 *tw2, double *out, hipStream_t st) {
    if (g.D != 3) return hipErrorInvalidValue;
    static bool done[10] = {};
    auto tune = [&](int N, auto run) -> hipError_t {
        int k = N == 16 ? 0 : N == 32 ? 1 : N == 64 ? 2 : N == 96 ? 3 : N == 128 ? 4 : N == 192 ? 5 : N == 256 ? 6
              : N == 512 ? 7 : 8;
        if (done[k]) return hipSuccess;
        hipEvent_t e0, e1;
        CHK(hipEventCreate(&e0));
        CHK(hipEventCreate(&e1));
        float best = 1e30f;
        int bestSel = 1;
        for (int sel = 0; sel < 3; sel++) {
            fftk::bsel_slot(N, true) = sel;
            CHK(run());
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
        fftk::bsel_slot(N, true) = bestSel;
        done[k] = true;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        return hipSuccess;
    };
    CHK(tune(g.s, [&]() { return fft32_forward(g, field, nullptr, nullptr, spec, tw, st); }));
    if (pad) CHK(tune(2 * g.s, [&]() { return hartree_isolated32(g, field, pad, khat, tw2, 1.f, out, st); }));
    return hipSuccess;
}

hipError_t to_float(size_t n, const double *a, float *b, hipStream_t st) {
    hipLaunchKernelGGL(k_to_float, dim3(2048), dim3(256), 0, st, n, a, b);
    return hipGetLastError();
}
hipError_t to_float2(size_t n, const double2 *a, float2 *b, hipStream_t st) {
    hipLaunchKernelGGL(k_to_float2, dim3(2048), dim3(256), 0, st, n, a, b);
    return hipGetLastError();
}

}  // namespace mk
