// fft_core.h -- LDS-resident Stockham FFT pass for gfx950 (device templates).
//
// A 1D/2D/3D transform is a sequence of axis passes.  One pass moves every
// line of one axis through LDS exactly once: a 256-thread workgroup loads B
// lines (B consecutive values of the fast line coordinate, so global reads and
// writes are contiguous runs), runs the Stockham radix stages in LDS with
// twiddles from an LDS table, and stores the B lines back.  Loads and stores
// go through functors, which is where the centred-grid layout, zero padding,
// k-space multipliers, scaling and the periodic-image fill are fused, so no
// pass ever touches HBM twice.
//
// Indexing: buffers are [n0][n1][n2] with n2 contiguous; a 2D transform uses
// n0 = 1.  Axis 2 lines are (x, y), axis 1 lines are (x, z), axis 0 lines are
// (y, z).  Forward uses exp(-2 pi i jk/N) and is unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace fftk {

template <typename T> struct c2;
template <> struct c2<double> { using type = double2; };
template <> struct c2<float> { using type = float2; };

template <typename T2> __device__ __forceinline__ T2 cmul(T2 a, T2 b) {
    T2 r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
}
template <typename T2> __device__ __forceinline__ T2 cadd(T2 a, T2 b) {
    T2 r;
    r.x = a.x + b.x;
    r.y = a.y + b.y;
    return r;
}
template <typename T2> __device__ __forceinline__ T2 csub(T2 a, T2 b) {
    T2 r;
    r.x = a.x - b.x;
    r.y = a.y - b.y;
    return r;
}
// multiply by sign*i (sign = -1 forward, +1 backward)
template <typename T2> __device__ __forceinline__ T2 mul_si(T2 a, int sign) {
    T2 r;
    if (sign < 0) {
        r.x = a.y;
        r.y = -a.x;
    } else {
        r.x = -a.y;
        r.y = a.x;
    }
    return r;
}

// In-register small DFTs, out[k] = sum_j in[j] exp(sign 2 pi i jk/R).
template <int R, typename T2> struct SmallDFT;

template <typename T2> struct SmallDFT<2, T2> {
    __device__ __forceinline__ static void run(T2 *v, int) {
        T2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <typename T2> struct SmallDFT<3, T2> {
    __device__ __forceinline__ static void run(T2 *v, int sign) {
        using T = decltype(v[0].x);
        const T c = (T)-0.5, sn = (T)0.86602540378443864676 * (T)sign;
        T2 a = v[0], b = v[1], d = v[2];
        T2 s = cadd(b, d), t = csub(b, d);
        T2 m;
        m.x = a.x + c * s.x;
        m.y = a.y + c * s.y;
        v[0] = cadd(a, s);
        // exp(sign*2pi i/3) = c + i sn
        v[1].x = m.x - sn * t.y;
        v[1].y = m.y + sn * t.x;
        v[2].x = m.x + sn * t.y;
        v[2].y = m.y - sn * t.x;
    }
};

template <typename T2> struct SmallDFT<4, T2> {
    __device__ __forceinline__ static void run(T2 *v, int sign) {
        T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        T2 b0 = cadd(v[1], v[3]), b1 = mul_si(csub(v[1], v[3]), sign);
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};

template <typename T2> struct SmallDFT<8, T2> {
    __device__ __forceinline__ static void run(T2 *v, int sign) {
        using T = decltype(v[0].x);
        const T h = (T)0.70710678118654752440;
        // even/odd split: two radix-4 DFTs, then combine with W8^k
        T2 e[4] = {v[0], v[2], v[4], v[6]};
        T2 o[4] = {v[1], v[3], v[5], v[7]};
        SmallDFT<4, T2>::run(e, sign);
        SmallDFT<4, T2>::run(o, sign);
        // W8^1 = h(1 + sign i), W8^2 = sign i, W8^3 = h(-1 + sign i)
        T2 w1o, w2o, w3o;
        w1o.x = h * (o[1].x - (T)sign * o[1].y);
        w1o.y = h * (o[1].y + (T)sign * o[1].x);
        w2o = mul_si(o[2], sign);
        w3o.x = h * (-o[3].x - (T)sign * o[3].y);
        w3o.y = h * (-o[3].y + (T)sign * o[3].x);
        v[0] = cadd(e[0], o[0]);
        v[4] = csub(e[0], o[0]);
        v[1] = cadd(e[1], w1o);
        v[5] = csub(e[1], w1o);
        v[2] = cadd(e[2], w2o);
        v[6] = csub(e[2], w2o);
        v[3] = cadd(e[3], w3o);
        v[7] = csub(e[3], w3o);
    }
};

constexpr int kThreads = 256;

// Radix plan per length: first radix R1 (evaluated straight from the load
// functor), middle radices in LDS, last radix written straight through the
// store functor.
template <int N> struct Plan;
template <> struct Plan<16> { static constexpr int n = 2, r[3] = {4, 4, 1}; };
template <> struct Plan<32> { static constexpr int n = 2, r[3] = {8, 4, 1}; };
template <> struct Plan<64> { static constexpr int n = 2, r[3] = {8, 8, 1}; };
template <> struct Plan<96> { static constexpr int n = 3, r[3] = {8, 4, 3}; };
template <> struct Plan<128> { static constexpr int n = 3, r[3] = {8, 8, 2}; };
template <> struct Plan<192> { static constexpr int n = 3, r[3] = {8, 8, 3}; };
template <> struct Plan<256> { static constexpr int n = 3, r[3] = {8, 8, 4}; };
template <> struct Plan<512> { static constexpr int n = 3, r[3] = {8, 8, 8}; };

// Butterfly index j of a stage with N/R butterflies per line -> (line b,
// butterfly jj).  For lines along the contiguous axis (AXIS 2) consecutive
// threads take consecutive butterflies of one line; otherwise consecutive
// lines (contiguous in memory) -- either way global accesses are runs.
template <int N, int R, int B, int AXIS>
__device__ __forceinline__ void split(int j, int &b, int &jj) {
    if constexpr (AXIS == 2) {
        b = j / (N / R);
        jj = j - b * (N / R);
    } else {
        jj = j / B;
        b = j - jj * B;
    }
}

// One Stockham stage (Govindaraju et al., SC'08 formulation) over B lines.
// SRC/DST 0: LDS (row pitch P); SRC 1: the load functor; DST 1: the store
// functor.  Ns = product of the earlier radices.
template <int N, int B, int P, int R, int AXIS, int SRC, int DST, typename T2, class Ld, class St>
__device__ __forceinline__ void stage(T2 *lds, const T2 *tw, int Ns, int sign, const Ld &ld, const St &st, int slow,
                                      int fast0, int L1) {
    constexpr int NB = B * (N / R);
    constexpr int PER = (NB + kThreads - 1) / kThreads;
    const int tid = threadIdx.x;
    T2 v[PER][R];
    int dst[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        int j = tid + q * kThreads;
        dst[q] = -1;
        if (j < NB) {
            int b, jj;
            split<N, R, B, AXIS>(j, b, jj);
            const int k = jj % Ns, f = fast0 + b;
            const bool live = f < L1;
            if constexpr (SRC == 1) {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    int t = jj + r * (N / R);
                    T2 z;
                    z.x = 0;
                    z.y = 0;
                    if (live) {
                        if constexpr (AXIS == 2) z = ld(slow, f, t);
                        else if constexpr (AXIS == 1) z = ld(slow, t, f);
                        else z = ld(t, slow, f);
                    }
                    v[q][r] = z;
                }
            } else {
                const T2 *line = lds + b * P;
#pragma unroll
                for (int r = 0; r < R; r++) v[q][r] = line[jj + r * (N / R)];
            }
            if (Ns > 1) {
                const int step = N / (Ns * R);
#pragma unroll
                for (int r = 1; r < R; r++) {
                    T2 w = tw[(r * k * step) % N];
                    if (sign > 0) w.y = -w.y;
                    v[q][r] = cmul(v[q][r], w);
                }
            }
            SmallDFT<R, T2>::run(v[q], sign);
            const int o = (jj / Ns) * Ns * R + k;
            if constexpr (DST == 1) {
                if (live) {
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        int t = o + r * Ns;
                        if constexpr (AXIS == 2) st(slow, f, t, v[q][r]);
                        else if constexpr (AXIS == 1) st(slow, t, f, v[q][r]);
                        else st(t, slow, f, v[q][r]);
                    }
                }
            } else {
                dst[q] = b * P + o;
            }
        }
    }
    if constexpr (DST == 0) {
        if constexpr (SRC == 0) __syncthreads();  // every read of this stage is done
#pragma unroll
        for (int q = 0; q < PER; q++) {
            if (dst[q] >= 0) {
#pragma unroll
                for (int r = 0; r < R; r++) lds[dst[q] + r * Ns] = v[q][r];
            }
        }
        __syncthreads();
    }
}

// Pass geometry: the axis length N is a template parameter; L0/L1 are the
// numbers of lines along the slow and fast line coordinates (so pruned
// passes of a zero-padded transform simply launch fewer lines).
struct PassGeom {
    int L0, L1;       // line extents: axis2 -> (x<L0, y<L1); axis1 -> (x<L0, z<L1); axis0 -> (y<L0, z<L1)
    int sign;
};

// Load functor: T2 operator()(int x, int y, int z) const
// Store functor: void operator()(int x, int y, int z, T2 v) const
// The first radix stage reads through the load functor and the last writes
// through the store functor, so the B lines make one round trip through LDS
// per middle stage plus one.
template <int N, int B, int AXIS, typename T2, class Ld, class St>
__global__ __launch_bounds__(kThreads) void fft_pass(Ld ld, St st, PassGeom g, const T2 *__restrict__ twg) {
    constexpr int P = N + 1;  // pad: keep column accesses off one bank
    constexpr int R1 = Plan<N>::r[0], R2 = Plan<N>::r[1], R3 = Plan<N>::r[2];
    __shared__ T2 lds[B * P];
    __shared__ T2 tw[N];
    const int tid = threadIdx.x;
    for (int t = tid; t < N; t += kThreads) tw[t] = twg[t];
    const int tilesFast = (g.L1 + B - 1) / B;
    const int slow = blockIdx.x / tilesFast;
    const int fast0 = (blockIdx.x - slow * tilesFast) * B;
    if constexpr (Plan<N>::n == 2) {
        stage<N, B, P, R1, AXIS, 1, 0>(lds, tw, 1, g.sign, ld, st, slow, fast0, g.L1);
        stage<N, B, P, R2, AXIS, 0, 1>(lds, tw, R1, g.sign, ld, st, slow, fast0, g.L1);
    } else {
        stage<N, B, P, R1, AXIS, 1, 0>(lds, tw, 1, g.sign, ld, st, slow, fast0, g.L1);
        stage<N, B, P, R2, AXIS, 0, 0>(lds, tw, R1, g.sign, ld, st, slow, fast0, g.L1);
        stage<N, B, P, R3, AXIS, 0, 1>(lds, tw, R1 * R2, g.sign, ld, st, slow, fast0, g.L1);
    }
}

template <int N> constexpr int lines_per_block() {
    return N <= 64 ? 32 : (N <= 192 ? 16 : 8);
}

// Block shape choice per transform length and precision, set by the
// autotuner (mk::fft_autotune): 0 = half the default lines per block,
// 1 = the default, 2 = twice the default.  Host-side table, default 1.
inline int &bsel_slot(int N, bool fp32) {
    static int table[2][10] = {{1, 1, 1, 1, 1, 1, 1, 1, 1, 1}, {1, 1, 1, 1, 1, 1, 1, 1, 1, 1}};
    int k = N == 16 ? 0 : N == 32 ? 1 : N == 64 ? 2 : N == 96 ? 3 : N == 128 ? 4 : N == 192 ? 5 : N == 256 ? 6
          : N == 512 ? 7 : 8;
    return table[fp32 ? 1 : 0][k];
}

// Every length and precision to one block shape (the tuner then leaves them)
inline void bsel_force(int sel) {
    for (int N : {16, 32, 64, 96, 128, 192, 256, 512}) bsel_slot(N, false) = bsel_slot(N, true) = sel;
}

template <int N, int B, int AXIS, typename T2, class Ld, class St>
inline hipError_t launch_pass_b(Ld ld, St st, int L0, int L1, int sign, const T2 *tw, hipStream_t stream) {
    PassGeom g{L0, L1, sign};
    unsigned blocks = (unsigned)(L0 * ((L1 + B - 1) / B));
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL((fft_pass<N, B, AXIS, T2, Ld, St>), dim3(blocks), dim3(kThreads), 0, stream, ld, st, g, tw);
    return hipGetLastError();
}

// LDS of a pass (lines + twiddles) within the 160 KB of a CU
template <int N, int B, typename T2> constexpr bool lds_fits() { return (B * (N + 1) + N) * sizeof(T2) <= 160 * 1024; }

template <int N, int AXIS, typename T2, class Ld, class St>
inline hipError_t launch_pass(Ld ld, St st, int L0, int L1, int sign, const T2 *tw, hipStream_t stream) {
    // FP32 lines are half the bytes: twice as many per block keeps the LDS tile and the contiguous runs
    constexpr bool f32 = sizeof(T2) == 8;
    constexpr int B1 = lines_per_block<N>() * (f32 ? 2 : 1);
    constexpr int B0 = B1 / 2 > 0 ? B1 / 2 : 1, B2 = B1 * 2;
    const int sel = bsel_slot(N, f32);
    if (sel == 0) return launch_pass_b<N, B0, AXIS>(ld, st, L0, L1, sign, tw, stream);
    if constexpr (lds_fits<N, B2, T2>()) {
        if (sel == 2) return launch_pass_b<N, B2, AXIS>(ld, st, L0, L1, sign, tw, stream);
    }
    return launch_pass_b<N, B1, AXIS>(ld, st, L0, L1, sign, tw, stream);
}

}  // namespace fftk
