"""ORACLE loader -- test infrastructure only.

ctypes access to oracle/liboracle.so, the plain-C FP64 restatement of the
reference's SC loop.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this module; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
SRC = [os.path.join(HERE, "dpft_oracle.c")]


def build():
    """gcc -O2 -fopenmp the oracle into oracle/liboracle.so."""
    cmd = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", "-o", LIB + ".tmp"] + SRC + ["-lm"]
    subprocess.run(cmd, check=True)
    os.replace(LIB + ".tmp", LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        from mpdpft_amd._abi import Params, Result  # shared ABI structs only
        if not os.path.exists(LIB):
            build()
        os.environ.setdefault("MPDPFT_TABLES", os.path.join(os.path.dirname(HERE), "mpdpft_amd", "lib",
                                                            "nait_tables.bin"))
        L = ctypes.CDLL(LIB)
        P = ctypes.POINTER(Params)
        R = ctypes.POINTER(Result)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_solve.argtypes = [P, R, dp, dp]
        L.oracle_fft_centered.argtypes = [P, ctypes.c_int, dp]
        L.oracle_dft_brute.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_fft1d.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_delfield.argtypes = [P, dp, ctypes.c_int, dp, dp, dp, dp, dp]
        L.oracle_vint.argtypes = [P, ctypes.c_int, ctypes.c_int, dp, dp]
        L.oracle_density.argtypes = [P, ctypes.c_int, ctypes.c_double, dp, dp, dp]
        L.oracle_env.argtypes = [P, dp]
        L.oracle_integrate.argtypes = [P, dp]
        L.oracle_integrate.restype = ctypes.c_double
        L.oracle_pw92.argtypes = [ctypes.c_double, dp, dp]
        L.oracle_pw92.restype = None
        L.oracle_pulay_coefficients.argtypes = [ctypes.c_int, dp, dp, dp]
        L.oracle_hartree_isolated_direct.argtypes = [P, dp, dp]
        L.oracle_cube_inverse_r.restype = ctypes.c_double
        L.oracle_enforce.argtypes = [P, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_enforce_trace.argtypes = [P, dp, dp, dp, dp, ip, dp, dp, ip, ctypes.c_int, ip]
        L.oracle_convergence_sequence.argtypes = [P, ctypes.c_int, dp, dp, dp, dp, dp]
        L.oracle_mix_sequence.argtypes = [P, ctypes.c_int, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
        L.oracle_time_iterations.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.oracle_time_iterations.restype = ctypes.c_double
        for n in ("oracle_curlyA", "oracle_airy", "oracle_exp_i0"):
            getattr(L, n).argtypes = [ctypes.c_double]
            getattr(L, n).restype = ctypes.c_double
        L.oracle_qmc_beta.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_qmc_beta.restype = ctypes.c_double
        u64p = ctypes.POINTER(ctypes.c_uint64)
        L.oracle_mt64_seed.argtypes = [u64p, ctypes.POINTER(ctypes.c_int), ctypes.c_uint64]
        L.oracle_mt64_next.argtypes = [u64p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_mt64_next.restype = ctypes.c_uint64
        L.oracle_rn.argtypes = [u64p, ctypes.POINTER(ctypes.c_int)]
        L.oracle_rn.restype = ctypes.c_double
        L.oracle_bessel_jn.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_bessel_jn.restype = ctypes.c_double
        L.oracle_polylog.argtypes = [ctypes.c_int, ctypes.c_double]
        L.oracle_nait_point.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int, ctypes.c_int, ctypes.c_double]
        L.oracle_nait_point.restype = ctypes.c_double
        L.oracle_ekin_point.argtypes = [ctypes.c_double] * 5 + [ctypes.c_int]
        L.oracle_ekin_point.restype = ctypes.c_double
        L.oracle_polylog.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def grid_size(p):
    return (p.steps + 1) ** p.DIM


def solve(p, v0=None):
    """The SC loop from V = Env, or from v0 ([S][G]) when given."""
    from mpdpft_amd._abi import Result
    G = grid_size(p)
    r = Result()
    den = np.zeros(p.S * G)
    V = np.zeros(p.S * G)
    if v0 is None:
        rc = lib().oracle_solve(ctypes.byref(p), ctypes.byref(r), _dp(den), _dp(V))
    else:
        v0 = np.ascontiguousarray(v0, dtype=np.float64).reshape(-1)
        rc = lib().oracle_solve_v0(ctypes.byref(p), _dp(v0), ctypes.byref(r), _dp(den), _dp(V))
    return rc, r, den.reshape(p.S, G), V.reshape(p.S, G)


def fft_centered(p, field, sign):
    a = np.ascontiguousarray(field, dtype=np.complex128).copy()
    rc = lib().oracle_fft_centered(ctypes.byref(p), sign, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == 0
    return a


def fft1d(x, sign, brute=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    f = lib().oracle_dft_brute if brute else lib().oracle_fft1d
    f(len(x), sign, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def delfield(p, f, method=5):
    G = grid_size(p)
    f = np.ascontiguousarray(f, dtype=np.float64)
    outs = [np.zeros(G) for _ in range(5)]
    lib().oracle_delfield(ctypes.byref(p), _dp(f), method, *[_dp(o) for o in outs])
    return dict(zip(("dx", "dy", "dz", "gradsq", "laplacian"), outs))


def load_eta(path):
    """Load a TabFunc_2D3D_eta.dat file for interaction 24."""
    L = lib()
    L.oracle_load_eta.argtypes = [ctypes.c_char_p]
    rc = L.oracle_load_eta(str(path).encode())
    assert rc == 0, rc


def energy_density(p, type_, s, den):
    """The energy integrand of interaction 14 or 24 on den ([S][G])."""
    L = lib()
    P = ctypes.POINTER(type(p))
    L.oracle_energy_density.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double)]
    G = grid_size(p)
    den = np.ascontiguousarray(den, dtype=np.float64).reshape(-1)
    out = np.zeros(G)
    rc = L.oracle_energy_density(ctypes.byref(p), type_, s, _dp(den), _dp(out))
    assert rc == 0, rc
    return out


def pulay_momentum_gram(p, res):
    """Gram matrix of residuals res [I][S][G] in the momentum-space metric."""
    L = lib()
    P = ctypes.POINTER(type(p))
    L.oracle_pulay_momentum_gram.argtypes = [P, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                             ctypes.POINTER(ctypes.c_double)]
    res = np.ascontiguousarray(res, dtype=np.float64)
    I = res.shape[0]
    M = np.zeros((I, I))
    rc = L.oracle_pulay_momentum_gram(ctypes.byref(p), I, _dp(res), _dp(M))
    assert rc == 0, rc
    return M


def vint(p, type_, s, den):
    G = grid_size(p)
    den = np.ascontiguousarray(den, dtype=np.float64).reshape(-1)
    out = np.zeros(G)
    rc = lib().oracle_vint(ctypes.byref(p), type_, s, _dp(den), _dp(out))
    assert rc == 0, rc
    return out


def density(p, s, mu, v):
    G = grid_size(p)
    v = np.ascontiguousarray(v, dtype=np.float64)
    out = np.zeros(G)
    N = ctypes.c_double()
    rc = lib().oracle_density(ctypes.byref(p), s, mu, _dp(v), _dp(out), ctypes.byref(N))
    assert rc == 0, rc
    return out, N.value


def enforce(p, v, mu):
    """GetDensities + EnforceConstraints on potential v ([S][G]) from mu:
    (mu, N, density [S][G], bracket failures)."""
    G = grid_size(p)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, dtype=np.float64).copy()
    N = np.zeros(p.S)
    den = np.zeros(p.S * G)
    nb = ctypes.c_int()
    rc = lib().oracle_enforce(ctypes.byref(p), _dp(v), _dp(m), _dp(N), _dp(den), ctypes.byref(nb))
    assert rc == 0, rc
    return m, N, den.reshape(p.S, G), nb.value


def enforce_trace(p, v, mu, max_evals=4096):
    """enforce() with the record of its density evaluations: (mu, N, bracket
    failures, trace) where trace is a list of (phase, mu, N) with phase 1 =
    walking mu down, 2 = walking up, 3 = bisection."""
    G = grid_size(p)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
    m = np.ascontiguousarray(mu, dtype=np.float64).copy()
    N = np.zeros(p.S)
    tm, tN = np.zeros(max_evals), np.zeros(max_evals)
    tp = (ctypes.c_int * max_evals)()
    nb, nt = ctypes.c_int(), ctypes.c_int()
    rc = lib().oracle_enforce_trace(ctypes.byref(p), _dp(v), _dp(m), _dp(N), None, ctypes.byref(nb), _dp(tm), _dp(tN),
                                    tp, max_evals, ctypes.byref(nt))
    assert rc == 0, rc
    return m, N, nb.value, [(tp[i], tm[i], tN[i]) for i in range(nt.value)]


def convergence_sequence(p, dens, Vs, Ns, mus):
    """CheckConvergence over states k = 0..n-1 (arrays [n][S][G], [n][S][G],
    [n][S], [n][S]): (CC per iteration, exit iteration or -1)."""
    dens = np.ascontiguousarray(dens, dtype=np.float64)
    Vs = np.ascontiguousarray(Vs, dtype=np.float64)
    Ns = np.ascontiguousarray(Ns, dtype=np.float64)
    mus = np.ascontiguousarray(mus, dtype=np.float64)
    n = dens.shape[0]
    CC = np.zeros(n)
    ex = lib().oracle_convergence_sequence(ctypes.byref(p), n, _dp(dens), _dp(Vs), _dp(Ns), _dp(mus), _dp(CC))
    return CC, ex


def mix_sequence(p, n_in0, n_outs):
    """AdmixDensities called with SCcount 0, 1, ... on n_outs[k], starting
    from n_in0: (mixed density [S][G], Pulay rejections)."""
    G = grid_size(p)
    n0 = np.ascontiguousarray(n_in0, dtype=np.float64).reshape(-1)
    no = np.ascontiguousarray(n_outs, dtype=np.float64).reshape(-1)
    k = no.size // (p.S * G)
    out = np.zeros(p.S * G)
    rj = ctypes.c_int()
    lib().oracle_mix_sequence(ctypes.byref(p), k, _dp(n0), _dp(no), _dp(out), ctypes.byref(rj))
    return out.reshape(p.S, G), rj.value


def env(p):
    out = np.zeros(p.S * grid_size(p))
    assert lib().oracle_env(ctypes.byref(p), _dp(out)) == 0
    return out.reshape(p.S, -1)


def integrate(p, f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    return lib().oracle_integrate(ctypes.byref(p), _dp(f))


def pw92(n):
    e, v = ctypes.c_double(), ctypes.c_double()
    lib().oracle_pw92(n, ctypes.byref(e), ctypes.byref(v))
    return e.value, v.value


def pulay_coefficients(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    c = np.zeros(n)
    tv = ctypes.c_double()
    ok = lib().oracle_pulay_coefficients(n, _dp(A), _dp(c), ctypes.byref(tv))
    return ok, c, tv.value


def hartree_isolated_direct(p, den):
    out = np.zeros(grid_size(p))
    den = np.ascontiguousarray(den, dtype=np.float64)
    lib().oracle_hartree_isolated_direct(ctypes.byref(p), _dp(den), _dp(out))
    return out


def cube_inverse_r():
    return lib().oracle_cube_inverse_r()


def time_iterations(p, n_iter):
    nb = ctypes.c_int()
    t = lib().oracle_time_iterations(ctypes.byref(p), n_iter, ctypes.byref(nb))
    return t, nb.value


def curlyA(y):
    return lib().oracle_curlyA(float(y))


def polylog(j, x):
    """Li_{3/2-j}(-e^x) as the nAiT density evaluates it."""
    return lib().oracle_polylog(int(j), float(x))


def airy(y):
    return lib().oracle_airy(float(y))


def nait_point(U, ar, lap, T, thr, bp, max_iter, acc):
    """The nAiT density at one point (Units 1, degeneracy 1)."""
    return lib().oracle_nait_point(float(U), float(ar), float(lap), float(T), float(thr), int(bp), int(max_iter),
                                   float(acc))


def ekin_point(U, a, lap, thr, acc, max_it):
    """EkinAiT03D's kinetic energy density at one point (Units 1, no prefactor)."""
    return lib().oracle_ekin_point(float(U), float(a), float(lap), float(thr), float(acc), int(max_it))


def qmc_beta(der, x):
    """The fitted b(eta) (der 0) / b'(eta) (der 1) of the renormalised 2D contact."""
    return lib().oracle_qmc_beta(int(der), float(x))


def exp_i0(x):
    """e^-x I0(x) (2D periodic Hartree compensation)."""
    return lib().oracle_exp_i0(float(x))


def noise_draws(seed, n, raw=False):
    """The first n draws of ImposeNoise's generator seeded with seed: the
    MT19937-64 outputs (raw) or the uniform values on [-1, 1)."""
    L = lib()
    mt = (ctypes.c_uint64 * 312)()
    mti = ctypes.c_int(0)
    L.oracle_mt64_seed(mt, ctypes.byref(mti), int(seed))
    f = L.oracle_mt64_next if raw else L.oracle_rn
    return [f(mt, ctypes.byref(mti)) for _ in range(n)]


def bessel_jn(n, x):
    """J_n(x), n = 0..3, elementwise over x >= 0 (Getn3p's kernel)."""
    f = lib().oracle_bessel_jn
    return np.array([f(int(n), float(v)) for v in np.atleast_1d(x)])
