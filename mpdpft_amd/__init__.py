"""mpdpft_amd -- MI355X-native self-consistent DPFT loop (drop-in for mpDPFT's GetData).

The compute path is libmpdpft.so (gfx950 HIP kernels + C++ host loop) behind
the C ABI in include/mpdpft.h; this package is its Python binding.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import Params, Result, load_library

__all__ = ["Params", "Result", "default_params", "parse_input", "Context", "Batch", "batch_solve_noalloc",
           "batch_fields_len", "run_dir", "set_data_dir", "load_library", "MpdpftError"]


class MpdpftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("mpdpft error %d: %s" % (code, msg))
        self.code = code


def default_params():
    p = Params()
    load_library().mpdpft_params_default(ctypes.byref(p))
    return p


def parse_input(path, base=None):
    p = base if base is not None else default_params()
    rc = load_library().mpdpft_parse_input(str(path).encode(), ctypes.byref(p))
    if rc != _abi.OK:
        raise MpdpftError(rc, "cannot parse %s" % path)
    return p


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """One solver context on one GPU (owns its device buffers and stream)."""

    def __init__(self, params, device=-1):
        self._lib = load_library()
        self._h = ctypes.c_void_p()
        self.params = Params.from_buffer_copy(params)  # set_centres edits this copy, not the caller's
        rc = self._lib.mpdpft_create(ctypes.byref(params), device, ctypes.byref(self._h))
        if rc != _abi.OK:
            msg = self._lib.mpdpft_last_error(self._h).decode() if self._h else ""
            self._lib.mpdpft_destroy(self._h)
            self._h = ctypes.c_void_p()
            raise MpdpftError(rc, msg)
        self.G = self._lib.mpdpft_grid_size(self._h)
        self.S = params.S

    def close(self):
        if self._h:
            self._lib.mpdpft_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def _check(self, rc, allow=(_abi.OK,)):
        if rc not in allow:
            raise MpdpftError(rc, self._lib.mpdpft_last_error(self._h).decode())
        return rc

    def solve(self):
        r = Result()
        rc = self._lib.mpdpft_solve(self._h, ctypes.byref(r))
        self._check(rc, (_abi.OK, _abi.ENOTCONV))
        return r

    def loop_begin(self):
        """Start a loop without iterating (mpdpft_loop_begin)."""
        self._check(self._lib.mpdpft_loop_begin(self._h))

    def step(self):
        """One SC iteration (mpdpft_step): (Result with the loop state, exit_loop)."""
        r = Result()
        ex = ctypes.c_int()
        self._check(self._lib.mpdpft_step(self._h, ctypes.byref(r), ctypes.byref(ex)))
        return r, bool(ex.value)

    def energy(self):
        """GetEnergy of the current fields (mpdpft_energy): the whole Result."""
        r = Result()
        self._check(self._lib.mpdpft_energy(self._h, ctypes.byref(r)))
        return r

    def enforce_constraints(self, mu):
        """GetDensities + EnforceConstraints on the current potential from mu
        (S values): (mu, N, status) with status OK or ENOBRACKET."""
        m = np.ascontiguousarray(mu, dtype=np.float64).copy()
        N = np.zeros(self.S)
        rc = self._check(self._lib.mpdpft_enforce_constraints(self._h, _dp(m), _dp(N)), (_abi.OK, _abi.ENOBRACKET))
        return m, N, rc

    def mix(self, sc_count, n_out):
        """AdmixDensities of n_out ([S][G]) into the current density
        (mpdpft_mix): OK, or EPULAY when the step fell back to linear mixing."""
        n = np.ascontiguousarray(n_out, dtype=np.float64).reshape(-1)
        return self._check(self._lib.mpdpft_mix(self._h, int(sc_count), _dp(n)), (_abi.OK, _abi.EPULAY))

    def inject_fault(self, kind, at):
        """Forcing hook at SC iteration `at` (mpdpft_inject_fault); at < 0 clears."""
        self._check(self._lib.mpdpft_inject_fault(self._h, int(kind), int(at)))

    def _field(self, fn):
        a = np.empty(self.S * self.G)
        self._check(fn(self._h, _dp(a)))
        return a.reshape(self.S, self.G)

    def density(self):
        return self._field(self._lib.mpdpft_get_density)

    def potential(self):
        return self._field(self._lib.mpdpft_get_potential)

    def environment(self):
        return self._field(self._lib.mpdpft_get_environment)

    def set_potential(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        self._check(self._lib.mpdpft_set_potential(self._h, _dp(v)))

    def fft(self, field, sign):
        """Centred transform of a complex field of G points (fftParallel)."""
        a = np.ascontiguousarray(field, dtype=np.complex128).copy()
        self._check(self._lib.mpdpft_fft(self._h, sign, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
        return a

    def delfield(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        outs = [np.zeros(self.G) for _ in range(5)]
        self._check(self._lib.mpdpft_delfield(self._h, _dp(f), *[_dp(o) for o in outs]))
        return dict(zip(("dx", "dy", "dz", "gradsq", "laplacian"), outs))

    def interaction_potential(self, type_, s, den):
        den = np.ascontiguousarray(den, dtype=np.float64).reshape(-1)
        out = np.zeros(self.G)
        self._check(self._lib.mpdpft_interaction_potential(self._h, type_, s, _dp(den), _dp(out)))
        return out

    def enable_timing(self, on=True):
        self._check(self._lib.mpdpft_enable_timing(self._h, 1 if on else 0))

    def stage_stats(self):
        """{stage: dict(launches, total_ms, units, bytes, flops)} of the GPU stage timers."""
        n = ctypes.c_int()
        arr = (_abi.StageStats * 64)()
        self._check(self._lib.mpdpft_get_stage_stats(self._h, arr, 64, ctypes.byref(n)))
        return {arr[i].name.decode(): dict(launches=arr[i].launches, total_ms=arr[i].total_ms, units=arr[i].units,
                                           bytes=arr[i].bytes, flops=arr[i].flops) for i in range(n.value)}

    def set_initial_potential(self, v):
        """Start the next solve from v ([S][G]) instead of V = Env; None resets."""
        if v is None:
            self._check(self._lib.mpdpft_set_initial_potential(self._h, None))
            return
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(-1)
        self._check(self._lib.mpdpft_set_initial_potential(self._h, _dp(v)))

    def adjoint(self, dL_dn, tol=0.0, max_iterations=0):
        """dL/dEnv of a loss of the last converged density (mpdpft_adjoint):
        dL_dn is dL/dn at every grid point (one species).  Returns (lambda,
        info dict)."""
        g = np.ascontiguousarray(dL_dn, dtype=np.float64).reshape(-1)
        if g.size != self.G:
            raise ValueError("dL_dn needs %d values" % self.G)
        lam = np.zeros(self.G)
        info = _abi.AdjointInfo()
        info.tol = tol
        info.max_iterations = max_iterations
        rc = self._check(self._lib.mpdpft_adjoint(self._h, _dp(g), _dp(lam), ctypes.byref(info)),
                         (_abi.OK, _abi.ENOTCONV))
        return lam, dict(iterations=info.iterations, relative_residual=info.relative_residual,
                         dN_dmu=info.dN_dmu, converged=rc == _abi.OK)

    def centre_gradient(self, lam):
        """sum_i lam_i dEnv_i/dR_c for the environment-41 centres, shape (n, 3)."""
        lam = np.ascontiguousarray(lam, dtype=np.float64).reshape(-1)
        out = np.zeros(3 * self.params.n_centres)
        self._check(self._lib.mpdpft_centre_gradient(self._h, _dp(lam), _dp(out)))
        return out.reshape(-1, 3)

    def centre_forces(self):
        """dE/dR_c of the last converged solve's total energy through the fixed point, shape (n, 3)."""
        out = np.zeros(3 * self.params.n_centres)
        self._check(self._lib.mpdpft_centre_forces(self._h, _dp(out)))
        return out.reshape(-1, 3)

    def task_gradient(self):
        """dE/dN of the last converged solve's total energy through the fixed point
        when the loop holds N (RelAcc > 0), dE/dmu when it holds mu (RelAcc = 0);
        one species (mpdpft_task_gradient)."""
        out = np.zeros(1)
        self._check(self._lib.mpdpft_task_gradient(self._h, _dp(out)))
        return float(out[0])

    def set_centres(self, centres):
        """Replace the environment-41 centres (rows x, y, z, Z, a) and rebuild Env."""
        c = np.ascontiguousarray(centres, dtype=np.float64).reshape(-1, 5)
        self._check(self._lib.mpdpft_set_centres(self._h, c.shape[0], _dp(c)))
        self.params.n_centres = c.shape[0]
        for i in range(c.shape[0]):
            for k in range(5):
                self.params.centres[i][k] = c[i, k]

    def set_break_file(self, path):
        self._check(self._lib.mpdpft_set_break_file(self._h, str(path).encode() if path else None))

    def density_at(self, s, mu, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        out = np.zeros(self.G)
        N = ctypes.c_double()
        self._check(self._lib.mpdpft_density(self._h, s, mu, _dp(v), _dp(out), ctypes.byref(N)))
        return out, N.value


class Batch:
    """Many independent TF-family solves in one persistent launch
    (mpdpft_batch_*).  instances: a list of Params; fp32: FP32 field storage
    with an FP64 polish."""

    def __init__(self, instances, device=-1, fp32=False):
        self._lib = load_library()
        self.n = len(instances)
        self.params = list(instances)
        arr = (Params * self.n)(*instances)
        self._h = ctypes.c_void_p()
        rc = self._lib.mpdpft_batch_create(arr, self.n, device, _abi.BATCH_FP32 if fp32 else 0, ctypes.byref(self._h))
        if rc != _abi.OK:
            msg = self._lib.mpdpft_batch_last_error(self._h).decode() if self._h else ""
            self.close()
            raise MpdpftError(rc, "batch_create: " + msg)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mpdpft_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001
            pass

    def solve(self):
        res = (Result * self.n)()
        rc = self._lib.mpdpft_batch_solve(self._h, res)
        if rc != _abi.OK:
            raise MpdpftError(rc, "batch_solve: " + self._lib.mpdpft_batch_last_error(self._h).decode())
        return list(res)

    def seconds(self):
        return self._lib.mpdpft_batch_seconds(self._h)

    def phase_stats(self):
        """((fp32_seconds, fp64_seconds), (fp32_iterations, fp64_iterations),
        (fp32_bisection_steps, fp64_bisection_steps)) of the last solve."""
        sec = (ctypes.c_double * 2)()
        its = (ctypes.c_longlong * 2)()
        bis = (ctypes.c_longlong * 2)()
        self._check(self._lib.mpdpft_batch_phase_stats(self._h, sec, its, bis))
        return (sec[0], sec[1]), (its[0], its[1]), (bis[0], bis[1])

    def kernel_stats(self):
        """(kernel_ms[2], stage bytes[2], kernel bytes[2], per-instance
        microseconds[n]) of the last solve's persistent launches: [0] the FP32
        loop, [1] the FP64 loop or polish (HIP events on the batch's stream;
        stage bytes: the reference stages' inputs + outputs, kernel bytes: the
        fused sweeps' own accesses)."""
        ms = (ctypes.c_double * 2)()
        sb = (ctypes.c_double * 2)()
        kb = (ctypes.c_double * 2)()
        us = np.zeros(self.n)
        self._check(self._lib.mpdpft_batch_kernel_stats(self._h, ms, sb, kb, _dp(us)))
        return (ms[0], ms[1]), (sb[0], sb[1]), (kb[0], kb[1]), us

    def _check(self, rc):
        if rc != _abi.OK:
            raise MpdpftError(rc, self._lib.mpdpft_batch_last_error(self._h).decode())

    def fields(self, i):
        p = self.params[i]
        SG = p.S * (p.steps + 1) ** p.DIM
        den, V = np.zeros(SG), np.zeros(SG)
        rc = self._lib.mpdpft_batch_get_fields(self._h, i, _dp(den), _dp(V))
        if rc != _abi.OK:
            raise MpdpftError(rc, "batch_get_fields")
        return den.reshape(p.S, -1), V.reshape(p.S, -1)


def batch_fields_len(instances):
    """Doubles the converged fields of these instances take in the output of
    batch_solve_noalloc (instance i's [S][G] density then [S][G] potential at
    the sum of the earlier instances' 2 S G)."""
    return sum(2 * p.S * (p.steps + 1) ** p.DIM for p in instances)


def batch_solve_noalloc(instances, fields_ptr, fields_len, device=0, fp32=False):
    """The batched solve without device allocation (mpdpft_batch_solve_noalloc):
    fields_ptr is the address of device memory on `device` holding fields_len
    doubles, which receives the converged fields; returns the results."""
    lib = load_library()
    n = len(instances)
    arr = (Params * n)(*instances)
    res = (Result * n)()
    err = ctypes.create_string_buffer(512)
    rc = lib.mpdpft_batch_solve_noalloc(arr, n, device, _abi.BATCH_FP32 if fp32 else 0, ctypes.c_void_p(fields_ptr),
                                        fields_len, res, err, len(err))
    if rc != _abi.OK:
        raise MpdpftError(rc, "batch_solve_noalloc: " + err.value.decode())
    return list(res)


def import_v(path, params):
    """LoadV/ImportV: a potential file interpolated onto the lattice of params
    (host code, no GPU).  Returns ([S][G] array, warning count)."""
    lib = load_library()
    S = params.S
    G = (params.steps + 1) ** params.DIM
    out = np.zeros(S * G)
    w = ctypes.c_int()
    rc = lib.mpdpft_import_v(str(path).encode(), ctypes.byref(params), _dp(out), ctypes.byref(w))
    if rc != _abi.OK:
        raise MpdpftError(rc, "import_v(%s)" % path)
    return out.reshape(S, G), w.value


def set_data_dir(path):
    """Directory of the TabFunc_*.dat tables contexts read (TabFunc_2D3D_eta.dat
    for interaction 24); None or "" = the current directory."""
    load_library().mpdpft_set_data_dir(str(path or "").encode())


def run_dir(path, device=-1):
    r = Result()
    rc = load_library().mpdpft_run_dir(str(path).encode(), device, ctypes.byref(r))
    if rc not in (_abi.OK, _abi.ENOTCONV):
        raise MpdpftError(rc, "run_dir(%s)" % path)
    return r
