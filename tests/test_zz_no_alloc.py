"""mpdpft_batch_solve_noalloc (the batched solve with no device allocation,
its working memory in the library's static scratch array) against
mpdpft_batch_create + mpdpft_batch_solve on the same instances: two calls in a
row, the second with fewer and smaller instances, results and converged fields
at the FP64 bar (1e-10 relative, SC iteration count +-1; the kernels are the
same, so they agree exactly), and the host-side rejections of sizes beyond the
scratch array.  Runs as its own process or after the other GPU tests."""
import ctypes

import numpy as np
import pytest

from mpdpft_amd import Batch, MpdpftError, _abi, batch_fields_len, batch_solve_noalloc, configs, load_library

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def pick(counts):
    """The first instances of each grid size in the seeded C3 list:
    counts = {grid: how many}."""
    insts = configs.c3(n=512)
    out = []
    for g, k in counts.items():
        out += [p for p in insts if p.steps == g][:k]
    return out


class DeviceBuffer:
    """n doubles of device memory from the HIP runtime the library runs on
    (the caller's output buffer), filled with NaN."""

    def __init__(self, n):
        load_library()
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.n = n
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), 8 * n) == 0
        nan = np.full(n, np.nan)
        assert self.hip.hipMemcpy(self.ptr, nan.ctypes.data, 8 * n, 1) == 0  # host to device

    def data_ptr(self):
        return self.ptr.value

    def host(self):
        out = np.empty(self.n)
        assert self.hip.hipMemcpy(out.ctypes.data, self.ptr, 8 * self.n, 2) == 0  # device to host
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def device_fields(n):
    return DeviceBuffer(n)


def compare(instances, fp32=False):
    print("noalloc vs batch: %d instances, grids %s, fp32 %s" % (
        len(instances), sorted({p.steps for p in instances}), fp32), flush=True)
    L = batch_fields_len(instances)
    f = device_fields(L + 64)  # room beyond the fields: nothing is written there
    r_na = batch_solve_noalloc(instances, f.data_ptr(), L + 64, device=0, fp32=fp32)
    host = f.host()
    f.free()
    b = Batch(instances, device=0, fp32=fp32)
    r_b = b.solve()
    off = 0
    worst = 0.0
    for i, p in enumerate(instances):
        SG = p.S * (p.steps + 1) ** p.DIM
        den, V = b.fields(i)
        d_na, V_na = host[off:off + SG], host[off + SG:off + 2 * SG]
        off += 2 * SG
        a, c = r_na[i], r_b[i]
        assert a.converged and c.converged
        assert abs(a.SCcount - c.SCcount) <= 1
        assert a.LinMix == c.LinMix and a.PulMix == c.PulMix
        for s in range(p.S):
            assert abs(a.muVec[s] - c.muVec[s]) <= RTOL * abs(c.muVec[s])
            assert abs(a.FinalAbundances[s] - c.FinalAbundances[s]) <= RTOL * abs(c.FinalAbundances[s])
        assert abs(a.TotalEnergy - c.TotalEnergy) <= RTOL * abs(c.TotalEnergy)
        e = max(np.abs(d_na - den.reshape(-1)).max() / np.abs(den).max(),
                np.abs(V_na - V.reshape(-1)).max() / np.abs(V).max())
        assert e < RTOL
        worst = max(worst, e, abs(a.TotalEnergy - c.TotalEnergy) / abs(c.TotalEnergy))
    b.close()
    assert np.isnan(host[L:]).all()
    print("  worst relative difference %.2e" % worst, flush=True)


def test_noalloc_matches_batch_two_calls():
    # first call: two 256^2 instances (22 MB slots: three fit, so they run as
    # two-workgroup teams), two 128^2 and two 64^2; second: one 128^2, one 64^2
    compare(pick({256: 2, 128: 2, 64: 2}))
    compare(pick({128: 1, 64: 1}))


def test_noalloc_fp32_polish_matches_batch():
    compare(pick({128: 1, 64: 2}), fp32=True)


def test_noalloc_rejects_what_does_not_fit():
    small = pick({64: 1})
    L = batch_fields_len(small)
    f = device_fields(L)
    with pytest.raises(MpdpftError) as e:  # output shorter than the fields
        batch_solve_noalloc(small, f.data_ptr(), L - 1)
    assert e.value.code == _abi.EINVAL and "fields holds" in str(e.value)
    big = configs.c3(n=1)[0]  # one 512^2 two-species instance: one slot is 88 MB
    big.steps = 512
    Lb = batch_fields_len([big])
    fb = device_fields(Lb)
    with pytest.raises(MpdpftError) as e:
        batch_solve_noalloc([big], fb.data_ptr(), Lb)
    assert e.value.code == _abi.EINVAL and "scratch" in str(e.value)
    with pytest.raises(MpdpftError) as e:
        batch_solve_noalloc(small * (_abi.NOALLOC_MAX_INSTANCES + 1), f.data_ptr(), L)
    assert e.value.code == _abi.EINVAL
    f.free()
    fb.free()
