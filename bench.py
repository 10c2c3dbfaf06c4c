#!/usr/bin/env python3
"""Converged self-consistent DPFT solves per second on one MI355X.

Headline workload (a step): one batched solve of 512 independent 2D
two-species Thomas-Fermi + contact gases (seed 20260101: N per species in
[20, 400], anisotropic trap (x^2 + lambda^2 y^2)/2 with lambda in [1, 2],
contact g in [0.2, 2.0], grids 64^2/128^2/256^2 with probabilities
0.5/0.3/0.2 on a box 2.5 sqrt(2N), Pulay history 8, mixing 0.1, converged at
max|dV| < 1e-8) in one persistent launch, FP64 throughout.  `value` is
converged solves per second (512 x steps / wall time of the timed solves,
inputs resident on the device).  The same line carries:

  roofline      the persistent kernel (k_batch_solve, FP64): its algorithmic
                bytes per launch (each field access of every instance's loop,
                counted by the kernel itself) over its launch time from HIP
                events on the batch's stream, against 8 TB/s; the counter
                traffic comes from a rocprof summary under profiles/
  batch         the FP32 + FP64-polish batch (rate, |dE| against FP64 and
                against the oracle) and the plain FP64 GPU loop on the same
                512 instances (one context per instance, one after another)
  c1            one converged solve of the 3D drop-in workload (128^3, nAiT
                T = 0.05, isolated Hartree on the zero-padded 256^3 grid,
                Dirac + PW92): its nAiT kernel roofline (FP64 VALU) and the
                HBM rates of its transform and pointwise stages
  adjoint       the 96^3 differentiation workload: solve, implicit adjoint
                gradient of int (n - n_target)^2 and the total-energy forces
                with respect to the 96 centre coordinates, checked against
                central differences (h = 1e-4) on a few coordinates
  cpu_baseline  the oracle (plain C restatement, FP64) on a bounded sample of
                the same 512 instances: one instance per core, and in the
                notes sequentially on all cores and on one thread

Every leg runs under an internal deadline; a leg that does not fit is named
as skipped and the line is always printed.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--deadline S]
"""
import argparse
import json
import os
import subprocess
import sys
import time
import traceback

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged SC solves/s vs FP64 GPU at |ΔE|<1e-6; FFT/pointwise HBM GB/s % peak"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
N_BATCH = 512


def host_cpu():
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return name


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)


def read_summary_value(path, key):
    """'key = value' line of a committed rocprof summary, or None."""
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for ln in f:
            if ln.strip().startswith(key):
                try:
                    return float(ln.split("=", 1)[1].split()[0])
                except (ValueError, IndexError):
                    return None
    return None


# ---------------------------------------------------------------- CPU legs --
_CPU_WORKER = r"""
import ctypes, json, sys, time
sys.path.insert(0, %r)
from mpdpft_amd._abi import Params   # the struct layout only; the GPU library is not loaded here
from oracle import oracle as O
raw = open(sys.argv[1], "rb").read()
sz = ctypes.sizeof(Params)
idx = json.loads(sys.argv[2])
O.lib()
out = []
for i in idx:
    p = Params.from_buffer_copy(raw[i * sz:(i + 1) * sz])
    t0 = time.perf_counter()
    rc, r, den, V = O.solve(p)
    out.append([i, time.perf_counter() - t0, r.TotalEnergy, int(r.SCcount), int(r.converged)])
print(json.dumps(out))
""" % ROOT
_PARAMS_FILE = None


def _params_file(n):
    """The first n seeded C3 instances as raw mpdpft_params records (for the
    oracle's worker processes, which load no GPU library)."""
    global _PARAMS_FILE
    if _PARAMS_FILE is None:
        import tempfile
        from mpdpft_amd import configs
        fd, path = tempfile.mkstemp(prefix="c3_params_", suffix=".bin")
        with os.fdopen(fd, "wb") as f:
            for p in configs.c3(n=n):
                f.write(bytes(p))
        _PARAMS_FILE = path
    return _PARAMS_FILE


def _oracle_run(indices, threads):
    """Oracle solves of the given C3 instances in a child process with
    OMP_NUM_THREADS = threads (the oracle is plain C + OpenMP)."""
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    r = subprocess.run([sys.executable, "-c", _CPU_WORKER, _params_file(N_BATCH), json.dumps(list(indices))], env=env,
                       capture_output=True, text=True, check=True)
    return json.loads(r.stdout.strip().splitlines()[-1])


def cpu_baseline_leg(budget_s):
    """C3 on the host, on the first 2 x cores seeded instances: one instance
    per core (independent jobs side by side: the value), sequentially with
    all threads (the reference's runY jobs one after another) and sequentially
    on one thread (the last two on a prefix of that sample when the budget
    runs out)."""
    cores = host_threads()
    t_start = time.perf_counter()
    # one instance per core: 2 x cores instances, one worker process each (OMP 1)
    n_par = 2 * cores
    procs = []
    t0 = time.perf_counter()
    env = dict(os.environ, OMP_NUM_THREADS="1")
    for w in range(cores):
        idx = list(range(w, n_par, cores))
        procs.append(subprocess.Popen([sys.executable, "-c", _CPU_WORKER, _params_file(N_BATCH), json.dumps(idx)], env=env,
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    par = []
    for p in procs:
        out, err = p.communicate(timeout=max(30.0, budget_s))
        if p.returncode != 0:
            raise RuntimeError("oracle worker failed: " + err[-400:])
        par += json.loads(out.strip().splitlines()[-1])
    t_par = time.perf_counter() - t0
    # sequential on all threads, then sequential on one thread, over the same instances (the
    # first n_par, in chunks of 4 until each mode has used its share of the budget)
    def sequential(threads, share):
        done, t0, k = [], time.perf_counter(), 0
        while k < n_par and time.perf_counter() - t0 < share * budget_s:
            done += _oracle_run(range(k, min(k + 4, n_par)), threads)
            k += 4
        return done
    seq = sequential(cores, 0.25)
    t_seq = sum(x[1] for x in seq)
    one = sequential(1, 0.5)
    t_one = sum(x[1] for x in one)
    return {
        "value": round(len(par) / t_par, 3),
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": ("the oracle (plain C FP64 restatement of the SC loop) on the first %d of the 512 seeded "
                   "instances, one instance per core (%d worker processes, 1 thread each): %.2f s wall; "
                   "host %s" % (len(par), cores, t_par, host_cpu())),
        "sequential_all_threads": {"instances": len(seq), "seconds": round(t_seq, 3),
                                   "solves_per_s": round(len(seq) / t_seq, 3), "threads": cores},
        "sequential_one_thread": {"instances": len(one), "seconds": round(t_one, 3),
                                  "solves_per_s": round(len(one) / t_one, 3), "threads": 1},
        "oracle_energies": {str(x[0]): x[2] for x in par},
        "leg_seconds": round(time.perf_counter() - t_start, 1),
    }


# ---------------------------------------------------------------- GPU legs --
def batch_leg(steps, warmup, plain_n, oracle_energies):
    """The headline: FP64 batch solves/s; FP32 + polish; the plain FP64 loop."""
    import numpy as np
    from mpdpft_amd import Batch, Context, configs
    insts = configs.c3(n=N_BATCH)
    b64 = Batch(insts)
    for _ in range(warmup):
        b64.solve()
    ms = sb = kb = 0.0
    t0 = time.perf_counter()
    for _ in range(steps):
        r64 = b64.solve()
        (_, k_ms), (_, k_sb), (_, k_kb), us = b64.kernel_stats()
        ms += k_ms
        sb += k_sb
        kb += k_kb
    elapsed = time.perf_counter() - t0
    f64 = [b64.fields(i) for i in range(N_BATCH)]
    b64.close()
    grid = np.array([p.steps for p in insts])
    per_class = {str(g): {"instances": int((grid == g).sum()), "max_instance_ms": round(float(us[grid == g].max()) / 1e3, 2),
                          "mean_instance_ms": round(float(us[grid == g].mean()) / 1e3, 2)}
                 for g in (64, 128, 256)}
    avg_ms = ms / steps
    bytes_per_launch = sb / steps
    achieved = bytes_per_launch / (avg_ms / 1e3) / 1e9
    traffic = read_summary_value(os.path.join(ROOT, "profiles", "batch_c3_512_traffic.md"), "traffic_bytes_per_launch")
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": "k_batch_solve<double>",
            "avg_launch_ms": round(avg_ms, 3), "algorithmic_bytes_per_launch": bytes_per_launch,
            "fused_kernel_bytes_per_launch": kb / steps,
            "note": "one launch solves all 512 instances; algorithmic bytes = per SC iteration of every instance "
                    "the reference stages' inputs + outputs (UpdateV 3 SG, N(mu) G per evaluation, density 2 SG, "
                    "convergence 2 SG, Pulay Gram row (I+1) SG and combination (2I+1) SG or linear mixing 3 SG; "
                    "8-byte fields), counted per instance by the kernel; the fused sweeps themselves move "
                    "fused_kernel_bytes_per_launch"}
    # FP32 fields + FP64 polish
    b32 = Batch(insts, fp32=True)
    b32.solve()
    n32 = max(1, min(steps, 10))
    t0 = time.perf_counter()
    for _ in range(n32):
        r32 = b32.solve()
    t32 = (time.perf_counter() - t0) / n32
    (ms32, by32, _, _) = b32.kernel_stats()
    (ph_s, ph_it, ph_bis) = b32.phase_stats()
    f32 = [b32.fields(i) for i in range(N_BATCH)]
    b32.close()
    dE = max(abs(a.TotalEnergy - b.TotalEnergy) / abs(b.TotalEnergy) for a, b in zip(r32, r64))
    dmu = max(abs(a.muVec[s] - b.muVec[s]) / abs(b.muVec[s]) for a, b in zip(r32, r64) for s in range(a.S))
    dn = max(float(np.max(np.abs(a[0] - b[0])) / np.max(np.abs(b[0]))) for a, b in zip(f32, f64))
    out = {
        "workload": "C3: 512 2D two-species TF + contact gases, seed 20260101, grids 64/128/256 (p 0.5/0.3/0.2), "
                    "box 2.5 sqrt(2N), Pulay 8, theta 0.1, max|dV|<1e-8",
        "fp64_batch": {"solves_per_s": round(N_BATCH * steps / elapsed, 2), "seconds_per_batch": round(elapsed / steps, 4),
                       "converged": sum(r.converged for r in r64), "sc_iterations": int(sum(r.SCcount for r in r64)),
                       "per_grid": per_class},
        "fp32_polish_batch": {"solves_per_s": round(N_BATCH / t32, 2), "seconds_per_batch": round(t32, 4),
                              "converged": sum(r.converged for r in r32),
                              "fp32_seconds": round(ph_s[0], 4), "polish_seconds": round(ph_s[1], 4),
                              "fp32_kernel_ms": round(ms32[0], 3), "polish_kernel_ms": round(ms32[1], 3),
                              "fp32_algorithmic_GBps": round(by32[0] / (ms32[0] / 1e3) / 1e9, 1) if ms32[0] else None,
                              "fp32_iterations": int(ph_it[0]), "polish_iterations": int(ph_it[1]),
                              "fp32_bisection_steps": int(ph_bis[0]), "polish_bisection_steps": int(ph_bis[1]),
                              "max_rel_dE_vs_fp64": dE, "max_rel_dmu_vs_fp64": dmu,
                              "max_dn_over_max_n_vs_fp64": dn},
    }
    if oracle_energies:
        eo = max(abs(r32[int(i)].TotalEnergy - e) / abs(e) for i, e in oracle_energies.items())
        e64 = max(abs(r64[int(i)].TotalEnergy - e) / abs(e) for i, e in oracle_energies.items())
        out["fp32_polish_batch"]["max_rel_dE_vs_oracle"] = eo
        out["fp64_batch"]["max_rel_dE_vs_oracle"] = e64
        out["oracle_instances"] = len(oracle_energies)
    # the plain FP64 GPU loop: one context per instance, one after another
    plain_n = min(plain_n, N_BATCH)
    if plain_n > 0:
        t_plain = 0.0
        match = 0
        for i in range(plain_n):
            ctx = Context(insts[i])
            t0 = time.perf_counter()
            r = ctx.solve()
            t_plain += time.perf_counter() - t0
            ctx.close()
            match += int(abs(r.TotalEnergy - r64[i].TotalEnergy) <= 1e-10 * abs(r.TotalEnergy))
            if i % 128 == 127:
                print("bench: plain loop %d/%d" % (i + 1, plain_n), file=sys.stderr, flush=True)
        plain_rate = plain_n / t_plain
        out["plain_fp64_loop"] = {"instances": plain_n, "seconds": round(t_plain, 3), "solves_per_s": round(plain_rate, 2),
                                  "energy_within_1e-10_of_fp64_batch": match}
        out["speedup_fp64_batch_vs_plain_fp64"] = round(N_BATCH * steps / elapsed / plain_rate, 2)
        out["speedup_fp32_polish_vs_plain_fp64"] = round(N_BATCH / t32 / plain_rate, 2)
    return N_BATCH * steps / elapsed, elapsed / steps, roof, out


def c1_leg(grid, maxsc):
    """One converged drop-in solve: nAiT kernel roofline and stage HBM rates."""
    from mpdpft_amd import Context, configs
    p = configs.c1(steps=grid, density_expression=5, hartree_mode=1, maxsc=maxsc)
    ctx = Context(p)
    ctx.enable_timing(True)
    t0 = time.perf_counter()
    res = ctx.solve()
    t = time.perf_counter() - t0
    stats = ctx.stage_stats()
    ctx.close()
    nait = stats.get("nait_density", {})
    roof = None
    if nait.get("launches"):
        avg_s = nait["total_ms"] / nait["launches"] / 1e3
        fl = nait["flops"] / nait["launches"]
        ach = fl / avg_s / 1e12
        traffic = read_summary_value(os.path.join(ROOT, "profiles", "nait_lanes_128_traffic.md"),
                                     "traffic_bytes_per_launch") if grid == 128 else None
        roof = {"bound": "valu", "achieved": round(ach, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(ach / FP64_PEAK_TFLOPS, 5), "traffic": traffic, "kernel": "k_nait_lanes",
                "avg_launch_ms": round(avg_s * 1e3, 4), "flops_per_launch": fl,
                "note": "FP64 vector (VALU) bound, no MFMA: integrand evaluations performed x 50 nominal FP64 flops"}
    hbm = {}
    for name in ("delfield", "hartree_isolated", "update_v", "tf_density", "pulay_gram", "pulay_mix"):
        st = stats.get(name)
        if st and st["launches"] and st["total_ms"] > 0:
            gbs = st["bytes"] / (st["total_ms"] / 1e3) / 1e9
            hbm[name] = {"GB/s": round(gbs, 1), "pct_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
                         "avg_ms": round(st["total_ms"] / st["launches"], 4), "launches": st["launches"]}
    return {
        "workload": "C1 drop-in: 3D nAiT(T=0.05) + isolated Hartree (zero-padded %d^3) + Dirac + PW92, %d^3 on a "
                    "20 bohr cube, Z=10 a=0.5 N=10, Pulay 10, theta 0.05, max|dV|<1e-8" % (2 * grid, grid),
        "seconds": round(t, 2), "solves_per_s": round(1.0 / t, 5), "converged": bool(res.converged),
        "sc_iterations": res.SCcount - 1, "density_evaluations": res.bisection_steps + p.S * (1 + res.SCcount),
        "total_energy": res.TotalEnergy, "mu": res.muVec[0], "nait_roofline": roof, "stages_hbm": hbm,
        "stage_ms": {k: round(v["total_ms"], 2) for k, v in stats.items() if v["launches"]},
    }


def adjoint_leg(steps, h, n_coords, deadline):
    """C4: converged solve, adjoint gradient of the density-mismatch loss and the
    total-energy forces for all 96 centre coordinates, central-difference checks."""
    import numpy as np
    from mpdpft_amd import Context, configs
    pos = configs.centre_positions(32, 424242)
    p = configs.c4(steps=steps, tol=1e-10, positions=pos)
    pt = configs.c4(steps=steps, tol=1e-10, positions=configs.centre_positions(32, 434343))
    EL = steps + 1
    w1 = np.ones(EL)
    w1[0] = w1[-1] = 0.5
    w = (w1[:, None, None] * w1[None, :, None] * w1[None, None, :]).reshape(-1) * (p.edge / steps) ** 3
    ct = Context(pt)
    ct.solve()
    nt = ct.density()[0].copy()
    ct.close()
    print("bench: adjoint target solved", file=sys.stderr, flush=True)
    ctx = Context(p)
    t0 = time.perf_counter()
    r = ctx.solve()
    t_solve = time.perf_counter() - t0
    n = ctx.density()[0].copy()
    V = ctx.potential().copy()
    t0 = time.perf_counter()
    lam, info = ctx.adjoint(2.0 * w * (n - nt))
    grad = ctx.centre_gradient(lam)
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    forces = ctx.centre_forces()
    t_forces = time.perf_counter() - t0
    ctx.close()
    C = np.array([[q[0], q[1], q[2], 1.0, 0.6] for q in pos])
    coords = [(0, 0), (5, 1), (11, 2), (17, 0), (23, 1), (31, 2)][:n_coords]
    checks = []
    fd = Context(configs.c4(steps=steps, tol=1e-12, maxsc=500, positions=pos))
    for ci, d in coords:
        if time.perf_counter() > deadline:
            break
        vals = []
        for sgn in (1, -1):
            Cp = C.copy()
            Cp[ci, d] += sgn * h
            fd.set_centres(Cp)
            fd.set_initial_potential(V)
            rr = fd.solve()
            vals.append((float(np.sum(w * (fd.density()[0] - nt) ** 2)), rr.TotalEnergy, bool(rr.converged)))
        fdL = (vals[0][0] - vals[1][0]) / (2 * h)
        fdE = (vals[0][1] - vals[1][1]) / (2 * h)
        print("bench: adjoint fd check centre %d %s done" % (ci, "xyz"[d]), file=sys.stderr, flush=True)
        checks.append({"centre": ci, "axis": "xyz"[d], "dL_adjoint": float(grad[ci, d]), "dL_fd": fdL,
                       "dL_rel_err": abs(grad[ci, d] - fdL) / abs(fdL), "dE_adjoint": float(forces[ci, d]),
                       "dE_fd": fdE, "dE_rel_err": abs(forces[ci, d] - fdE) / abs(fdE),
                       "fd_solves_converged": all(v[2] for v in vals)})
    fd.close()
    return {
        "workload": ("C4: %d^3 nAiT(T=0.05) + isolated Hartree + Dirac + PW92 on a 30 bohr cube, 32 softened Coulomb "
                     "centres (Z=1, a=0.6, seed 424242), N=32, max|dV|<1e-10; loss int (n - n_target)^2 (target "
                     "seed 434343) and total energy vs the 96 centre coordinates" % steps),
        "solve_seconds": round(t_solve, 3), "sc_iterations": r.SCcount - 1, "converged": bool(r.converged),
        "total_energy": r.TotalEnergy,
        "adjoint_seconds": round(t_adj, 4), "gmres_iterations": info["iterations"],
        "gmres_relative_residual": info["relative_residual"], "forces_seconds": round(t_forces, 4),
        "gradient_norm": float(np.linalg.norm(grad)), "force_norm": float(np.linalg.norm(forces)),
        "fd_h": h, "fd_checks": checks,
        "max_dL_rel_err": max((c["dL_rel_err"] for c in checks), default=None),
        "max_dE_rel_err": max((c["dE_rel_err"] for c in checks), default=None),
        "fd_gradient_cost_estimate_s": round(2 * 96 * t_solve, 1),
    }


def copy_leg():
    """A plain device copy of known size (256 MiB of FP64 read, 256 MiB
    written, three times): the calibration kernel for the FETCH_SIZE and
    WRITE_SIZE counters of a rocprof counter run."""
    import torch
    n = 1 << 25
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"bytes_read_per_launch": 8 * n, "bytes_written_per_launch": 8 * n, "launches": 3}


def main():
    t_begin = time.perf_counter()
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batched solves")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--deadline", type=float, default=450.0, help="seconds after which no further leg starts")
    ap.add_argument("--legs", default="cpu,batch,c1,adjoint",
                    help="comma-separated subset of cpu, batch, plain, c1, adjoint, copy (counter calibration)")
    ap.add_argument("--plain", type=int, default=512, help="instances the plain FP64 loop solves")
    ap.add_argument("--c1-grid", type=int, default=128)
    ap.add_argument("--maxsc", type=int, default=300)
    ap.add_argument("--adjoint-grid", type=int, default=96)
    ap.add_argument("--fd-coords", type=int, default=6)
    ap.add_argument("--cpu-budget", type=float, default=40.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the batched instances would shard across GPUs by queue share, which is "
              "not built; --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    deadline = t_begin + args.deadline
    legs = set(args.legs.split(","))
    os.environ.setdefault("MPDPFT_VERBOSE", "0")

    skipped, errors = {}, {}
    value = ms_per_step = None
    roof = cpu = batch = c1 = adjoint = copy = None  # noqa: F841

    def run(name, est_s, fn):
        if name not in legs:
            skipped[name] = "not requested"
            return None
        left = deadline - time.perf_counter()
        if left < est_s:
            skipped[name] = "deadline: %.0f s left, leg needs about %.0f s" % (left, est_s)
            return None
        print("bench: leg %s (%.0f s elapsed)" % (name, time.perf_counter() - t_begin), file=sys.stderr, flush=True)
        try:
            return fn()
        except Exception as e:  # noqa: BLE001
            errors[name] = "%s: %s" % (type(e).__name__, e)
            traceback.print_exc()
            return None

    try:
        # the host legs run before the GPU is touched (child processes only)
        cpu = run("cpu", 20, lambda: cpu_baseline_leg(args.cpu_budget))
        oracle_E = cpu.pop("oracle_energies", None) if cpu else None
        got = run("batch", 30, lambda: batch_leg(args.steps, args.warmup,
                                                 args.plain if "plain" in legs or "batch" in legs else 0, oracle_E))
        if got:
            value, sec, roof, batch = got
            ms_per_step = 1e3 * sec
        c1 = run("c1", 70, lambda: c1_leg(args.c1_grid, args.maxsc))
        adjoint = run("adjoint", 60, lambda: adjoint_leg(args.adjoint_grid, 1e-4, args.fd_coords, deadline))
        copy = run("copy", 5, copy_leg) if "copy" in legs else None
    finally:
        if _PARAMS_FILE and os.path.exists(_PARAMS_FILE):
            os.unlink(_PARAMS_FILE)
        line = {
            "metric": METRIC,
            "value": round(value, 2) if value else None,
            "unit": "solves/s",
            "n_gpus": 1,
            "steps": args.steps,
            "warmup": args.warmup,
            "ms_per_step": round(ms_per_step, 3) if ms_per_step else None,
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": "f64",
            "data": "synthetic",
            "config": {"workload": "C3 batched: 512 seeded 2D two-species TF + contact gases (64^2/128^2/256^2), "
                                   "FP64, one persistent launch per solve", "instances": N_BATCH, "seed": 20260101},
            "roofline": roof,
            "cpu_baseline": cpu,
            "batch": batch,
            "c1": c1,
            "adjoint": adjoint,
            "skipped_legs": {k: v for k, v in skipped.items() if k != "copy"},
            "leg_errors": errors,
            "bench_seconds": round(time.perf_counter() - t_begin, 1),
        }
        print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
