#!/usr/bin/env python3
"""Converged self-consistent DPFT solves per second on one MI355X.

Workload (the FP64 GPU drop-in configuration): spin-unpolarised electrons,
N = 10, in the softened Coulomb well -Z/sqrt(r^2 + a^2) (Z = 10, a = 0.5 bohr)
on a 128^3 grid over a 20 bohr cube; density from the Airy-averaged
quantum-corrected functional (nAiT, T = 0.05 Ha) with |grad V|^2 and lap V
from FFT derivatives smoothed by a k-space Gaussian of width 0.3 bohr^-1;
Hartree by an isolated-boundary convolution on the zero-padded 256^3 grid;
Dirac exchange + PW92 correlation; Pulay history 10, mixing 0.05, converged
when max|V - V_old| < 1e-8 for three consecutive iterations.

A step is one converged solve from V = V_ext.  The JSON line carries the
roofline of the dominant kernel (the nAiT density kernel, FP64) timed with
HIP events on the solver's stream, the HBM rates of the FFT stages, and a
CPU baseline from the oracle on a bounded sample.

The "batch" object adds the batched workload: 512 seeded 2D two-species TF +
contact gases (grids 64^2/128^2/256^2, seed 20260101) solved in one persistent
launch, in FP64 and as FP32 with an FP64 polish, against the plain FP64 GPU
loop solving the same instances one after another; the polish results are
compared with the FP64 results (energy, density, mu) and, on a sample, with
the oracle, which is also timed there as the CPU baseline of that workload.

The "adjoint" object adds the differentiation workload: a 96^3 cluster of 32
softened Coulomb centres (Z = 1, a = 0.6, seed 424242) on a 30 bohr cube with
N = 32 and the functional above; the loss int (n - n_target)^2 (target: the
centres of seed 434343) and the total energy are differentiated with respect to
the 96 centre coordinates by the implicit adjoint of the converged fixed point,
and one coordinate of each is checked against central differences (h = 2e-4,
which keeps the O(h^2) truncation near 2e-8 relative while halving the effect
of the solves' stopping noise over 2h) of solves started from the converged
potential and stopped at max|dV| < 2e-11.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--grid 128]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "converged SC solves/s vs FP64 GPU at |ΔE|<1e-6; FFT/pointwise HBM GB/s % peak"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def host_cpu():
    """Model name of the host CPU (first entry of /proc/cpuinfo) and os.cpu_count()."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    name = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return "%s (%s logical CPUs visible)" % (name, os.cpu_count())


def cpu_baseline(p, mu, V, density_evals, budget_s):
    """Oracle (plain C, OpenMP over all host threads) timing of one nAiT
    density evaluation at the solve's own grid and converged potential,
    scaled by the number of density evaluations one converged solve makes."""
    import numpy as np
    from oracle import oracle as O
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    O.lib()
    v = np.ascontiguousarray(V[0])
    t0 = time.perf_counter()
    _, N = O.density(p, 0, mu, v)
    t_eval = time.perf_counter() - t0
    solve_s = t_eval * density_evals
    return {
        "value": 1.0 / solve_s,
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": ("one nAiT density evaluation (DelField + per-point adaptive Airy quadrature) of the oracle at "
                   "%d^3 on the converged potential took %.2f s; a solve makes %d such evaluations "
                   "(mu bracketing/bisection included) -> %.0f s per solve; host: %s" % (
                       p.steps, t_eval, density_evals, solve_s, host_cpu())),
    }


def batch_leg(n, plain_n, cpu_n, cpu_on):
    """C3: the batched solver (FP64; FP32 + FP64 polish) against the plain FP64
    loop on the same instances; errors of the polished results."""
    import numpy as np
    from mpdpft_amd import Batch, Context, configs
    insts = configs.c3(n=n)
    Batch(insts[:4]).solve()  # loads the kernels
    b64 = Batch(insts)
    r64 = b64.solve()
    t64 = b64.seconds()
    f64 = [b64.fields(i) for i in range(n)]
    b64.close()
    b32 = Batch(insts, fp32=True)
    r32 = b32.solve()
    t32 = b32.seconds()
    (ph_s, ph_it, ph_bis) = b32.phase_stats()
    f32 = [b32.fields(i) for i in range(n)]
    b32.close()
    dE = max(abs(a.TotalEnergy - b.TotalEnergy) / abs(b.TotalEnergy) for a, b in zip(r32, r64))
    dmu = max(abs(a.muVec[s] - b.muVec[s]) / abs(b.muVec[s]) for a, b in zip(r32, r64) for s in range(a.S))
    dn = max(float(np.max(np.abs(a[0] - b[0])) / np.max(np.abs(b[0]))) for a, b in zip(f32, f64))
    plain_n = min(plain_n, n)
    t_plain = 0.0
    plain_match = 0
    for i in range(plain_n):
        ctx = Context(insts[i])
        t0 = time.perf_counter()
        r = ctx.solve()
        t_plain += time.perf_counter() - t0
        ctx.close()
        plain_match += int(abs(r.TotalEnergy - r64[i].TotalEnergy) <= 1e-10 * abs(r.TotalEnergy)
                           and r.SCcount == r64[i].SCcount)
    plain_rate = plain_n / t_plain
    out = {
        "workload": "C3: %d 2D two-species TF + contact gases, seed 20260101, grids 64/128/256 (p 0.5/0.3/0.2), "
                    "box 2.5 sqrt(2N), Pulay 8, theta 0.1, max|dV|<1e-8" % n,
        "instances": n,
        "grid_points": int(sum(p.S * (p.steps + 1) ** 2 for p in insts)),
        "fp64_batch": {"seconds": round(t64, 4), "solves_per_s": round(n / t64, 2),
                       "converged": sum(r.converged for r in r64),
                       "sc_iterations": int(sum(r.SCcount for r in r64))},
        "fp32_polish_batch": {"seconds": round(t32, 4), "solves_per_s": round(n / t32, 2),
                              "converged": sum(r.converged for r in r32),
                              "fp32_seconds": round(ph_s[0], 4), "polish_seconds": round(ph_s[1], 4),
                              "fp32_iterations": int(ph_it[0]), "polish_iterations": int(ph_it[1]),
                              "fp32_bisection_steps": int(ph_bis[0]), "polish_bisection_steps": int(ph_bis[1]),
                              "max_rel_dE_vs_fp64": dE, "max_rel_dmu_vs_fp64": dmu,
                              "max_dn_over_max_n_vs_fp64": dn},
        "plain_fp64_loop": {"instances": plain_n, "seconds": round(t_plain, 4),
                            "solves_per_s": round(plain_rate, 2),
                            "identical_to_fp64_batch": plain_match},
        "speedup_fp32_polish_vs_plain_fp64": round((n / t32) / plain_rate, 2),
        "speedup_fp64_batch_vs_plain_fp64": round((n / t64) / plain_rate, 2),
    }
    if cpu_on:
        from oracle import oracle as O
        O.lib()
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        t_cpu = 0.0
        eE = en = emu = 0.0
        cpu_n = min(cpu_n, n)
        for i in range(cpu_n):
            t0 = time.perf_counter()
            _, r, den, _ = O.solve(insts[i])
            t_cpu += time.perf_counter() - t0
            eE = max(eE, abs(r32[i].TotalEnergy - r.TotalEnergy) / abs(r.TotalEnergy))
            emu = max(emu, max(abs(r32[i].muVec[s] - r.muVec[s]) / abs(r.muVec[s]) for s in range(r.S)))
            en = max(en, float(np.max(np.abs(f32[i][0] - den)) / np.max(np.abs(den))))
        out["fp32_polish_batch"].update({"max_rel_dE_vs_oracle": eE, "max_rel_dmu_vs_oracle": emu,
                                         "max_dn_over_max_n_vs_oracle": en, "oracle_instances": cpu_n})
        out["cpu_baseline"] = {"value": round(cpu_n / t_cpu, 3), "unit": "solves/s", "cores": cores,
                               "kind": "port",
                               "sample": "the oracle's SC loop on the first %d instances, one after another "
                                         "(OpenMP over %d threads): %.2f s; host: %s" % (cpu_n, cores, t_cpu,
                                                                                         host_cpu())}
    return out


def adjoint_leg(steps, h):
    """C4: converged solve, adjoint gradient of the density-mismatch loss and
    total-energy forces for all centre coordinates, FD checks."""
    import numpy as np
    from mpdpft_amd import Context, configs
    os.environ["MPDPFT_VERBOSE"] = "0"
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
    # central differences from solves stopped at max|dV| < 2e-11 (their stopping noise over 2h is the
    # difference's error), each started from the converged V
    C = np.array([[q[0], q[1], q[2], 1.0, 0.6] for q in pos])
    ci, d = 0, 0
    vals = []
    fd = Context(configs.c4(steps=steps, tol=2e-11, maxsc=500, positions=pos))
    for sgn in (1, -1):
        Cp = C.copy()
        Cp[ci, d] += sgn * h
        fd.set_centres(Cp)
        fd.set_initial_potential(V)
        rr = fd.solve()
        vals.append((float(np.sum(w * (fd.density()[0] - nt) ** 2)), rr.TotalEnergy, bool(rr.converged)))
    fd.close()
    fdL = (vals[0][0] - vals[1][0]) / (2 * h)
    fdE = (vals[0][1] - vals[1][1]) / (2 * h)
    return {
        "workload": "C4: 96^3 nAiT(T=0.05) + isolated Hartree + Dirac + PW92 on a 30 bohr cube, 32 softened "
                    "Coulomb centres (Z=1, a=0.6, seed 424242), N=32, max|dV|<1e-10; loss int (n - n_target)^2 "
                    "(target seed 434343) and total energy vs the 96 centre coordinates" if steps == 96 else
                    "C4 geometry at %d^3" % steps,
        "solve_seconds": round(t_solve, 3), "sc_iterations": r.SCcount - 1, "converged": bool(r.converged),
        "total_energy": r.TotalEnergy,
        "adjoint_seconds": round(t_adj, 4), "gmres_iterations": info["iterations"],
        "gmres_relative_residual": info["relative_residual"], "forces_seconds": round(t_forces, 4),
        "gradient_norm": float(np.linalg.norm(grad)), "force_norm": float(np.linalg.norm(forces)),
        "fd_check": {"coordinate": "centre 0 x", "h": h, "dL_adjoint": float(grad[ci, d]), "dL_fd": fdL,
                     "dL_rel_err": abs(grad[ci, d] - fdL) / abs(fdL), "dE_adjoint": float(forces[ci, d]),
                     "dE_fd": fdE, "dE_rel_err": abs(forces[ci, d] - fdE) / abs(fdE),
                     "fd_solves_converged": all(v[2] for v in vals)},
        "fd_gradient_cost_estimate_s": round(2 * 96 * t_solve, 1),
    }


def precision_leg(steps):
    """C2: the same solve in FP64 and as Precision 1 (FP32 transforms and pointwise
    potentials until the FP32 floor, then FP64 polish); errors and times."""
    import numpy as np
    from mpdpft_amd import Context, configs
    out = {"workload": "C2: %d^3 nAiT(T=0.05) + isolated Hartree + Dirac + PW92 on a 24 bohr cube, Z=18, a=0.4, "
                       "N=18, max|dV|<1e-8" % steps}
    res = {}
    for prec in (0, 1):
        ctx = Context(configs.c2(steps=steps, precision=prec))
        ctx.enable_timing(True)
        t0 = time.perf_counter()
        r = ctx.solve()
        t = time.perf_counter() - t0
        res[prec] = (r, ctx.density()[0].copy(), t)
        stats = ctx.stage_stats()
        ctx.close()
        out["fp64" if prec == 0 else "precision1"] = {
            "seconds": round(t, 2), "sc_iterations": r.SCcount - 1, "fp32_iterations": r.fp32_iterations,
            "converged": bool(r.converged), "total_energy": r.TotalEnergy, "mu": r.muVec[0],
            "stage_ms": {k: round(v["total_ms"], 1) for k, v in stats.items() if v["launches"]}}
    (r0, n0, _), (r1, n1, _) = res[0], res[1]
    out["rel_dE"] = abs(r1.TotalEnergy - r0.TotalEnergy) / abs(r0.TotalEnergy)
    out["rel_dmu"] = abs(r1.muVec[0] - r0.muVec[0]) / abs(r0.muVec[0])
    out["max_dn_over_max_n"] = float(np.abs(n1 - n0).max() / np.abs(n0).max())
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--maxsc", type=int, default=300)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=30.0)
    ap.add_argument("--batch", type=int, default=512, help="instances of the batched workload (0: skip)")
    ap.add_argument("--batch-plain", type=int, default=512,
                    help="instances the plain FP64 loop solves for the batch comparison")
    ap.add_argument("--batch-cpu", type=int, default=16, help="instances the oracle solves for the batch baseline")
    ap.add_argument("--adjoint-grid", type=int, default=96, help="grid of the differentiation workload (0: skip)")
    ap.add_argument("--precision-grid", type=int, default=0,
                    help="grid of the precision-variant comparison (256 = C2; 0: skip, about 15 min at 256)")
    ap.add_argument("--traffic", default=os.path.join(ROOT, "profiles", "nait_lanes_128_traffic.md"),
                    help="rocprof summary holding 'traffic_bytes_per_launch = N' for the nAiT kernel at this grid")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported (the batched and sharded "
              "layouts are not built)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    os.environ.setdefault("MPDPFT_VERBOSE", "1")  # one line per SC iteration on stderr
    from mpdpft_amd import Context, configs

    p = configs.c1(steps=args.grid, density_expression=5, hartree_mode=1, maxsc=args.maxsc)
    ctx = Context(p)
    for _ in range(args.warmup):
        ctx.solve()
    ctx.enable_timing(True)  # resets the stage statistics
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = ctx.solve()  # returns after the stream has drained
    t1 = time.perf_counter()
    elapsed = t1 - t0
    stats = ctx.stage_stats()
    V = ctx.potential()

    nait = stats.get("nait_density", {})
    roof = None
    if nait.get("launches"):
        avg_s = nait["total_ms"] / nait["launches"] / 1e3
        flops_per_launch = nait["flops"] / nait["launches"]
        achieved = flops_per_launch / avg_s / 1e12
        traffic = None
        if args.grid == 128 and args.traffic and os.path.exists(args.traffic):
            for ln in open(args.traffic):
                if ln.startswith("traffic_bytes_per_launch"):
                    traffic = float(ln.split("=")[1])
        roof = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 5), "traffic": traffic,
                "kernel": "k_nait_lanes" if p.Mpp[0] <= 15 else "k_nait_reuse",
                "avg_launch_ms": round(avg_s * 1e3, 4),
                "flops_per_launch": flops_per_launch,
                "note": "FP64 vector (VALU) compute bound, no MFMA: integrand evaluations actually performed "
                        "x 50 nominal FP64 flops each, against the 78.6 TFLOP/s FP64 vector peak"}
    hbm = {}
    stages = {k: {"launches": v["launches"], "total_ms": round(v["total_ms"], 2)} for k, v in stats.items()}
    for name in ("delfield", "hartree_isolated", "update_v", "tf_density"):
        st = stats.get(name)
        if st and st["launches"] and st["total_ms"] > 0:
            gbs = st["bytes"] / (st["total_ms"] / 1e3) / 1e9
            hbm[name] = {"GB/s": round(gbs, 1), "pct_peak": round(100 * gbs / HBM_PEAK_GBS, 2),
                         "avg_ms": round(st["total_ms"] / st["launches"], 4), "launches": st["launches"]}
    if "update_v" in hbm:  # Dirac exchange and PW92 per point: transcendental (pow, log) bound, not HBM
        hbm["update_v"]["note"] = "FP64 transcendental bound (Dirac pow, PW92 log/pow per point)"

    density_evals = res.bisection_steps + p.S * (1 + res.SCcount)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(p, res.muVec[0], V, density_evals, args.cpu_budget)
        except Exception as e:  # noqa: BLE001
            cpu = {"value": None, "error": str(e)}

    batch = None
    if args.batch > 0:
        os.environ["MPDPFT_VERBOSE"] = "0"  # 512 plain solves: no per-iteration lines
        print("bench: batched workload, %d instances" % args.batch, file=sys.stderr, flush=True)
        batch = batch_leg(args.batch, args.batch_plain, args.batch_cpu, not args.no_cpu_baseline)

    precision = None
    if args.precision_grid > 0:
        print("bench: precision variant, %d^3" % args.precision_grid, file=sys.stderr, flush=True)
        precision = precision_leg(args.precision_grid)

    adjoint = None
    if args.adjoint_grid > 0:
        print("bench: differentiation workload, %d^3" % args.adjoint_grid, file=sys.stderr, flush=True)
        adjoint = adjoint_leg(args.adjoint_grid, 2e-4)

    line = {
        "metric": METRIC,
        "value": args.steps / elapsed,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "C1 drop-in: 3D nAiT(T=0.05) + isolated Hartree (zero-padded %d^3) + Dirac + PW92, "
                               "%d^3 on a 20 bohr cube, Z=10 a=0.5 N=10, Pulay 10, theta 0.05, max|dV|<1e-8"
                               % (2 * args.grid, args.grid),
                   "grid": args.grid, "converged": bool(res.converged), "sc_iterations": res.SCcount - 1,
                   "density_evaluations": density_evals, "total_energy": res.TotalEnergy, "mu": res.muVec[0]},
        "roofline": roof,
        "cpu_baseline": cpu,
        "stages_hbm": hbm,
        "stage_ms": stages,
        "batch": batch,
        "adjoint": adjoint,
        "precision": precision,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
