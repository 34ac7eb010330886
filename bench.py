#!/usr/bin/env python3
"""Time-to-converged FLEX self-energy on one MI355X: mixed precision + warm
start versus the plain FP64 cold start, same GPU, same workload.

Workload (default): the mixed-precision FLEX lattice -- 4-orbital synthetic 3D
tight binding (seeded real hoppings to 2nd in-plane neighbours, amplitudes in
[-0.4, 0.4] eV, interlayer x0.2), 32x32x16 k-mesh, T = 0.015 eV, 1024 fermionic
Matsubara frequencies, Kanamori U = 1.0 eV, J = 0.2 eV, 0.9 electrons per
orbital.  Both arms run the reference's loop (libs/src/fself.f90:287-448):
linear mixing pp = 0.5, m_diis = 1, sub_sigma = 1, sw_rescale, stop when
esterr < 1e-8.

  plain arm : FP64 everywhere, cold start G0 at the non-interacting mu.
  mixed arm : FP32 transforms of rolling increments around FP64 chi0 / Sigma
              accumulators (vertex LU, Dyson, mu search, mixing and the
              convergence test in FP64), started from the warm-start Sigma.

A step is one timed pair of solves.  `value` is the speedup
t_plain / t_mixed of the device-resident solve times (inputs in HBM, the
copies of the host ABI excluded; the PCIe-inclusive wall times are reported
beside them).  The dominant kernel's roofline comes from hipEvents around
every FFT pass on the library stream during the timed mixed solve; a pass's
algorithmic bytes are what its fused load and store must move: a full grid
read or written counts N b (N grid elements, b bytes each), an irreducible
(k, w) slice counts its own size, so a plain pass is 2 N b and each fused
grid operand adds N b.

The CPU baseline is the numpy/scipy restatement of the reference (oracle/)
on the host cores: every component of a reference FLEX iteration timed on a
bounded sample of this workload and scaled to one iteration, times the plain
arm's iteration count; beside it, one whole oracle iteration of the 1-orbital
64x64 x 1024 workload, measured.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 3]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "time-to-converged FLEX Σ (tol 1e-8): FP32+warm-start vs plain FP64, same GPU"
HBM_PEAK_GBS = 8000.0
# both arms compute chi0 with the FP64 Matsubara tail correction
# (libs/src/fchi.f90:852-892; sw_tail as main.py:165 sets it)
SW_TAIL = True
# HBM bytes per launch of the FP32 convolution kernels (box_fft_kernel,
# box_accum_kernel and the Sigma increment's contraction kernel) on
# the 4-orbital 32x32x16 x 1024-Matsubara workload (tail-corrected chi0, warm
# start), averaged over the dispatches of the timed mixed solve: rocprofv3
# FETCH_SIZE x 2 + WRITE_SIZE (KiB x 1024), both calibrated on a device copy of
# known size (FETCH_SIZE reads half the bytes on this device), per kernel
# class, weighted by the classes' dispatch counts in the full solve.  Measured
# offline: counter runs cannot be part of the timed run.
FP32_BOX_TRAFFIC_BYTES = {3: 1357645944}


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def solve(F, L, st, precision, eps, scf_loop=300, seed=None, profile=False):
    """One mkself solve; returns (sigma, mu, stats, profile or None)."""
    F.set_precision(precision)
    sw_in = seed is not None
    if sw_in:
        F.set_sigma_seed(*seed)
    if profile:
        F.profile(True)
    t0 = time.perf_counter()
    sig, mu = L.flex(st, eps=eps, pp=0.5, m_diis=1, scf_loop=scf_loop, sw_tail=SW_TAIL, sw_rescale=True,
                     sub_sigma=1, sw_in=sw_in)
    wall = time.perf_counter() - t0
    prof = None
    if profile:
        prof = F.profile_read()
        F.profile(False)
    stats = F.flex_stats()
    stats["wall_seconds"] = wall
    return sig, mu, stats, prof


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(st, iters_plain, threads):
    """The oracle (numpy/scipy restatement of the reference, oracle/) timed on
    the host cores on bounded samples of this workload, every component of a
    reference FLEX iteration with the tail correction (libs/src/fself.f90:361-430):
      * the bubble unit of fchi.f90:559-611 (scatter, 2 forward 4D FFTs,
        product with the (-r,-tau) reversal, inverse FFT, gather) on the full
        complex128 grid; an iteration runs 3 P of them for chi0 (conv[G],
        conv[G0] and the tau reference, fchi.f90:852-892) and Nchi^2 for
        Sigma (fself.f90:199-269);
      * the vertex (fUJ.f90:318-365) on a sample of the (q, i nu) points,
        scaled to all of them;
      * Dyson (fself.f90:43-157) and the eigenvalues of the mu search
        (fself.f90:716-740) on a sample of k points with all frequencies,
        scaled to all k, plus 30 filling evaluations (the Brent search's
        typical count) likewise scaled;
    and, measured whole, one oracle FLEX iteration (the final mu search
    included) of the 1-orbital 64x64 x 1024 workload."""
    import oracle.pyrpa_oracle as O
    O.FFT_WORKERS = threads
    Nx, Ny, Nz, Nw = st["Nx"], st["Ny"], st["Nz"], st["Nw"]
    No = st["Norb"]
    Nchi = No * No
    P = Nchi * (Nchi + 1) // 2
    eig, uni = st["eig"], st["uni"]  # (Nk, No), (Nk, No, No) C views of Fortran eig(No,Nk), uni(No,No,Nk)
    Nk = eig.shape[0]
    parts = {}
    # bubble unit (G0_00 as gen_green0 builds it, fself.f90:1-41)
    wn = (2.0 * np.arange(Nw) + 1.0) * np.pi * st["temp"]
    wts = np.abs(uni[:, :, 0]) ** 2
    g00 = (wts[:, None, :] / (1j * wn[None, :, None] + st["mu"] - eig[:, None, :])).sum(axis=2)
    Gf = g00[:, :, None, None]
    t0 = time.perf_counter()
    g13 = O.fft4(O._full_green(Gf, 0, 0, st["kmap"], st["invk"], Nx, Ny, Nz))
    g42 = O.fft4(O._full_green(Gf, 0, 0, st["kmap"], st["invk"], Nx, Ny, Nz))
    tmp = O.ifft4(-g13 * O._reverse(g42))
    O._gather(tmp, st["kmap"], st["invk"], Nk, Nw)
    t_unit = time.perf_counter() - t0
    del g13, g42, tmp, Gf, g00
    parts["bubble_units"] = (3 * P + Nchi * Nchi) * t_unit
    # vertex on a sample of (q, i nu) points
    rng = np.random.default_rng(1)
    ns = min(Nk * Nw, 16384)
    chi = 0.05 * (rng.standard_normal((Nchi, Nchi, 1, ns)) + 1j * rng.standard_normal((Nchi, Nchi, 1, ns)))
    t0 = time.perf_counter()
    O.get_Vsigma_nosoc_flex(chi, st["Smat"], st["Cmat"])
    parts["vertex"] = (time.perf_counter() - t0) * (Nk * Nw) / ns
    # Dyson and the mu search on a sample of k points
    ks = min(Nk, 64)
    sig = 0.05 * (rng.standard_normal((No, No, Nw, ks)) + 1j * rng.standard_normal((No, No, Nw, ks)))
    t0 = time.perf_counter()
    O.gen_green(sig, st["hamk"][:ks], st["mu"], st["temp"])
    fl = O._Flex(st["hamk"][:ks], eig[:ks], st["fill"], st["temp"], 1, No, ks, Nw)
    fl.precompute(np.asarray(sig).T)
    for i in range(30):
        fl.get_rn(st["mu"] + 1e-3 * i)
    parts["dyson_and_mu"] = (time.perf_counter() - t0) * Nk / ks
    t_iter = sum(parts.values())
    # one whole oracle FLEX iteration of the 1-orbital 64x64 x 1024 workload
    from pyrpa_amd import lattice as L
    c1 = L.config(1)
    t0 = time.perf_counter()
    O.mkself(c1["Smat"], c1["Cmat"], c1["kmap"], c1["invk"], c1["olist"], c1["hamk"], c1["eig"], c1["uni"],
             c1["mu"], c1["fill"], c1["temp"], c1["Nw"], c1["Nx"], c1["Ny"], c1["Nz"], False, False, eps=0.0,
             pp=0.5, m_diis=1, scf_loop=1, sw_rescale=True, sw_tail=SW_TAIL)
    t_c1 = time.perf_counter() - t0
    O.FFT_WORKERS = 1
    return {"value": round(t_iter * iters_plain, 1), "unit": "s to converge (extrapolated)", "cores": threads,
            "kind": "port", "cpu": _cpu_model(),
            "sample": (f"oracle on {threads} host threads ({_cpu_model()}): one bubble unit of the reference "
                       f"algorithm on the {Nx}x{Ny}x{Nz}x{2 * Nw} complex128 grid {t_unit:.2f} s, x (3P + Nchi^2) = "
                       f"{3 * P + Nchi * Nchi} units; vertex on {ns} of {Nk * Nw} (q, i nu) points; Dyson and the "
                       f"mu search on {ks} of {Nk} k points; {t_iter:.0f} s per FLEX iteration x {iters_plain} "
                       f"iterations of the plain arm. Measured whole: one oracle FLEX iteration of the 1-orbital "
                       f"64x64 x 1024 workload {t_c1:.1f} s"),
            "s_per_iteration": round(t_iter, 1), "s_per_iteration_parts": {k: round(v, 1) for k, v in parts.items()},
            "one_orbital_64x64_1024_s_per_iteration_measured": round(t_c1, 2)}


def fft_engine(st, reps):
    """The transform engine alone (gmod_fft_device) on the workload's full
    (k-mesh x 2 Nw) grid: `reps` in-place forward 4D transforms per precision,
    timed with events on the stream the library launches on (synchronised by
    the entry point).  One pass moves 2 N b bytes (N grid elements of b bytes)."""
    from pyrpa_amd import flibs as F
    dims = (st["Nx"], st["Ny"], st["Nz"], 2 * st["Nw"])
    n = int(np.prod(dims))
    out = {}
    rng = np.random.default_rng(0)
    for name, dt, b in (("fp32", np.complex64, 8), ("fp64", np.complex128, 16)):
        buf = F.DeviceArray.from_host((rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dt))
        tmp = F.DeviceArray(n, dt)
        F.profile(True)
        F.fft_device(buf, tmp, dims, sign=-1, reps=1)       # plan + warm-up
        F.profile(False)
        F.profile(True)
        t0 = time.perf_counter()
        F.fft_device(buf, tmp, dims, sign=-1, reps=reps)
        wall = time.perf_counter() - t0
        prof = F.profile_read()
        F.profile(False)
        pc = prof["fft_pass_fp32" if b == 8 else "fft_pass"]
        passes = pc["launches"] // reps
        gbs = pc["bytes"] / (pc["ms"] * 1e-3) / 1e9
        out[name] = {"passes_per_transform": int(passes), "ms_per_transform": round(1e3 * wall / reps, 4),
                     "us_per_pass": round(1e3 * pc["ms"] / max(pc["launches"], 1), 2),
                     "algorithmic_gbs": round(gbs, 1), "frac_of_hbm": round(gbs / HBM_PEAK_GBS, 4)}
        log(f"fft {name}: {out[name]}")
        buf.free()
        tmp.free()
    return {"metric": "4D FFT pass throughput (algorithmic bytes / pass time)", "grid": list(dims),
            "reps": reps, **out}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", type=int, default=3, help="workload (1..4); the metric is quoted on 3")
    ap.add_argument("--eps", type=float, default=1e-8)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=min(16, os.cpu_count() or 1),
                    help="host threads of the CPU baseline (16: the host CPU share of one GPU)")
    ap.add_argument("--warm", default="coarse", choices=["coarse", "none"], help="warm start of the mixed arm")
    ap.add_argument("--fft", type=int, default=0, metavar="REPS",
                    help="time only the transform engine on the workload's grid (REPS transforms per precision)")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE {world}): the FLEX "
              "path is benchmarked on a single MI355X; sharding across GPUs is a design note, not a run.",
              file=sys.stderr)
        sys.exit(2)

    from pyrpa_amd import flibs as F
    from pyrpa_amd import lattice as L
    from pyrpa_amd import warmstart as W

    t0 = time.perf_counter()
    st = L.config(args.config)
    if args.fft:
        print(json.dumps(fft_engine(st, args.fft)), flush=True)
        return
    log(f"workload {args.config} built in {time.perf_counter() - t0:.1f} s: mesh {st['Nx']}x{st['Ny']}x{st['Nz']} "
        f"Nw={st['Nw']} Norb={st['Norb']} Nk={len(st['klist'])} mu0={st['mu']:.6f}")

    for w in range(args.warmup):
        solve(F, L, st, 0, 0.0, scf_loop=2)
        solve(F, L, st, 1, 0.0, scf_loop=3)
        W.warm_start(st, args.warm, sw_tail=SW_TAIL)
        log(f"warmup {w + 1}/{args.warmup} done")

    steps = []
    for k in range(args.steps):
        sig64, mu64, s64, p64 = solve(F, L, st, 0, args.eps, profile=True)
        log(f"step {k + 1}: plain FP64 cold: {int(s64['iterations'])} iterations, {s64['solve_seconds']:.2f} s "
            f"(wall {s64['wall_seconds']:.2f} s), esterr {s64['esterr']:.2e}")
        tw = time.perf_counter()
        seed, winfo = W.warm_start(st, args.warm, sw_tail=SW_TAIL)
        t_warm = time.perf_counter() - tw
        if seed is not None:
            winfo["seed_rel_err_vs_converged"] = rel(seed[0], sig64)
            log(f"step {k + 1}: warm start {winfo['kind']}: {t_warm:.2f} s, seed max|dSigma|/max|Sigma| vs the "
                f"converged FP64 Sigma {winfo['seed_rel_err_vs_converged']:.2e}")
        sigm, mum, sm, pm = solve(F, L, st, 1, args.eps, seed=seed, profile=True)
        t_mixed = t_warm + sm["solve_seconds"]
        log(f"step {k + 1}: mixed + warm start ({winfo['kind']}, {t_warm:.2f} s): {int(sm['iterations'])} "
            f"iterations ({int(sm['increment_iterations'])} FP32-increment), {sm['solve_seconds']:.2f} s, "
            f"esterr {sm['esterr']:.2e}; sigma rel err {rel(sigm, sig64):.2e}, |dmu| {abs(mum - mu64):.2e}")
        steps.append(dict(s64=s64, sm=sm, p64=p64, pm=pm, t_warm=t_warm, t_mixed=t_mixed, winfo=winfo,
                          sig_err=rel(sigm, sig64), mu_err=abs(mum - mu64)))

    best = steps[-1]
    t_plain = float(np.mean([s["s64"]["solve_seconds"] for s in steps]))
    t_mix = float(np.mean([s["t_mixed"] for s in steps]))
    speedup = t_plain / t_mix

    # roofline of the dominant kernel (the FFT pass) during the timed mixed solve
    prof = best["pm"]
    cls = "fft_pass_fp32" if prof["fft_pass_fp32"]["launches"] else "fft_pass"
    pc = prof[cls]
    avg_ms = pc["ms"] / max(pc["launches"], 1)
    bytes_per = pc["bytes"] / max(pc["launches"], 1)
    achieved = bytes_per / (avg_ms * 1e-3) / 1e9
    p64c = best["p64"]["fft_pass"]
    ach64 = (p64c["bytes"] / max(p64c["ms"], 1e-12)) / 1e6

    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(st, int(best["s64"]["iterations"]), args.cpu_threads)
            log(f"cpu baseline: {cpu['sample']}")
        except MemoryError:
            cpu = None

    out = {
        "metric": METRIC,
        "value": round(speedup, 3),
        "unit": "x (speedup of time-to-converge)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * (t_plain + t_mix), 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 transforms / f64 accumulate, vertex, Dyson, mu (mixed) vs f64 (plain)",
        "data": "synthetic (seeded lattice)",
        "config": {"workload": f"FLEX {st['Norb']}-orbital {st['Nx']}x{st['Ny']}x{st['Nz']} k-mesh x "
                               f"{st['Nw']} Matsubara, T={st['temp']}, U={st['U']}, J={st['J']}, "
                               f"rfill={st['fill']}, tol {args.eps:g}, linear mixing 0.5, "
                               f"{'tail-corrected chi0 (FP64 tail)' if SW_TAIL else 'sharp-cutoff chi0'}",
                   "warm_start": best["winfo"]["kind"], "parallelism": "single GPU"},
        "t_plain_fp64_s": round(t_plain, 3),
        "t_mixed_warm_s": round(t_mix, 3),
        "t_warm_start_s": round(best["t_warm"], 3),
        "warm_start": best["winfo"],
        "iterations_plain": int(best["s64"]["iterations"]),
        "iterations_mixed": int(best["sm"]["iterations"]),
        "ms_per_iteration_plain": round(1e3 * best["s64"]["solve_seconds"] / best["s64"]["iterations"], 1),
        "ms_per_iteration_mixed": round(1e3 * best["sm"]["solve_seconds"] / best["sm"]["iterations"], 1),
        "wall_plain_s": round(best["s64"]["wall_seconds"], 3),
        "wall_mixed_s": round(best["sm"]["wall_seconds"], 3),
        "sigma_rel_err_vs_fp64": best["sig_err"],
        "mu_abs_err_vs_fp64": best["mu_err"],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": FP32_BOX_TRAFFIC_BYTES.get(args.config) if cls == "fft_pass_fp32" else None,
                     "kernel": f"box_fft_kernel<{'float' if cls == 'fft_pass_fp32' else 'double'}>, "
                               f"box_accum_kernel (the FFT passes, all instantiations) and the Sigma increment's "
                               f"contraction (k_sigma_inc_contract)", "avg_launch_us": round(1e3 * avg_ms, 2),
                     "launches": int(pc["launches"]), "algorithmic_bytes_per_launch": bytes_per,
                     "bytes_model": "per pass: N b per full grid read or written (fused operands included), "
                                    "irreducible (k,w) slices at their size; average over the launches"},
        "roofline_fp64_fft_pass_gbs": round(ach64, 1),
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
