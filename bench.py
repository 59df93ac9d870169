"""Benchmark: H.psi applications/s of the stochastic-DFT Chebyshev filter.

Workload (default): Si64, 2x2x2 diamond supercell a = 20.52 bohr, 12 Ry
(|G|^2 <= 12 bohr^-2, ncoef 3016), 64^3 grid, 256 Rademacher orbitals from
seed 1, Chebyshev degree L = 1024, kT = 0.01 Ha, analytic pseudopotentials.

A step is one full filter pass over the orbital block: L-1 = 1023 Chebyshev
steps, i.e. 256 x 1023 H.psi applications (filterChebyPolyHerm with the
orbitals resident in HBM).  The headline value is the FP32-MFMA precision
form; the plain FP64 form is timed on the same inputs, and the FP32 result's
error against the FP64 one is reported (density L1 per electron, band energy
per atom).  Per-kernel durations come from HIP events on the library's stream.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(REPO, "stochastic-density-functional-theory-piny_amd")
for p in (REPO, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6   # vector (and matrix) FP64
FP32_VALU_PEAK_TFLOPS = 157.3
FP32_VALU_UNPACKED_TFLOPS = 78.6  # one FP32 FMA per lane per clock (no v_pk_*)
FP32_MFMA_PEAK_TFLOPS = 157.3   # dense MFMA fp32
FP64_MFMA_PEAK_TFLOPS = 78.6    # dense MFMA fp64


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def plane_kernel_model(sysm, ncol, npairs, wc, wv, M):
    """Algorithmic bytes / flops of one k_plane launch over all pairs:
    W read + written once, v read once; FFT flops 5 n log2 n per line."""
    n = sysm.n[0]
    bytes_ = 2.0 * npairs * n * ncol * wc + float(n) ** 3 * wv
    fft = 5.0 * n * math.log2(n)
    flops = npairs * n * ((2 * M + 2 * n) * fft + 2.0 * 2 * n * n)
    return bytes_, flops


def survey_bytes_per_hpsi(sysm, wc, wv):
    """The per-H.psi byte model B = 2 N w_c + N w_v / 2 + 80 ncoef of a
    full-grid FFT sandwich (three transposing passes per side)."""
    N = float(sysm.ngrid)
    return 2.0 * N * wc + N * wv / 2.0 + 80.0 * sysm.ncoef


def host_info():
    """The host the CPU baseline runs on: CPUs this process may use (nproc),
    the machine's CPU count, the lscpu model name, OMP_NUM_THREADS."""
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        nproc = os.cpu_count() or 1
    model = None
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            if line.split(":")[0].strip() == "Model name":
                model = line.split(":", 1)[1].strip()
                break
    except Exception:
        pass
    if model is None:
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
    return {"nproc": nproc, "cpu_count": os.cpu_count(), "model": model,
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(sysm, chi, coeffs, scale, emean, nsteps, workers, runs=3, warmup=3):
    """Time the CPU FP64 oracle (same algorithm: pair-packed 3D FFTs, v.psi,
    G-space KB, kinetic, recurrence, accumulate) over the whole orbital
    block: `warmup` untimed H.psi (plans, caches), then `runs` runs of
    `nsteps` Chebyshev steps from chi; returns the median rate and the total
    timed seconds."""
    from oracle import sdft_oracle as O
    O.set_fft_workers(workers)
    for _ in range(warmup):
        O.hpsi_raw(chi, sysm)
    rates, total = [], 0.0
    for _ in range(runs):
        c = chi.copy()
        t_prev = chi.copy()
        acc = coeffs[0] * chi
        t0 = time.perf_counter()
        for k in range(1, nsteps + 1):
            f = O.hpsi_raw(c, sysm)
            O.norm_h_cheby(c, f, t_prev, k, scale, emean)
            if k > 1:
                t_prev = c
            c = f
            acc = acc + coeffs[k] * c
        dt = time.perf_counter() - t0
        rates.append(chi.shape[0] * nsteps / dt)
        total += dt
    return float(np.median(rates)), total


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="si64")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=10,
                    help="timed Chebyshev steps per CPU-baseline run (3 runs, median)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 343-fragment batch measurement")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the Si512 / Si1728 block measurements")
    ap.add_argument("--workloads-only", action="store_true",
                    help="only the Si512 / Si1728 block measurements (JSON line of its own)")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the 343-fragment batch measurement (JSON line of its own)")
    ap.add_argument("--batch-config", default=None,
                    help="fixed KB split,overlap,plane_tile[,fused_expand] for the batch (e.g. "
                         "2,0,0,1) instead of the autotuner")
    ap.add_argument("--config", default=None,
                    help="plane_tile,kb_wg_target,fused_expand for the single-system filter "
                         "instead of the tuner's pick")
    ap.add_argument("--pmc-run", action="store_true",
                    help="short run for hardware counters: an 8-term FP32 and FP64 filter of "
                         "the workload, then a 256 MiB plain copy (counter calibration)")
    args = ap.parse_args()
    if args.pmc_run:
        pmc_run(args.workload)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the filter path runs on a single MI355X "
              "(orbital sharding across GPUs is a design note, not a launched mode); "
              "use --gpus 1", file=sys.stderr)
        sys.exit(2)

    from sdft_amd import hip, stodft, systems

    if args.batch_only:
        print(json.dumps(batch_bench(args, 0.2)), flush=True)
        return
    if args.workloads_only:
        print(json.dumps({k: workload_bench(k) for k in ("si512", "si1728")}), flush=True)
        return
    wl = systems.WORKLOADS[args.workload]
    t0 = time.perf_counter()
    sysm = systems.build_system(wl.name, wl.a0, wl.rep, wl.n)
    chi = systems.noise_orbitals(sysm, wl.nstate, wl.seed)
    log("setup: %s ncoef=%d nproj=%d grid=%d^3 (%.1fs)" % (wl.name, sysm.ncoef, sysm.nproj,
                                                          wl.n, time.perf_counter() - t0))
    ctx = hip.Context(sysm)
    # spectral range by power iteration on the GPU (genEnergyMax/Min)
    emax = stodft.gen_energy_max(ctx, niter=500)
    emin = stodft.gen_energy_min(ctx, emax, niter=3000)
    lo, hi = stodft.spectral_bounds(sysm)
    emin, emax = max(emin, lo), min(emax, hi)
    L = wl.degree
    beta = 1.0 / wl.kt
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    ne = sysm.nelec
    mom = ctx.cheby_moments(chi, L, scale, emean, hip.PREC_FP64)
    mu, ne_mu = ctx.chem_pot(L, emin, emax, beta, mom, wl.nstate, ne, 0.0, 0.5)
    coeffs = ctx.cheby_coeffs(L, emin, emax, mu, beta, 2)
    log("spectral range [%.4f, %.4f] mu=%.6f N(mu)=%.8f" % (emin, emax, mu, ne_mu))

    nst = wl.nstate
    hpsi_per_step = nst * (L - 1)
    results = {}
    tuned = {}
    for prec, tag in ((hip.PREC_FP32, "fp32"), (hip.PREC_FP64, "fp64")):
        # launch configuration for this precision (plane tile, KB workgroups)
        if args.config:
            cfg = tuple(int(x) for x in args.config.split(","))
            ctx.set_config(*cfg)
            cand = np.zeros(0)
        else:
            cfg, cand = ctx.tune(nst, prec, steps=4)
        tuned[tag] = {"plane_tile": cfg[0], "kb_wg_target": cfg[1], "fused_expand": cfg[2],
                      "candidates_ms_per_step": np.round(cand, 4).tolist()}
        log("%s tune: plane_tile=%d kb_wg_target=%d fused_expand=%d candidates %s" % (
            tag, cfg[0], cfg[1], cfg[2], np.round(cand, 4).tolist()))
        blk = hip.Block(ctx, nst, prec)
        blk.set(chi)
        for _ in range(args.warmup):
            blk.filter(coeffs, scale, emean)
        t1 = time.perf_counter()
        for _ in range(args.steps):
            blk.filter(coeffs, scale, emean)
        dt = time.perf_counter() - t1
        out = blk.get(0)
        # per-kernel averages: one more filter with HIP events recorded around
        # every kernel on the stream it runs on (same two-stream schedule as
        # the timed filters; the events themselves cost a few percent, so the
        # headline rate comes from the unprofiled filters above)
        ctx.set_profiling(True)
        t2 = time.perf_counter()
        blk.filter(coeffs, scale, emean)
        dt_prof = time.perf_counter() - t2
        ctx.set_profiling(False)
        kt = ctx.kernel_times()
        blk.close()
        log("%s: profiled filter %.2f ms (unprofiled %.2f ms)" % (tag, 1e3 * dt_prof,
                                                                 1e3 * dt / args.steps))
        # the separate moments pass (calcChebyMoments: L/2 H.psi steps per
        # orbital, moments by the doubling identities), one warm call and one
        # timed call; includes the host -> device copy of the block (~12 MB)
        ctx.cheby_moments(chi, L, scale, emean, prec)
        t3 = time.perf_counter()
        mom_p = ctx.cheby_moments(chi, L, scale, emean, prec)
        dt_mom = time.perf_counter() - t3
        mom_err = float(np.max(np.abs(mom_p - mom)) / np.max(np.abs(mom)))
        results[tag] = dict(ms_per_step=1e3 * dt / args.steps,
                            hpsi_per_s=hpsi_per_step * args.steps / dt, filtered=out, kernels=kt,
                            moments=dict(ms=round(1e3 * dt_mom, 3),
                                         hpsi_per_s=round(nst * (L // 2) / dt_mom, 1),
                                         max_rel_diff_vs_fp64=mom_err))
        log("%s: %.2f ms/filter, %.4g H.psi/s; kernels %s" % (
            tag, results[tag]["ms_per_step"], results[tag]["hpsi_per_s"],
            {k: (round(v[0], 3), v[1]) for k, v in kt.items()}))

    # FP32 vs FP64 error on the filtered orbitals
    occ_over = 2.0 / nst
    rho64 = ctx.density(results["fp64"]["filtered"], occ_over)
    rho32 = ctx.density(results["fp32"]["filtered"], occ_over)
    N = sysm.ngrid
    ne64 = rho64.sum() / N
    dens_l1 = float(np.abs(rho32 - rho64).sum() / N / ne64)

    def band_energy(C, rho):
        eke, enl = ctx.energies(C, 2.0, nst)
        return eke + enl + float(rho @ sysm.vks) / N

    e64 = band_energy(results["fp64"]["filtered"], rho64)
    e32 = band_energy(results["fp32"]["filtered"], rho32)
    de_atom = abs(e32 - e64) / sysm.natm
    f64 = ctx.nl_forces(results["fp64"]["filtered"], 2.0, nst)
    f32 = ctx.nl_forces(results["fp32"]["filtered"], 2.0, nst)
    df = float(np.max(np.abs(f32 - f64)))
    log("FP32 vs FP64: density L1/e %.3e, energy/atom %.3e Ha (E64 %.6f, Ne %.6f)" % (
        dens_l1, de_atom, e64, ne64))

    # v_KS from the density (Hartree + external + LDA on the device)
    ks = ks_bench(ctx, sysm, rho64, not args.no_cpu_baseline)

    # roofline of the dominant kernel (FP32 form)
    kt32 = results["fp32"]["kernels"]
    dom = max(kt32.items(), key=lambda kv: kv[1][0])
    dom_name, (dom_ms, dom_launch) = dom[0], dom[1]
    avg_s = dom_ms / dom_launch / 1e3
    ncol = ctx_ncol(sysm)
    M = 2 * int(np.abs(sysm.k).max()) + 1
    bytes_, flops = plane_kernel_model(sysm, ncol, nst // 2, 8, 4, M)
    roofline = None
    # the plane kernel's instantiation: tile variant, pruned passes when 4K < n
    pruned = "true" if 4 * int(np.abs(sysm.k).max()) < wl.n else "false"
    if dom_name == "plane_bc_vpsi":
        gbs = bytes_ / avg_s / 1e9
        tfl = flops / avg_s / 1e12
        kname = plane_kernel_name("float", wl.n, tuned["fp32"]["plane_tile"], pruned)
        roofline = {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(gbs / HBM_PEAK_GBS, 4),
                    "traffic": measured_traffic(kname, wl),
                    "kernel": kname + " (b/c transforms + v.psi)",
                    "avg_launch_ms": round(avg_s * 1e3, 4),
                    "algorithmic_bytes": bytes_, "algorithmic_flops": flops,
                    "achieved_tflops": round(tfl, 2),
                    "valu_peak_tflops": FP32_VALU_PEAK_TFLOPS,
                    "valu_frac": round(tfl / FP32_VALU_PEAK_TFLOPS, 4),
                    # the build compiles without packed FP32, whose rate per
                    # lane is that of FP64: the reachable VALU peak is half
                    "valu_peak_unpacked_tflops": FP32_VALU_UNPACKED_TFLOPS,
                    "valu_frac_unpacked": round(tfl / FP32_VALU_UNPACKED_TFLOPS, 4)}
    else:
        roofline = {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": None, "traffic": None, "kernel": dom_name,
                    "avg_launch_ms": round(avg_s * 1e3, 4)}

    # KB products on the MFMA units: algorithmic flops per launch (projection
    # Z = A X and force F = Y At, each 2 x nproj x 2 ncoef x orbitals) over
    # the event-timed launch, which runs overlapped with the FFT kernels on
    # its own stream (so these are lower bounds of the standalone rate)
    kb_flops = 2.0 * sysm.nproj * 2 * sysm.ncoef * nst
    kb_mfma = {"flops_per_launch": kb_flops,
               "note": "event time of each launch while overlapped with the FFT kernels "
                       "on the main stream; kb_project includes its split-K reduction"}
    for tag, peak in (("fp32", FP32_MFMA_PEAK_TFLOPS), ("fp64", FP64_MFMA_PEAK_TFLOPS)):
        kt = results[tag]["kernels"]
        for name in ("kb_project", "kb_force"):
            if name in kt and kt[name][1] > 0:
                ms = kt[name][0] / kt[name][1]
                tf = kb_flops / (ms * 1e-3) / 1e12
                kb_mfma["%s_%s" % (tag, name)] = {"ms": round(ms, 4), "tflops": round(tf, 2),
                                                  "peak_tflops": peak,
                                                  "frac": round(tf / peak, 4)}

    cpu = None
    if not args.no_cpu_baseline:
        host = host_info()
        workers = host["nproc"]
        rate, dt = cpu_baseline(sysm, chi, coeffs, scale, emean, args.cpu_steps, workers)
        cpu = {"value": round(rate, 2), "unit": "H.psi/s", "cores": workers, "kind": "port",
               "host": host,
               "sample": "median of 3 runs of %d Chebyshev steps x %d orbitals of the same %s "
                         "workload after 3 untimed H.psi, FP64 numpy/scipy.fft oracle with "
                         "%d FFT threads (every CPU this process may run on) (%.1f s timed)"
                         % (args.cpu_steps, nst, wl.name, workers, dt)}
        log("cpu baseline: %.1f H.psi/s on %d threads (%s)" % (rate, workers, host["model"]))

    batch = None if args.no_batch else batch_bench(args, mu)
    workloads = None
    if not args.no_workloads:
        workloads = {k: workload_bench(k) for k in ("si512", "si1728")}

    r32, r64 = results["fp32"], results["fp64"]
    survey_frac32 = survey_bytes_per_hpsi(sysm, 8, 4) * r32["hpsi_per_s"] / (HBM_PEAK_GBS * 1e9)
    line = {
        "metric": "Hψ applications/s (orbital×Cheb step), FP32-MFMA vs FP64 at ≤1e-4 Ha/atom",
        "value": round(r32["hpsi_per_s"], 1),
        "unit": "Hψ/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r32["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 (FFT, v.psi, KB MFMA) + fp64 (recurrence, accumulators, reductions)",
        "data": "synthetic (seeded Rademacher noise, analytic pseudopotentials)",
        "config": {"workload": wl.name, "grid": wl.n, "ncoef": sysm.ncoef, "orbitals": nst,
                   "degree": L, "kT": wl.kt, "natm": sysm.natm, "nproj": sysm.nproj,
                   "hpsi_per_step": hpsi_per_step},
        "fp64": {"value": round(r64["hpsi_per_s"], 1), "ms_per_step": round(r64["ms_per_step"], 3)},
        "fp32_over_fp64": round(r32["hpsi_per_s"] / r64["hpsi_per_s"], 3),
        "fp32_error": {"density_l1_per_electron": dens_l1, "energy_per_atom_ha": de_atom,
                       "nl_force_ha_per_bohr": df, "nl_force_max_ha_per_bohr": float(np.abs(f64).max())},
        "survey_model_hbm_frac": round(survey_frac32, 4),
        "tuned": tuned,
        "kernels_fp32_ms": {k: round(v[0] / v[1], 4) for k, v in r32["kernels"].items()},
        "kernels_fp64_ms": {k: round(v[0] / v[1], 4) for k, v in r64["kernels"].items()},
        "roofline": roofline,
        "kb_mfma": kb_mfma,
        "moments_pass": {"hpsi_per_pass": nst * (L // 2),
                         "fp32": r32["moments"], "fp64": r64["moments"],
                         "note": "calcChebyMoments over the block (L/2 H.psi steps per orbital), "
                                 "host arrays in and out"},
        "cpu_baseline": cpu,
        "fragment_batch": batch,
        "ks_potential": ks,
        "workloads": workloads,
    }
    print(json.dumps(line), flush=True)


def workload_bench(key, nsteps=24, block=128):
    """One 128-orbital block of the Si512 / Si1728 workload (systems.
    workload_system: real-space KB projectors in the library's per-atom
    form) through `nsteps` Chebyshev steps of its degree-L expansion: the
    per-step cost does not depend on L, so the rate of a short pass is the
    rate of the full one.  FP32 and FP64 forms; per-kernel event times of one
    profiled pass; the NL kernels (real-space product and v.psi + scatter)
    against the FFT kernels (expand, the two plane stages, contract); and the
    HBM rate of each kernel from its algorithmic bytes:
      plane_bc_to_r: W read + the pair grids written (n ncol + n^3) w_c per pair
      rs_project:    the grids read at the projectors' points (each point once
                     per atom that covers it) + the beta values and indices
      rs_apply_vpsi: the grids read and written + v + the inverse lists
      plane_r_to_bc: the grids read + W written."""
    from sdft_amd import hip, stodft, systems
    wl = systems.WORKLOADS[key]
    t0 = time.perf_counter()
    sysm, rs = systems.workload_system(key)
    ctx = hip.Context(sysm, with_nl=False)
    ctx.set_projectors_real(rs)
    emin, emax = stodft.spectral_bounds(sysm, rs)
    log("%s setup: ncoef %d, P %d, %d KB grid entries (%.1fs)" % (
        key, sysm.ncoef, len(rs.energy), rs.offsets[-1], time.perf_counter() - t0))
    L = wl.degree
    beta = 1.0 / wl.kt
    coeffs = ctx.cheby_coeffs(L, emin, emax, 0.3, beta, 2)[:nsteps + 1]
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    N = sysm.ngrid
    npairs = block // 2
    ncol = ctx_ncol(sysm)
    nent = int(rs.offsets[-1])
    # distinct (atom point list, channel group) entries: 4 channels share a list
    nent_g = nent // 4
    out = {"workload": wl.name, "grid": wl.n, "ncoef": sysm.ncoef, "natm": sysm.natm,
           "nproj": int(len(rs.energy)), "kb_grid_entries": nent, "block": block,
           "degree": L, "timed_steps": nsteps, "nl_form": "per-atom real space"}
    for prec, tag in ((hip.PREC_FP32, "fp32"), (hip.PREC_FP64, "fp64")):
        wc, wv = (8, 4) if prec == hip.PREC_FP32 else (16, 8)
        blk = hip.Block(ctx, block, prec)
        blk.noise(wl.seed, 0)
        blk.filter(coeffs, scale, emean)  # warm
        t1 = time.perf_counter()
        blk.filter(coeffs, scale, emean)
        dt = time.perf_counter() - t1
        ctx.set_profiling(True)
        blk.filter(coeffs, scale, emean)
        ctx.set_profiling(False)
        kt = ctx.kernel_times()
        blk.close()
        per = {k: v[0] / v[1] for k, v in kt.items() if v[1] > 0}
        steps = nsteps  # Chebyshev steps applied (L' - 1 with L' = nsteps + 1)
        nl_ms = sum(kt[k][0] for k in ("rs_project", "rs_apply_vpsi") if k in kt) / steps
        fft_ms = sum(kt[k][0] for k in ("expand_a", "plane_bc_to_r", "plane_r_to_bc",
                                        "contract_a") if k in kt) / steps
        grid_b = float(N) * wc * npairs
        wbytes = float(wl.n) * ncol * wc * npairs
        qc = 4  # pairs per product workgroup
        nch = -(-npairs // qc)
        wr = wc // 2  # one real value
        # product: the grids at every (atom, point) entry, beta + index per
        # pair chunk; apply: the grids read and written, v, the point offsets
        # and the (slot, weight) lists once per point
        model = {"plane_bc_to_r": wbytes + grid_b,
                 "rs_project": float(nent_g) * wc * npairs + float(nent_g) * (4 + 4 * wr) * nch,
                 "rs_apply_vpsi": 2.0 * grid_b + float(N) * (wv + 8) + float(nent) * (4 + wr),
                 "plane_r_to_bc": grid_b + wbytes}
        kern = {}
        for k, ms in per.items():
            d = {"ms": round(ms, 4)}
            if k in model:
                gbs = model[k] / (ms * 1e-3) / 1e9
                d.update({"algorithmic_bytes": model[k], "gbs": round(gbs, 1),
                          "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)})
            kern[k] = d
        out[tag] = {"hpsi_per_s": round(block * nsteps / dt, 1),
                    "ms_per_step": round(1e3 * dt / nsteps, 4),
                    "nl_ms_per_step": round(nl_ms, 4), "fft_ms_per_step": round(fft_ms, 4),
                    "kernels": kern}
        log("%s %s: %.4g H.psi/s, %.3f ms/step (NL %.3f, FFT %.3f ms/step)" % (
            key, tag, block * nsteps / dt, 1e3 * dt / nsteps, nl_ms, fft_ms))
    out["fp32_over_fp64"] = round(out["fp32"]["hpsi_per_s"] / out["fp64"]["hpsi_per_s"], 3)
    out["end_to_end"] = workload_errors(ctx, sysm, rs, wl, emin, emax, block)
    ctx.close()
    return out


def workload_errors(ctx, sysm, rs, wl, emin, emax, block, degree=256):
    """The workload's whole path on one 128-orbital block at a reduced degree
    (the precision comparison, not the expansion's accuracy, is the point):
    FP64 moments -> mu for N_e = 4 natm (tolerance 1e-6 electrons) -> sqrt(f)
    coefficients -> the filter in both forms -> per form the density, the
    total energy as calcTotEnergy assembles it without the ion-ion term
    (E_kin + E_NL + E_H + E_ext + E_xc, stodft/calc-energy.c:329-333) and the
    Hellmann-Feynman forces on every atom (NL in the per-atom real-space form
    + local); FP32 against FP64."""
    from sdft_amd import hip, systems
    t0 = time.perf_counter()
    beta = 1.0 / wl.kt
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    chi = ctx.noise(block, wl.seed)
    mom = ctx.cheby_moments(chi, degree, scale, emean)
    mu, nmu = ctx.chem_pot(degree, emin, emax, beta, mom, block, sysm.nelec, 0.3, 0.4, tol=1e-6)
    coeffs = ctx.cheby_coeffs(degree, emin, emax, mu, beta, 2)
    dens = systems.density_sphere(sysm)
    vext = systems.vloc_density(sysm, dens, separable=True)
    vloc = systems.vloc_form_factor(sysm, dens)
    types = np.zeros(sysm.natm, dtype=np.int32)
    ctx.set_projector_gradients(systems.real_space_projector_gradients(sysm, rs))
    ctx.ks_init(dens)
    res = {}
    for prec, tag in ((hip.PREC_FP64, "fp64"), (hip.PREC_FP32, "fp32")):
        ctx.set_vks(sysm.vks)
        C, _ = ctx.cheby_filter(chi, coeffs, scale, emean, prec=prec)
        eke, enl = ctx.energies(C, 2.0, block)
        rho = ctx.density(C, 2.0 / block)
        _, en = ctx.ks_potential(rho, vext, sysm.volume, xc=0)
        e = eke + enl + en["eh"] + en["eext"] + en["exc"]
        f = ctx.nl_forces(C, 2.0, block) + ctx.local_forces(sysm.pos, types, vloc, dens.gvec)
        res[tag] = (e, f, rho)
    (e64, f64, r64), (e32, f32, r32) = res["fp64"], res["fp32"]
    N = sysm.ngrid
    ne64 = float(r64.sum()) / N
    out = {"degree": degree, "mu": mu, "n_mu": nmu, "n_e": sysm.nelec,
           "energy_fp64_ha": e64, "energy_per_atom_err_ha": abs(e32 - e64) / sysm.natm,
           "force_max_ha_per_bohr": float(np.abs(f64).max()),
           "force_err_max_ha_per_bohr": float(np.abs(f32 - f64).max()),
           "density_l1_per_electron": float(np.abs(r32 - r64).sum() / N / ne64),
           "seconds": round(time.perf_counter() - t0, 1)}
    log("%s end to end (L=%d): mu %.6f N %.7f, E %.6f Ha, FP32 err %.3g Ha/atom, forces %.3g "
        "Ha/bohr (|F| max %.3g), density L1/e %.3g" % (
            wl.name, degree, mu, nmu, e64, out["energy_per_atom_err_ha"],
            out["force_err_max_ha_per_bohr"], out["force_max_ha_per_bohr"],
            out["density_l1_per_electron"]))
    return out


def ks_bench(ctx, sysm, rho, with_cpu, reps=20):
    """sdft_hip_ks_potential on the workload's own FP64 density: wall time
    per call (host density in, v_KS set on the device and copied out, the
    four energies), and the FP64 numpy oracle's time for one call."""
    from sdft_amd import systems
    dens = systems.density_sphere(sysm)
    vext = systems.vloc_density(sysm, dens)
    ctx.ks_init(dens)
    ctx.ks_potential(rho, vext, sysm.volume, xc=0)
    t0 = time.perf_counter()
    for _ in range(reps):
        vks, en = ctx.ks_potential(rho, vext, sysm.volume, xc=0)
    dt = (time.perf_counter() - t0) / reps
    out = {"ms_per_call": round(1e3 * dt, 3), "density_sphere_ncoef": dens.ncoef, "xc": "pz_lda",
           "energies": {k: float(v) for k, v in en.items()}}
    if with_cpu:
        from oracle import sdft_oracle as O
        t1 = time.perf_counter()
        ref, _ = O.ks_potential(rho, vext, dens, sysm.volume, 0)
        out["cpu_oracle_ms"] = round(1e3 * (time.perf_counter() - t1), 1)
        out["max_rel_diff_vs_oracle"] = float(np.max(np.abs(vks - ref)) / np.max(np.abs(ref)))
    ctx.set_vks(sysm.vks)
    log("ks_potential: %.3f ms/call (density sphere %d coefficients)" % (1e3 * dt, dens.ncoef))
    return out


def _build_fragment(spec):
    from threadpoolctl import threadpool_limits

    from sdft_amd import stodft, systems
    with threadpool_limits(1):  # one BLAS thread per pool worker
        s = systems.build_fragment(spec)
        lo, hi = stodft.spectral_bounds(s)
    return s, lo, hi


def batch_bench(args, mu, degree=512, kt=0.01, seed=3):
    """The fragment batch: 343 fragments of the Si1728 cell (8-64 atoms, seed
    3), grids 24^3-48^3, 32-128 orbitals each, Chebyshev degree 512, all
    filtered together (sdft_hip_batch_filter: graph-replayed steps, launch
    configuration from the autotuner).  FP32 form timed and checked against
    the FP64 form on a sample of fragments; mu is shared by all fragments."""
    from concurrent.futures import ProcessPoolExecutor

    from sdft_amd import hip, systems
    t0 = time.perf_counter()
    specs = systems.fragment_specs(343, seed)
    workers = min(16, os.cpu_count() or 1)
    with ProcessPoolExecutor(workers) as ex:
        built = list(ex.map(_build_fragment, specs, chunksize=4))
    log("batch setup: %d fragments built (%.1fs)" % (len(built), time.perf_counter() - t0))
    ctxs = [hip.Context(s) for s, _, _ in built]
    L = degree
    beta = 1.0 / kt
    coeffs = np.stack([c.cheby_coeffs(L, lo, hi, mu, beta, 2)
                       for c, (_, lo, hi) in zip(ctxs, built)])
    scale = np.array([2.0 / (hi - lo) for _, lo, hi in built])
    emean = np.array([0.5 * (hi + lo) for _, lo, hi in built])
    nst = [sp.nstate for sp in specs]
    first = np.concatenate([[0], np.cumsum(nst)[:-1]]).astype(int)
    hpsi = float(sum(nst)) * (L - 1)
    grids = {}
    for sp in specs:
        grids[sp.n] = grids.get(sp.n, 0) + 1
    log("batch: %d orbitals, grids %s (%.1fs)" % (sum(nst), grids, time.perf_counter() - t0))
    sample = list(range(0, len(specs), 43))
    out, res = {}, {}
    for prec, tag in ((hip.PREC_FP32, "fp32"), (hip.PREC_FP64, "fp64")):
        blocks = [hip.Block(c, n, prec) for c, n in zip(ctxs, nst)]
        for b, f0 in zip(blocks, first):
            b.noise(seed, int(f0))
        bat = hip.Batch(blocks)
        if args.batch_config:
            cfg = [int(x) for x in args.batch_config.split(",")]
            split, ov, pv = cfg[:3]
            fu = cfg[3] if len(cfg) > 3 else 1
            bat.config(split, ov, pv, fu)
            cand = np.zeros(0)
        else:
            (split, ov, pv, fu), cand = bat.autotune(steps=6)
        log("%s autotune: split=%d overlap=%d plane_tile=%d fused_expand=%d "
            "candidates(ms/step)=%s" % (tag, split, ov, pv, fu, np.round(cand, 3).tolist()))
        steps = args.steps if prec == hip.PREC_FP32 else 1
        for _ in range(args.warmup):
            bat.filter(coeffs, scale, emean)
        t1 = time.perf_counter()
        for _ in range(steps):
            bat.filter(coeffs, scale, emean)
        dt = (time.perf_counter() - t1) / steps
        t2 = time.perf_counter()
        bat.filter(coeffs, scale, emean, use_graph=False)
        dt_direct = time.perf_counter() - t2
        # the whole filter as one graph launch (the first call captures it)
        bat.filter(coeffs, scale, emean, use_graph=2)
        t4 = time.perf_counter()
        bat.filter(coeffs, scale, emean, use_graph=2)
        dt_whole = time.perf_counter() - t4
        out[tag] = [blocks[i].get(0) for i in sample]
        bat.close()
        # the same fragments one after another (the reference's serial
        # fragment loop, frag-scf.c:211-248): each block filtered alone
        # (sdft_hip_block_filter, the contexts' default launch configuration);
        # one untimed short pass over all first
        for f, b in enumerate(blocks):
            b.filter(coeffs[f][:3], scale[f], emean[f])
        t3 = time.perf_counter()
        for f, b in enumerate(blocks):
            b.filter(coeffs[f], scale[f], emean[f])
        dt_serial = time.perf_counter() - t3
        res[tag] = dict(s_per_filter=dt, hpsi_per_s=hpsi / dt, direct_s=dt_direct,
                        whole_graph_s=dt_whole, serial_s=dt_serial, split=split, overlap=ov,
                        plane_tile=pv, fused_expand=fu, cand=cand)
        log("%s batch: %.3f s/filter (graph), %.3f s (direct), %.3f s (whole filter as one "
            "graph), %.4g H.psi/s; serial block filters %.3f s (batch/serial speed-up %.2f)" % (
                tag, dt, dt_direct, dt_whole, hpsi / dt, dt_serial, dt_serial / dt))
        for b in blocks:
            b.close()
    # FP32 vs FP64 on the sampled fragments
    de, dl1 = [], []
    for j, i in enumerate(sample):
        sysm, c = built[i][0], ctxs[i]
        N = sysm.ngrid
        occ_over = 2.0 / nst[i]
        e = []
        rhos = []
        for tag in ("fp64", "fp32"):
            rho = c.density(out[tag][j], occ_over)
            eke, enl = c.energies(out[tag][j], 2.0, nst[i])
            e.append(eke + enl + float(rho @ sysm.vks) / N)
            rhos.append(rho)
        ne = rhos[0].sum() / N
        dl1.append(float(np.abs(rhos[1] - rhos[0]).sum() / N / ne))
        de.append(abs(e[1] - e[0]) / sysm.natm)
    r32, r64 = res["fp32"], res["fp64"]
    return {
        "workload": "343 fragments of Si1728 (8-64 atoms, seed 3), grids 24-48^3, "
                    "32-128 orbitals, L=%d" % L,
        "grids": {str(k): v for k, v in sorted(grids.items())},
        "orbitals": int(sum(nst)),
        "hpsi_per_filter": hpsi,
        "value_fp32": round(r32["hpsi_per_s"], 1),
        "s_per_filter_fp32": round(r32["s_per_filter"], 4),
        "s_per_filter_fp32_direct_launches": round(r32["direct_s"], 4),
        "s_per_filter_fp32_whole_filter_graph": round(r32["whole_graph_s"], 4),
        "s_per_filter_fp64_whole_filter_graph": round(r64["whole_graph_s"], 4),
        "serial_block_filters_s_fp32": round(r32["serial_s"], 4),
        "serial_block_filters_s_fp64": round(r64["serial_s"], 4),
        "batch_over_serial_fp32": round(r32["serial_s"] / r32["s_per_filter"], 3),
        "batch_over_serial_fp64": round(r64["serial_s"] / r64["s_per_filter"], 3),
        "value_fp64": round(r64["hpsi_per_s"], 1),
        "s_per_filter_fp64": round(r64["s_per_filter"], 4),
        "fp32_over_fp64": round(r32["hpsi_per_s"] / r64["hpsi_per_s"], 3),
        "autotune_fp32": {"kb_split": r32["split"], "kb_overlap": r32["overlap"],
                          "plane_tile": r32["plane_tile"], "fused_expand": r32["fused_expand"],
                          "candidates_ms_per_step": np.round(r32["cand"], 4).tolist()},
        "autotune_fp64": {"kb_split": r64["split"], "kb_overlap": r64["overlap"],
                          "plane_tile": r64["plane_tile"], "fused_expand": r64["fused_expand"]},
        "fp32_error_sample": {"fragments": sample,
                              "energy_per_atom_ha_max": float(max(de)),
                              "density_l1_per_electron_max": float(max(dl1))},
        "mu": mu,
    }


# plane-kernel variants (b-line tile V, threads per workgroup, radix
# schedules of the c- and b-lines), as the library's tuner numbers them; the
# variants above 1 exist at n = 64 only
PLANE_VARIANTS = {0: (0, 256, 0, 0), 1: (1, 256, 0, 0), 2: (0, 256, 0, 1), 3: (1, 512, 1, 0),
                  4: (1, 512, 1, 1), 5: (0, 512, 1, 1)}


def plane_kernel_name(t, n, variant, pruned):
    """The plane kernel's instantiation as rocprofv3 names it."""
    v, nt, sc, sb = PLANE_VARIANTS[variant if n == 64 else variant & 1]
    return "k_plane<%s, %d, %d, %s, %d, %d, %d>" % (t, n, v, pruned, nt, sc, sb)


def measured_traffic(kernel, wl):
    """HBM bytes per dispatch of `kernel` from the committed counter summary
    (rocprofv3 FETCH_SIZE / WRITE_SIZE scaled by the copy-kernel factors)."""
    path = os.path.join(REPO, "profiles", "hbm_traffic_%s.json" % wl.name.split("-")[0])
    try:
        import json as _j
        d = _j.load(open(path))
        k = d["kernels"][kernel]
        return round(k["read_bytes"] + k["write_bytes"])
    except Exception:
        return None


def pmc_run(workload):
    """Few dispatches of each hot kernel for `rocprofv3 --pmc` passes."""
    from sdft_amd import hip, stodft, systems
    wl = systems.WORKLOADS[workload]
    sysm = systems.build_system(wl.name, wl.a0, wl.rep, wl.n)
    chi = systems.noise_orbitals(sysm, wl.nstate, wl.seed)
    ctx = hip.Context(sysm)
    emin, emax = stodft.spectral_bounds(sysm)
    scale, emean = 2.0 / (emax - emin), 0.5 * (emin + emax)
    coeffs = ctx.cheby_coeffs(8, emin, emax, 0.3, 1.0 / wl.kt, 2)
    for prec in (hip.PREC_FP32, hip.PREC_FP64):
        blk = hip.Block(ctx, wl.nstate, prec)
        blk.set(chi)
        blk.filter(coeffs, scale, emean)
        blk.close()
        log("pmc-run: filter prec=%d done" % prec)
    hip.selftest_copy(256 << 20, 3)
    log("pmc-run: copy 256 MiB x3 done")


def ctx_ncol(sysm):
    """Number of (kb, kc) columns of the full sphere."""
    k = sysm.k[:-1]
    cols = set(map(tuple, k[:, 1:].tolist())) | set(map(tuple, (-k[:, 1:]).tolist()))
    cols.add((0, 0))
    return len(cols)


if __name__ == "__main__":
    main()
