#!/usr/bin/env python3
"""Bead-updates/s of the bisection-move hot path on one MI355X.

A step is one sweep of every chain of the batch: one fused sweep_kernel
launch (choose section, node caches, every Sample's levels L..0, write back).
Default workload: 1024 independent 2D quantum-dot chains, N ~ U{2..20} split
between two spin species, M in {128,256,512,1024}, omega ~ U[0.5,2], SHO nodes
on even chains and free-particle nodes on odd chains, 4-level bisection, run
in mixed precision.  Beside it, on the same uniforms: the FP64 run of the same
batch (energy error and FP64 batched rate) and the FP64 single-chain rate of
the 3+3-electron SHONodes dot (M=400).  The CPU baseline is the oracle, the
FP64 C restatement, on a bounded sample of the same chains on all host cores.
Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pi-qmc_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import piqmc_capi as capi  # noqa: E402

METRIC = "bead-updates/sec on 1 MI355X, mixed-batched vs FP64 at fixed energy error"
HBM_PEAK_GBS = 8000.0
FP64_VECTOR_TFLOPS = 78.6


def measured_traffic(workload_name, precision):
    """HBM bytes per step (every sweep_kernel launch of one sweep) from the
    committed rocprofv3 FETCH_SIZE and WRITE_SIZE summary of this workload
    (copy-calibrated), or None."""
    path = os.path.join(ROOT, "profiles", f"sweep_kernel_{workload_name}_{precision}.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_step:"):
                    return float(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def workload(name, nchain):
    if name == "cfg3":
        return capi.cfg3_descs(nchain), "1024 uneven 2D dots (N 2-20, M 128-1024, SHO/free nodes, L=4)"
    if name == "cfg5":
        return capi.cfg5_descs(nchain), "4096 uneven 3D dots (N 2-64, M 256-2048, SHO nodes, L=4)"
    if name == "cfg2":
        return [capi.cfg2_desc()] * nchain, "3+3 e- 2D dot, SHONodes, M=400, L=4"
    if name == "cfg1":
        return [capi.cfg1_desc()] * nchain, "2 e- 3D dot, Coulomb, M=200, L=3"
    if name == "cfg4":
        return [capi.cfg4_desc()] * nchain, "32 e + 32 h, box L=10, FreeParticleNodes 32x32, M=1024, L=5"
    raise SystemExit(f"unknown workload {name}")


def bytes_per_update(d, w=8):
    """Compulsory bytes per bead-update: section gather + moved-bead scatter,
    w*D*[N(2^L+1) + (2^L-1)] / (2^L-1) (one moving particle)."""
    L = d.nlevel
    return w * d.ndim * (d.npart * ((1 << L) + 1) + ((1 << L) - 1)) / ((1 << L) - 1)


def flops_per_update(d):
    """FP64-equivalent flops per bead-update (sqrt/log/exp/div count 1):
    2 lam_L [P(8D+6K+12) + (6D+2) + S(11D+8)] + (20D+10) + node terms."""
    L, D = d.nlevel, d.ndim
    lam = ((1 << (L + 1)) - 1) / ((1 << L) - 1)
    P = d.npart - 1
    K = d.coulomb.norder if d.coulomb.enabled else 0
    S = 1 if d.nsho else 0
    f = 2 * lam * (P * (8 * D + 6 * K + 12) + (6 * D + 2) + S * (11 * D + 8)) + (20 * D + 10)
    if d.nnodes:
        ns = max(d.species[d.nodes[a].species].count for a in range(d.nnodes))
        f += ns * ns * (3 * D + 7) + 2 * ns ** 3 + ((1 << L) + 1) / ((1 << L) - 1) * 4 * ns * ns * (3 * D + 8)
    return f


def make_batch(descs, precision, seed):
    b = capi.Batch(descs, precision=precision)
    for c, d in enumerate(descs):
        b.set_paths(c, capi.initial_beads(d, 1000 + c))
    b.seed(seed, 0)
    return b


def all_stats(b):
    st = (capi.ChainStats * len(b.descs))()
    b.lib.pq_get_all_stats.restype = C.c_int
    b.lib.pq_get_all_stats.argtypes = [C.c_void_p, C.POINTER(capi.ChainStats)]
    capi.check(b.lib.pq_get_all_stats(b.h, st))
    return st


def timed_sweeps(b, steps):
    """steps launches bracketed by HIP events on the batch stream; returns the
    per-launch kernel times (ms) and the wall time of the whole timed region."""
    lib = b.lib
    lib.pq_run_sweeps_timed.restype = C.c_int
    lib.pq_run_sweeps_timed.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    ms = []
    b.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        v = C.c_float()
        capi.check(lib.pq_run_sweeps_timed(b.h, 1, C.byref(v)))
        ms.append(v.value)
    b.sync()
    return ms, time.perf_counter() - t0


TUNE_AFTER = 20   # sweeps from the replicated start before the autotuner times anything


def run_gpu(descs, precision, steps, warmup, seed, autotune=True, total_warmup=None, schedule=None):
    """warmup sweeps (the autotuner's timing sweeps count among them, and
    total_warmup, when given, pads the warmup to that many sweeps), then steps
    timed sweeps.  The autotuner runs after TUNE_AFTER sweeps, when the chains
    have left the replicated start and the node records are warm.
    schedule = (team, inverse, tables, node state[, heavy chains, heavy
    team]) fixes the schedule instead of autotuning."""
    b = make_batch(descs, precision, seed)
    tuned, done = None, 0
    if schedule is not None:
        b.set_schedule(schedule[0], schedule[1])
        b.set_staging(schedule[2], schedule[3])
        if len(schedule) > 4:
            b.set_split(schedule[4], schedule[5])
    elif autotune:
        b.run_sweeps(TUNE_AFTER)
        tuned, done = b.autotune(2)
        done += TUNE_AFTER
    extra = warmup if total_warmup is None else max(0, total_warmup - done)
    b.run_sweeps(extra)
    b.sync()
    done += extra
    capi.check(b.lib.pq_reset_stats(b.h))
    ms, wall = timed_sweeps(b, steps)
    st = all_stats(b)
    bu = np.array([s.bead_updates for s in st], dtype=np.float64)
    abu = np.array([s.accepted_bead_updates for s in st], dtype=np.float64)
    energy = b.energy()
    team, inv, lds = b.schedule()
    tab, fst = b.staging()
    heavy, heavy_team = b.split()
    sched = {"team": team, "inverse_lds": inv, "tables_lds": tab, "node_state_lds": fst,
             "lds_bytes": lds, "heavy_chains": heavy, "heavy_team": heavy_team, "autotune_ms": tuned}
    return b, dict(ms=ms, wall=wall, bu=bu, abu=abu, energy=energy, schedule=sched, warmup_done=done)


def all_paths(b):
    arr = np.empty(int(sum(d.nbeads() for d in b.descs)), dtype=np.float64)
    b.lib.pq_get_all_paths.restype = C.c_int
    b.lib.pq_get_all_paths.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int64]
    capi.check(b.lib.pq_get_all_paths(b.h, capi.dptr(arr), arr.size))
    return arr


def identical_chains(bm, br):
    """Chains whose final paths are bitwise the same in both batches."""
    pm, pr = all_paths(bm), all_paths(br)
    same, off = [], 0
    for d in bm.descs:
        n = d.nbeads()
        same.append(bool(np.array_equal(pm[off:off + n], pr[off:off + n])))
        off += n
    return np.array(same)


def paired_energy(bm, br, em, er):
    """Mixed vs FP64 on the same seeds and sweeps.  Both precisions move beads
    in FP64, so a chain whose accept/reject decisions never differed ends on
    bitwise the same path; its energy difference is the estimator's own.  The
    paired difference over all chains (diverged chains included) gives the
    energy error with its standard error."""
    same = identical_chains(bm, br)
    de = np.asarray(em) - np.asarray(er)
    n = len(de)
    return {"chains": n, "identical_decision_chains": int(same.sum()),
            "max_abs_dE_identical": float(np.max(np.abs(de[same]))) if same.any() else None,
            "mean_dE": float(de.mean()), "stderr_dE": float(de.std(ddof=1) / np.sqrt(n)) if n > 1 else None,
            "mean_E_fp64": float(np.mean(er)),
            "note": "per-chain thermodynamic energy of the final paths, same Philox streams and sweeps"}


def paired_observables(bm, br):
    """The other estimators of the mixed run against the FP64 run on the same
    streams: virial (nwindow 8) and Coulomb energies per chain (largest
    difference on the chains with identical decisions, mean difference with
    its standard error over all chains) and the integer g(r) counts of the
    two species (bit-identical on identical chains; total counts otherwise)."""
    same = identical_chains(bm, br)
    out = {}
    for name, fm, fr in (("virial_energy", bm.virial_energy(8), br.virial_energy(8)),
                         ("coulomb_energy", bm.coulomb_energy(1.0), br.coulomb_energy(1.0))):
        de = fm - fr
        out[name] = {"max_abs_d_identical": float(np.max(np.abs(de[same]))) if same.any() else None,
                     "mean_d": float(de.mean()), "stderr_d": float(de.std(ddof=1) / np.sqrt(len(de))),
                     "mean_fp64": float(fr.mean())}
    hm, hr = bm.pair_histogram(0, 1, 0.0, 8.0, 256), br.pair_histogram(0, 1, 0.0, 8.0, 256)
    out["pair_histogram_0_1"] = {"bins": 256, "rmax": 8.0,
                                 "identical_chains_bit_exact": bool(np.array_equal(hm[same], hr[same])),
                                 "total_counts_mixed": int(hm.sum()), "total_counts_fp64": int(hr.sum())}
    return out


def cpu_baseline(descs, budget_s=12.0):
    """The oracle (FP64 C restatement) on host cores: a bounded sample of the
    workload's chains, one chain per thread (OpenMP)."""
    import oracle_lib as ol
    lib = ol.load()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    nch = min(len(descs), max(cores, 2 * cores))
    sample = descs[:nch]
    arr = capi.descs_array(sample)
    beads = np.concatenate([capi.initial_beads(d, 1000 + c).ravel() for c, d in enumerate(sample)])
    # calibrate one sweep, then scale the sweep count to the time budget
    t0 = time.perf_counter()
    bu = lib.oq_run_batch(arr, nch, capi.dptr(beads), 1, 7, 0, cores, None)
    t1 = time.perf_counter() - t0
    nsweep = max(1, min(1000, int(budget_s / max(t1, 1e-3))))
    t0 = time.perf_counter()
    bu = lib.oq_run_batch(arr, nch, capi.dptr(beads), nsweep, 7, 0, cores, None)
    dt = time.perf_counter() - t0
    return dict(value=bu / dt, unit="bead-updates/s", cores=cores, kind="port",
                sample=f"{nch} chains of the workload x {nsweep} sweeps, oracle (FP64 C, "
                       f"-O3 -march=x86-64-v3, OpenMP {cores} threads), {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cfg3")
    ap.add_argument("--nchain", type=int, default=None)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-pair", action="store_true")
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--schedule", default=None,
                    help="team,inverse_lds,tables_lds,node_state_lds[,heavy_chains,heavy_team] "
                         "(e.g. 128,0,1,1,64,512): fixed schedule, no autotuning")
    ap.add_argument("--calib-copy", type=int, default=0,
                    help="MiB copied by a plain device copy kernel first (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described "
                                   "in DESIGN.md and not part of the benchmark", "n_gpus": args.gpus}))
        sys.exit(2)
    nchain = args.nchain or {"cfg3": 1024, "cfg5": 4096}.get(args.workload, 1)
    descs, wdesc = workload(args.workload, nchain)
    prec = capi.PQ_PREC_MIXED if args.precision == "mixed" else capi.PQ_PREC_FP64
    if args.calib_copy:
        # a plain copy of known size: its FETCH_SIZE/WRITE_SIZE scale the counters
        import torch
        src = torch.ones(args.calib_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    sched = tuple(int(v) for v in args.schedule.split(",")) if args.schedule else None
    b_main, main_run = run_gpu(descs, prec, args.steps, args.warmup, args.seed, schedule=sched)
    bu_total = float(main_run["bu"].sum())
    value = bu_total / main_run["wall"]
    kernel_ms = float(np.mean(main_run["ms"]))
    w = 8
    bytes_launch = float(sum(bu_c * bytes_per_update(d, w) for bu_c, d in zip(main_run["bu"], descs))) / args.steps
    flops_launch = float(sum(bu_c * flops_per_update(d) for bu_c, d in zip(main_run["bu"], descs))) / args.steps
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(args.workload, args.precision)
    out = {
        "metric": METRIC, "value": value, "unit": "bead-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "warmup_sweeps_done": main_run["warmup_done"],
        "ms_per_step": 1e3 * main_run["wall"] / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64" if prec == capi.PQ_PREC_MIXED else "fp64",
        "data": "synthetic (seeded Gaussian beads, replicated across slices for node chains)",
        "config": {"workload": f"{args.workload}: {wdesc}", "chains": len(descs),
                   "precision": args.precision, "step": "one sweep of every chain (one kernel launch)",
                   "schedule": main_run["schedule"]},
        "accepted_bead_updates_per_s": float(main_run["abu"].sum()) / main_run["wall"],
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src, "algorithmic_bytes": bytes_launch,
                     "kernel": "sweep_kernel", "kernel_ms": kernel_ms,
                     "fp64_equiv_tflops": flops_launch / (kernel_ms * 1e-3) / 1e12,
                     "valu_frac_of_fp64_peak": flops_launch / (kernel_ms * 1e-3) / 1e12 / FP64_VECTOR_TFLOPS},
    }
    if not args.no_fp64_pair and prec == capi.PQ_PREC_MIXED:
        # the same number of sweeps (the mixed run's autotuning sweeps are
        # matched by plain warmup sweeps) on the library's default schedule:
        # schedules change where data lives, not the decisions, and the mixed
        # run's LDS placement does not suit FP64 tables and node state
        b64, ref = run_gpu(descs, capi.PQ_PREC_FP64, args.steps, args.warmup, args.seed,
                           autotune=False, total_warmup=main_run["warmup_done"])
        out["fp64_batched"] = {"value": float(ref["bu"].sum()) / ref["wall"],
                               "kernel_ms": float(np.mean(ref["ms"]))}
        out["energy_vs_fp64"] = paired_energy(b_main, b64, main_run["energy"], ref["energy"])
        out["observables_vs_fp64"] = paired_observables(b_main, b64)
        single = [capi.cfg2_desc()]
        _, one = run_gpu(single, capi.PQ_PREC_FP64, max(2, args.steps // 5), 1, args.seed, autotune=False)
        r1 = float(one["bu"].sum()) / one["wall"]
        out["fp64_single_chain"] = {"workload": "cfg2: 3+3 e- 2D dot, SHONodes, M=400", "value": r1}
        out["ratio_vs_fp64_single_chain"] = value / r1
    if not args.no_cpu:
        cb = cpu_baseline(descs)
        out["cpu_baseline"] = cb
        out["ratio_vs_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
