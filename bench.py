#!/usr/bin/env python3
"""Bead-updates/s of the bisection-move hot path on one MI355X.

A step is one sweep of every chain of the batch followed by one Measure: the
sweep is a captured HIP graph of the fused sweep_kernel launch(es) (choose
section, node caches, every Sample's levels L..0, write back), timed by HIP
events around its launch on the batch stream; the Measure is the thermal
energy of every chain and the 256-bin pair-correlation histogram of the
first species against both species, accumulated on the device.

Default workload: 4096 independent 3D quantum-dot chains, N ~ U{2..64} in two
species, M = 16 U{16..128}, SHONodes on both species (node matrices up to
32 x 32), 4-level bisection, mixed precision.  Beside it: the same batch in
FP64 on the same uniforms (FP64 batched rate, paired energies); the
1024-chain uneven 2D batch as a second line; a paired mixed/FP64 energy study
over 1000 measured sweeps of the 2D batch; the FP64 single-chain rate of the
3+3-electron SHONodes dot; and the oracle, the FP64 C restatement, on a
bounded sample of the headline chains on all host cores and on one core.
Prints one JSON line.
"""
import argparse
import ctypes as C
import json
import math
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
FP64_PEAK_TFLOPS = 78.6
TUNE_SWEEPS = 2         # sweeps per timed schedule in the autotuner
STEADY = {"batch3d": 120, "batch2d": 220}   # sweeps from the replicated start before the warm-up steps
PAIR_CF = ([0, 0, 1], 0.0, 8.0, 256)  # species 0 against species 0 and 1, r in [0, 8), 256 bins


def log(msg):
    """Progress on stderr (stdout carries the one JSON line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def profile_figure(workload_name, precision, key):
    """A per-step figure (`key: value` line) of the committed rocprofv3
    summary of this workload, or None: traffic_bytes_per_step (every
    sweep_kernel launch of one step, copy-calibrated FETCH_SIZE +
    WRITE_SIZE), mfma_fp64_flops_per_step (SQ_INSTS_VALU_MFMA_MOPS_F64)."""
    path = os.path.join(ROOT, "profiles", f"sweep_kernel_{workload_name}_{precision}.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def measured_traffic(workload_name, precision):
    return profile_figure(workload_name, precision, "traffic_bytes_per_step")


def workload(name, nchain):
    if name == "batch2d":
        return capi.batch2d_descs(nchain), "1024 uneven 2D dots (N 2-20, M 128-1024, SHO/free nodes, L=4)"
    if name == "batch3d":
        return capi.batch3d_descs(nchain), "4096 uneven 3D dots (N 2-64 in two species, M 256-2048, SHO nodes, L=4)"
    if name == "dot2d":
        return [capi.dot2d_3p3_desc()] * nchain, "3+3 e- 2D dot, SHONodes, M=400, L=4"
    if name == "dot3d":
        return [capi.dot3d_2e_desc()] * nchain, "2 e- 3D dot, Coulomb, M=200, L=3"
    if name == "ehbox":
        return [capi.ehbox_desc()] * nchain, "32 e + 32 h, box L=10, FreeParticleNodes 32x32, M=1024, L=5"
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
    capi.check(b.lib.pq_get_all_stats(b.h, st))
    return st


def all_paths(b):
    arr = np.empty(int(sum(d.nbeads() for d in b.descs)), dtype=np.float64)
    capi.check(b.lib.pq_get_all_paths(b.h, capi.dptr(arr), arr.size))
    return arr


def identical_chains(bm, br, em, er):
    """Chains whose energies were bitwise equal at every measured sweep and
    whose final paths are bitwise equal in both batches: their decision
    streams never differed over the measurement.  (Equal final paths alone do
    not show it: a decision flipped at level 0 draws as many uniforms as the
    other outcome, so the streams stay aligned and a later accepted move over
    the same slices can make the paths equal again.)"""
    pm, pr = all_paths(bm), all_paths(br)
    same, off = [], 0
    for c, d in enumerate(bm.descs):
        n = d.nbeads()
        same.append(bool(np.array_equal(pm[off:off + n], pr[off:off + n])) and
                    bool(np.array_equal(em[:, c], er[:, c])))
        off += n
    return np.array(same)


def tune(descs, precision, seed, steady):
    """The autotuner's pick on a batch of its own (`steady` sweeps from the
    replicated start, then the timed schedules), so that the measured runs
    keep one schedule from their first sweep: the node-matrix width a chain
    is evaluated with changes the rounding, and a paired mixed/FP64
    comparison needs the same arithmetic apart from the precision."""
    b = make_batch(descs, precision, seed)
    b.run_sweeps(steady)
    ms, _ = b.autotune(TUNE_SWEEPS)
    team, inv, _ = b.schedule()
    tab, fst = b.staging()
    heavy, heavy_team = b.split()
    del b
    log(f"autotuned: team {team} inverse {inv} tables {tab} state {fst} split {heavy}x{heavy_team} ({ms:.2f} ms)")
    return (team, inv, tab, fst, heavy, heavy_team), ms


def run_gpu(descs, precision, steps, warmup, seed, steady, schedule=None, pair_cf=None, tuned=None):
    """`steady` sweeps from the replicated start, then `warmup` untimed steps
    and `steps` timed steps, each step one graph-launched sweep plus the
    Measure, all under one schedule = (team, inverse, tables, node state[,
    heavy chains, heavy team]), or the library's default."""
    b = make_batch(descs, precision, seed)
    if pair_cf:
        b.set_measure_pair_cf(*pair_cf)
    log(f"{len(descs)} chains, precision {precision}: batch ready")
    if schedule is not None:
        b.set_schedule(schedule[0], schedule[1])
        b.set_staging(schedule[2], schedule[3])
        if len(schedule) > 4 and schedule[4] != 0:
            b.set_split(schedule[4], schedule[5])
    b.run_sweeps(steady)
    b.sync()
    done = steady
    log(f"{done} sweeps done")
    if warmup > 0:
        b.run_measured_timed(warmup, 1)
        done += warmup
    if pair_cf:
        b.collect_pair_cf()
    b.sync()
    capi.check(b.lib.pq_reset_stats(b.h))
    t0 = time.perf_counter()
    energies, sweep_ms = b.run_measured_timed(steps, 1)
    wall = time.perf_counter() - t0
    counts = b.collect_pair_cf() if pair_cf else None
    log(f"{steps} timed steps in {wall:.2f} s")
    st = all_stats(b)
    bu = np.array([s.bead_updates for s in st], dtype=np.float64)
    abu = np.array([s.accepted_bead_updates for s in st], dtype=np.float64)
    team, inv, lds = b.schedule()
    tab, fst = b.staging()
    heavy, heavy_team = b.split()
    sched = {"team": team, "inverse_lds": inv, "tables_lds": tab, "node_state_lds": fst,
             "lds_bytes": lds, "heavy_chains": heavy, "heavy_team": heavy_team, "autotune_ms": tuned}
    return b, dict(wall=wall, sweep_ms=sweep_ms, bu=bu, abu=abu, energies=energies, counts=counts,
                   schedule=sched, warmup_done=done)


def roofline(descs, run, steps):
    kernel_ms = float(np.mean(run["sweep_ms"]))
    bytes_launch = float(sum(bu * bytes_per_update(d) for bu, d in zip(run["bu"], descs))) / steps
    flops_launch = float(sum(bu * flops_per_update(d) for bu, d in zip(run["bu"], descs))) / steps
    achieved = bytes_launch / (kernel_ms * 1e-3) / 1e9
    return kernel_ms, {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": bytes_launch,
                       "kernel": "sweep_kernel", "kernel_ms": kernel_ms,
                       "fp64_equiv_tflops": flops_launch / (kernel_ms * 1e-3) / 1e12,
                       "fp64_frac": flops_launch / (kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS}


def paired_energy(bm, br, em, er, label):
    """Mixed vs FP64 on the same seeds and sweeps: per-chain energies averaged
    over the measured sweeps, their paired difference over all chains (mean
    and standard error), and the chains whose paths stayed bitwise identical
    (both precisions move beads in FP64 and estimate in FP64, so a chain whose
    decisions never differed has dE = 0 exactly)."""
    em, er = np.asarray(em), np.asarray(er)
    same = identical_chains(bm, br, em, er)
    de = em.mean(axis=0) - er.mean(axis=0)
    n = de.size
    err = float(de.std(ddof=1) / math.sqrt(n)) if n > 1 else None
    out = {"measured_sweeps": int(em.shape[0]), "chains": n, "identical_decision_chains": int(same.sum()),
           "max_abs_dE_identical": float(np.max(np.abs(de[same]))) if same.any() else None,
           "mean_dE": float(de.mean()), "stderr_dE": err,
           "mean_E_fp64": float(er.mean()), "note": label}
    if err is not None:
        out["resolves_1e-5"] = bool(err < 1e-5)
    return out


def energy_study(nsweep, seed, steady):
    """The 1024-chain 2D batch in both precisions on the same Philox streams:
    `steady` sweeps, then nsweep sweeps each followed by a Measure.  Reports
    the paired difference of the per-chain time-averaged energies, and the
    two batch-mean energies with blocked errors (blocks of 50 sweeps)."""
    descs, _ = workload("batch2d", 1024)
    runs = {}
    for name, prec in (("mixed", capi.PQ_PREC_MIXED), ("fp64", capi.PQ_PREC_FP64)):
        b = make_batch(descs, prec, seed)
        b.run_sweeps(steady)
        e = b.run_measured(nsweep, 1)
        runs[name] = (b, e)
        log(f"energy study: {name} done")
    (bm, em), (br, er) = runs["mixed"], runs["fp64"]
    out = paired_energy(bm, br, em, er, f"1024-chain 2D batch, {nsweep} measured sweeps after {steady}, "
                                        "per-chain time averages")
    for name, e in (("mixed", em), ("fp64", er)):
        blocks = e.mean(axis=1)[: (nsweep // 50) * 50].reshape(-1, 50).mean(axis=1)
        out[f"E_{name}"] = float(blocks.mean())
        out[f"E_{name}_err"] = float(blocks.std(ddof=1) / math.sqrt(len(blocks)))
    d = out["E_mixed"] - out["E_fp64"]
    out["batch_means_agree_2sigma"] = bool(abs(d) <= 2 * math.hypot(out["E_mixed_err"], out["E_fp64_err"]))
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(descs, budget_s=12.0, single_s=8.0):
    """The oracle (FP64 C restatement) on host cores: a bounded sample of the
    workload's chains spread evenly over the batch, one chain per thread
    (OpenMP), every sweep followed by the same Measure as the device's
    (thermal energy and the pair-correlation counts), on all cores and on
    one core."""
    import oracle_lib as ol
    lib = ol.load()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    nch = min(len(descs), 2 * cores)
    stride = max(1, len(descs) // nch)
    sample = [descs[k * stride] for k in range(nch)]
    sp = np.ascontiguousarray(PAIR_CF[0], dtype=np.int32)

    def run(chains, nsweep, threads):
        arr = capi.descs_array(chains)
        beads = np.concatenate([capi.initial_beads(d, 1000 + c).ravel() for c, d in enumerate(chains)])
        es = np.zeros(len(chains))
        counts = np.zeros(len(chains) * PAIR_CF[3], dtype=np.uint64)
        t0 = time.perf_counter()
        bu = lib.oq_run_batch_measured(arr, len(chains), capi.dptr(beads), nsweep, 7, 0, threads, len(sp),
                                       capi.iptr(sp), PAIR_CF[1], PAIR_CF[2], PAIR_CF[3], capi.dptr(es),
                                       counts.ctypes.data_as(C.POINTER(C.c_uint64)))
        return bu, time.perf_counter() - t0

    _, t1 = run(sample, 1, cores)
    nsweep = max(1, min(1000, int(budget_s / max(t1, 1e-3))))
    bu, dt = run(sample, nsweep, cores)
    one = sample[::4]
    _, t1 = run(one, 1, 1)
    ns1 = max(1, min(1000, int(single_s / max(t1, 1e-3))))
    bu1, dt1 = run(one, ns1, 1)
    return dict(value=bu / dt, unit="bead-updates/s", cores=cores, kind="port",
                sample=f"{nch} chains spread over the workload (every {stride}th) x {nsweep} sweeps, each sweep "
                       f"followed by the Measure (thermal energy, 256-bin g(r)); oracle (FP64 C, -O3 "
                       f"-march=x86-64-v3, OpenMP {cores} threads), {dt:.1f} s",
                cpu_model=cpu_model(), nproc=os.cpu_count(),
                single_core={"value": bu1 / dt1, "sample": f"{len(one)} of those chains x {ns1} sweeps with the "
                                                            f"Measure on one thread, {dt1:.1f} s"})


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="batch3d")
    ap.add_argument("--nchain", type=int, default=None)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--steady", type=int, default=None,
                    help="sweeps from the replicated start before the warmup steps (tuning included)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64-pair", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="headline only: no second workload, energy study or single-chain rate")
    ap.add_argument("--energy-sweeps", type=int, default=1000)
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
    nchain = args.nchain or {"batch2d": 1024, "batch3d": 4096}.get(args.workload, 1)
    descs, wdesc = workload(args.workload, nchain)
    prec = capi.PQ_PREC_MIXED if args.precision == "mixed" else capi.PQ_PREC_FP64
    steady = args.steady if args.steady is not None else STEADY.get(args.workload, 40)
    if args.calib_copy:
        # a plain copy of known size: its FETCH_SIZE/WRITE_SIZE scale the counters
        import torch
        src = torch.ones(args.calib_copy * (1 << 20) // 8, dtype=torch.float64, device="cuda")
        dst = src.clone()
        torch.cuda.synchronize()
        del src, dst

    tuned = None
    if args.schedule:
        sched = tuple(int(v) for v in args.schedule.split(","))
    else:
        sched, tuned = tune(descs, prec, args.seed, steady)
    b_main, run = run_gpu(descs, prec, args.steps, args.warmup, args.seed, steady, schedule=sched, pair_cf=PAIR_CF,
                          tuned=tuned)
    bu_total = float(run["bu"].sum())
    value = bu_total / run["wall"]
    kernel_ms, roof = roofline(descs, run, args.steps)
    traffic, traffic_src = measured_traffic(args.workload, args.precision)
    roof["traffic"] = traffic
    roof["traffic_source"] = traffic_src
    mf, mf_src = profile_figure(args.workload, args.precision, "mfma_fp64_flops_per_step")
    if mf:
        # the node matrices' FP64 MFMA flops of one step (rocprofv3 count)
        # over the sweep span timed here
        roof["mfma_fp64"] = {"flops_per_step": mf, "achieved_tflops": mf / (kernel_ms * 1e-3) / 1e12,
                             "peak_tflops": FP64_PEAK_TFLOPS,
                             "frac": mf / (kernel_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, "source": mf_src}
    ms_step = 1e3 * run["wall"] / args.steps
    beads_per_measure = float(sum(d.nbeads() for d in descs))
    out = {
        "metric": METRIC, "value": value, "unit": "bead-updates/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "warmup_sweeps_done": run["warmup_done"],
        "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32+fp64" if prec == capi.PQ_PREC_MIXED else "fp64",
        "data": "synthetic (seeded Gaussian beads, replicated across slices for node chains)",
        "config": {"workload": f"{args.workload}: {wdesc}", "chains": len(descs),
                   "precision": args.precision,
                   "step": "one sweep of every chain (captured graph of the sweep_kernel launches) + one Measure "
                           "(thermal energy and 256-bin g(r) of every chain, on the device)",
                   "schedule": run["schedule"]},
        "accepted_bead_updates_per_s": float(run["abu"].sum()) / run["wall"],
        "measure_ms": ms_step - kernel_ms,
        "measure_bytes": 8 * beads_per_measure * 2,
        "roofline": roof,
        "pair_cf_counts": int(run["counts"].sum()) if run["counts"] is not None else None,
    }
    if not args.no_fp64_pair and prec == capi.PQ_PREC_MIXED:
        # the same sweeps from the same start with the mixed run's team size and
        # launch split from the first sweep, so that both runs evaluate node
        # matrices with the same Gauss-Jordan instance (LDS placement is left
        # automatic: it moves data, not arithmetic)
        sc = run["schedule"]
        b64, ref = run_gpu(descs, capi.PQ_PREC_FP64, args.steps, args.warmup, args.seed, steady, pair_cf=PAIR_CF,
                           schedule=(sc["team"], -1, -1, -1, sc["heavy_chains"], sc["heavy_team"]))
        out["fp64_batched"] = {"value": float(ref["bu"].sum()) / ref["wall"],
                               "kernel_ms": float(np.mean(ref["sweep_ms"]))}
        out["energy_vs_fp64"] = paired_energy(b_main, b64, run["energies"], ref["energies"],
                                              f"headline batch, the {args.steps} timed steps")
        del b64
    if not args.no_extras:
        if args.workload != "batch2d":
            d3, w3 = workload("batch2d", 1024)
            s3, t3ms = tune(d3, prec, args.seed, STEADY["batch2d"])
            b3, r3 = run_gpu(d3, prec, args.steps, args.warmup, args.seed, STEADY["batch2d"], schedule=s3,
                             pair_cf=PAIR_CF, tuned=t3ms)
            k3, roof3 = roofline(d3, r3, args.steps)
            t3, t3src = measured_traffic("batch2d", args.precision)
            roof3["traffic"], roof3["traffic_source"] = t3, t3src
            out["second_workload"] = {"workload": f"batch2d: {w3}", "value": float(r3["bu"].sum()) / r3["wall"],
                                      "ms_per_step": 1e3 * r3["wall"] / args.steps, "schedule": r3["schedule"],
                                      "roofline": roof3}
            del b3
        out["energy_study"] = energy_study(args.energy_sweeps, args.seed, STEADY["batch2d"])
        single = [capi.dot2d_3p3_desc()]
        _, one = run_gpu(single, capi.PQ_PREC_FP64, max(2, args.steps // 5), 1, args.seed, 20)
        r1 = float(one["bu"].sum()) / one["wall"]
        out["fp64_single_chain"] = {"workload": "dot2d: 3+3 e- 2D dot, SHONodes, M=400", "value": r1}
        out["ratio_vs_fp64_single_chain"] = value / r1
    if not args.no_cpu:
        log("cpu baseline")
        cb = cpu_baseline(descs)
        out["cpu_baseline"] = cb
        out["ratio_vs_cpu"] = value / cb["value"]
    print(json.dumps(out))


if __name__ == "__main__":
    main()
