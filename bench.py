"""Benchmark: K-sector ground states per second on one MI355X.

Default workload (the plain method at scale): spinless bilayer, N_phi = 16,
N_e = 16, all layer densities, sector (J,K) = (0,0), d/l = 1.0,
Delta_SAS = 0.02, gamma = 1, Lanczos cap 200, stop at |dE| < 1e-10
(2,348,125 states).  A step is one full ground-state solve (Lanczos pass 1,
tridiagonal pass 2, reconstruction and Rayleigh energy) with the Hamiltonian
resident in HBM; a step counts only if it reached |dE| < 1e-10.

    python bench.py [--steps K] [--warmup W] [--workload nphi16_all|nphi16_u8|nphi8_readme|sweep_nphi12|sweep_spin_nphi8|warm_nphi14]

Prints one JSON line.  The roofline object is for the dominant kernel (the
fused Lanczos SpMV), timed with HIP events on the library stream during the
timed solves.  cpu_baseline times the oracle's CSR SpMV (the reference's row
loop) over the whole matrix on 16 host threads and charges the reference's
fixed 200-step three-pass Lanczos (400 SpMVs) per ground state.  The default
run also times the sweep workloads (all 64 spinful N_phi = 8 sectors and all
144 N_phi = 12 sectors, --sweep-points parameter points each): batched
solves against plain single-problem solves, under "sweeps".
"""
import argparse
import json
from concurrent.futures import ThreadPoolExecutor
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "K-sector ground states/sec at |dE|<1e-10 (batched+warm vs plain); SpMV HBM GB/s"
HBM_PEAK_GBS = 8000.0
# The sweeps build one Hamiltonian per sector on several host threads: each
# thread queues its library work on a stream of its own, so one sector's
# short launches and host waits overlap another's
_BUILD_POOL = None


def build_map(f, xs):
    global _BUILD_POOL
    if _BUILD_POOL is None:
        _BUILD_POOL = ThreadPoolExecutor(4)
    return list(_BUILD_POOL.map(f, xs))

# HBM bytes per launch of the fused SpMV from rocprofv3 FETCH_SIZE / WRITE_SIZE
# passes, scaled by the same counters on a plain 1 GiB copy (profiles/);
# reported only when the live run uses the profiled workload and variant
MEASURED_TRAFFIC = {
    # k_jag<true, true, 8, 8>, internal row order, plain step order: 6,477,611.5 KiB FETCH_SIZE x 2.0 +
    # 36,706.9 KiB WRITE_SIZE x 1.0 per dispatch (profiles/k_jag_nphi16_counters.md)
    ("nphi16_all", (3, 64, 512, 2, 8)): (6477611.5 * 2.0 * 1024 + 36706.9 * 1024,
                                         "profiles/k_jag_nphi16_counters.md"),
}
# HBM bytes per batched K1 launch of the sweeps (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 KiB, mean over the
# run's K1 dispatches; profiles/*_counters.md), reported for the same batching mode
MEASURED_SWEEP_TRAFFIC = {
    ("sweep_spin_nphi8", "per-point"): (739444.2 * 2.0 * 1024 + 11263.0 * 1024,
                                        "profiles/k_b1j_sweep_spin_nphi8_counters.md"),
    ("sweep_nphi12", "shared"): (2438173.0 * 2.0 * 1024 + 69723.4 * 1024, "profiles/k_b1s_sweep_nphi12_counters.md"),
}
# fused Lanczos K1 kernel of each SpMV variant (lanczos.hip)
KERNEL_NAMES = {0: "k_lz1", 1: "k_pspmv<LZ=true>", 2: "k_sps<LZ=true>", 3: "k_jag<LZ=true>", 4: "k_t4<LZ=true>",
                5: "k_jagr<LZ=true>"}

WORKLOADS = {
    # name: (nspecies, nphi, nel, nel_up, J, K, d, delta_sas, label)
    "nphi16_all": (2, 16, 16, -1, 0, 0, 1.0, 0.02, "spinless bilayer Nphi=16 Ne=16 all densities (J,K)=(0,0) d=1.0 DSAS=0.02"),
    "nphi16_u8": (2, 16, 16, 8, 0, 0, 1.0, 0.02, "spinless bilayer Nphi=16 Ne=16 nel_up=8 (J,K)=(0,0) d=1.0 DSAS=0.02"),
    "nphi8_readme": (2, 8, 8, 4, 4, 4, 3.0, 0.0, "spinless bilayer Nphi=8 Ne=8 nel_up=4 (J,K)=(4,4) d=3.0"),
}


# Sweeps of all K-sectors x parameter points, solved by the batched engine.
SWEEPS = {
    "sweep_nphi12": dict(S=2, nphi=12, nel=12, nu=-1, nsu=-1, dv=0.0, dz=0.0,
               sectors=[(J, K) for J in range(12) for K in range(12)],
               points=[(round(0.2 * i, 1), ds) for i in range(16) for ds in (0.0, 0.05, 0.1)],
               label="spinless bilayer Nphi=12 Ne=12 all densities, all 144 (J,K) sectors, "
                     "d in {0,0.2,..,3.0} x DSAS in {0,0.05,0.1}"),
    # single layer, the Monolayer tree's K-resolved spectrum workload: all 108
    # (J, K) sectors of N_phi = 18, N_e = 6 (K < gcd = 6); one parameter point,
    # repeated (a step solves every sector once more, H resident)
    "mono_ek_nphi18": dict(S=1, nphi=18, nel=6, nu=-1, nsu=-1, dv=0.0, dz=0.0,
               sectors=[(J, K) for J in range(18) for K in range(6)],
               points=[(0.0, 0.0)] * 64,
               label="monolayer Nphi=18 Ne=6 (Monolayer conventions), all 108 (J,K) sectors, lx/ly=1"),
    "sweep_spin_nphi8": dict(S=4, nphi=8, nel=8, nu=4, nsu=4, dv=0.02, dz=0.01,
               sectors=[(J, K) for J in range(8) for K in range(8)],
               points=[(d, 0.05) for d in (0.5, 1.0, 1.5, 2.0)],
               label="spinful bilayer Nphi=8 Ne=8 nel_up=4 nel_su=4, all 64 (J,K) sectors, "
                     "DZ=0.01 DSAS=0.05 DV=0.02, d in {0.5,1.0,1.5,2.0}"),
}


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_full(H, iters_ref=200, threads=16, reps=3):
    """The oracle's CSR row loop (the reference's SpMV, Mat::operator*) over the
    whole matrix exported from the GPU build: one full, unsampled SpMV, median
    of `reps`, times the 2*iters_ref SpMVs of the reference's fixed-length
    three-pass Lanczos (pass 1, pass-3 replay, energy)."""
    import oracle as O
    t0 = time.time()
    rp, col, val = H.export()
    t_export = time.time() - t0
    n = H.n
    rng = np.random.default_rng(0)
    x = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex128)
    y = np.empty(n, np.complex128)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        O.lib().orc_spmv32(n, rp.ctypes.data, col.ctypes.data, val.ctypes.data, x.ctypes.data, y.ctypes.data,
                           threads)
        ts.append(time.perf_counter() - t0)
    t_spmv = float(np.median(ts))
    del rp, col, val
    t_solve = 2 * iters_ref * t_spmv
    return {
        "value": 1.0 / t_solve,
        "unit": "K-sector ground states/s",
        "cores": threads,
        "kind": "port",
        "cpu": cpu_model(),
        "sample": (f"oracle CSR SpMV over the full {n}-row, {H.nnz}-nnz matrix (not sampled), median of {reps}: "
                   f"{t_spmv:.3f} s ({[round(t, 3) for t in ts]}); x {2 * iters_ref} SpMVs of the reference's "
                   f"fixed {iters_ref}-step three-pass Lanczos = {t_solve:.0f} s per ground state; vector updates "
                   f"not counted; H resident (host copy exported in {t_export:.1f} s)"),
        "spmv_s": t_spmv,
        "threads": threads,
    }


def sweep_cpu_baseline(sw, iters_ref=200):
    """Oracle (CPU port of the reference: dense-row build, fixed 200-step
    three-pass Lanczos) on one sector of the sweep, 16 threads."""
    import oracle as O
    J, K = sw["sectors"][0]
    d, ds = sw["points"][0]
    if sw["S"] == 1:  # monolayer: the oracle's Monolayer-tree restatement (basis and rows together)
        t0 = time.time()
        m = O.monolayer(sw["nphi"], sw["nel"], J, K)
        t_build = time.time() - t0
        A = O.CsrMatrix(m["rowptr"], m["col"], m["val"])
        reps, t0 = 0, time.time()
        while reps < 5 or time.time() - t0 < 2.0:  # a sector takes milliseconds: repeat
            A.lanczos(lam=iters_ref, seed=1 + reps)
            reps += 1
        t_solve = (time.time() - t0) / reps
        return {
            "value": 1.0 / t_solve, "unit": "K-sector ground states/s", "cores": 1, "kind": "port",
            "sample": (f"oracle on sector (J,K)=({J},{K}) (N={A.n}, nnz={A.nnz}): fixed {iters_ref}-step three-pass "
                       f"Lanczos {t_solve * 1e3:.2f} ms (mean of {reps}, H resident); basis and rows "
                       f"{t_build:.2f} s"),
        }
    sec = O.sector(S=sw["S"], nphi=sw["nphi"], nel=sw["nel"], nel_up=sw["nu"], nel_su=sw["nsu"], J=J, K=K)
    par = O.params(sw["nphi"], d=d, dsas=ds, dv=sw["dv"], dz=sw["dz"])
    threads = min(16, os.cpu_count() or 1)
    t0 = time.time()
    OB = O.Basis(sec)
    t_basis = time.time() - t0
    t0 = time.time()
    OH = O.Hamiltonian(OB, par, nthreads=threads)
    t_build = time.time() - t0
    t0 = time.time()
    OH.lanczos(lam=iters_ref, seed=1)
    t_solve = time.time() - t0
    return {
        "value": 1.0 / t_solve,
        "unit": "K-sector ground states/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle on sector (J,K)=({J},{K}) at d={d}, DSAS={ds} (N={len(OB)}, nnz={OH.nnz}): "
                   f"fixed {iters_ref}-step three-pass Lanczos {t_solve:.2f} s (solve only, H resident); "
                   f"row build {t_build:.2f} s, basis {t_basis:.2f} s"),
    }


def run_sweep(args):
    import blqh_amd
    if args.events_all:
        blqh_amd.profile_enable(True)  # every batched launch event-timed: no graph replays
    sw = SWEEPS[args.workload]
    S, nphi, nel = sw["S"], sw["nphi"], sw["nel"]
    log(f"workload {args.workload}: {sw['label']}")
    t0 = time.time()
    bases = [blqh_amd.Basis(nphi, nel, sw["nu"], J, K, nspecies=S, nel_su=sw["nsu"]) for J, K in sw["sectors"]]
    t_basis = time.time() - t0
    log(f"{len(bases)} bases, N {min(b.n for b in bases)}..{max(b.n for b in bases)}, {t_basis:.1f} s")
    pts = sw["points"]
    nwarm = max(1, args.warmup)
    order = pts[:nwarm] + pts[nwarm:nwarm + args.steps]
    seeds = list(range(1, len(bases) + 1))

    # H(d, DSAS) = H(d, 0) + DSAS T: the tunnelling-free part is built once per d
    # (points are ordered d-major), the tunnelling matrix T once per sector
    # (it does not depend on d without theta_B phases), and every DSAS > 0
    # point is a merge of the two
    cache = {"d": None, "base": None, "tun": None}

    def par(d, ds):
        return blqh_amd.params(nphi, d=d, delta_sas=ds, delta_v=sw["dv"], delta_z=sw["dz"])

    def build(d, ds):
        t = time.time()
        if cache["d"] != d:
            cache["base"] = None
            cache["base"] = build_map(lambda B: blqh_amd.Hamiltonian(B, par(d, 0.0), terms=1), bases)
            cache["d"] = d
        if ds > 0:
            if cache["tun"] is None:
                cache["tun"] = build_map(lambda B: blqh_amd.Hamiltonian(B, par(d, 0.0), terms=2), bases)
            hs = build_map(lambda bt: blqh_amd.Hamiltonian.combine(bt[0], bt[1], ds), zip(cache["base"], cache["tun"]))
        else:
            hs = build_map(lambda H: H.prepare(), cache["base"])  # solved directly: its SpMV formats count as build
        return hs, time.time() - t

    # convergence checks: plain solves every --check-every steps, the batch
    # every --batch-check-every (a batch runs one speculative replay past
    # each check, so a shorter interval wastes fewer steps per problem)
    bchk = args.batch_check_every or args.check_every
    plain = None
    batch = None
    solved = solves = 0
    t_solve = t_build = 0.0
    iters, e_by_point = [], {}
    # continuation starts: every sector starts from its ground state at the
    # previous point, which the batch keeps on the device (conj(psi): psi =
    # sum c_l conj(v_l) is the eigenvector of conj(H)); run before the cold
    # batched solve of the same point, which checks it
    cont = {"solved": 0, "seconds": 0.0, "iters": [], "max_rel_dE0": 0.0, "mismatch": 0}
    for k, (d, ds) in enumerate(order):
        hs, tb = build(d, ds)
        warm = k < nwarm
        prev = batch.hams if batch is not None else None  # freed after the timed solve
        t = time.perf_counter()
        if batch is None:
            batch = blqh_amd.Batch(hs)  # autotunes the batched SpMV once for the sweep
        else:
            batch.rebind(hs)  # same sectors, next parameter point: pools and replay kept
        dt = time.perf_counter() - t
        wres = None
        if args.continuation and k > 0:
            t = time.perf_counter()
            wres = batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False,
                                 start_mix=args.start_mix, continue_prev=True,
                                 excited_mix=args.excited_mix)
            dtw = time.perf_counter() - t
        t = time.perf_counter()
        res = batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False)
        dt += time.perf_counter() - t
        t = time.time()
        del prev  # the previous point's Hamiltonians (device frees)
        tb += time.time() - t
        ok = sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in res)
        log(f"{'warmup' if warm else 'point'} d={d} DSAS={ds}: build {tb:.2f} s, batched solve {dt:.3f} s, "
            f"{ok}/{len(res)} converged, iterations {min(r['iterations'] for r in res)}.."
            f"{max(r['iterations'] for r in res)}")
        if wres is not None:
            wok = [r["converged"] == 1 and r["dE"] < 1e-10 for r in wres]
            rel = [abs(a["E0"] - b["E0"]) / abs(b["E0"]) for a, b in zip(wres, res)]
            bad = sum(int(x > 1e-10) for x, w in zip(rel, wok) if w)
            log(f"  continuation: batched solve {dtw:.3f} s, {sum(wok)}/{len(wres)} converged, iterations "
                f"{min(r['iterations'] for r in wres)}..{max(r['iterations'] for r in wres)}, "
                f"max |E0 - cold E0|/|E0| = {max(rel):.1e}")
            if not warm:
                cont["solved"] += sum(int(w and x <= 1e-10) for w, x in zip(wok, rel))
                cont["seconds"] += dtw
                cont["iters"] += [r["iterations"] for r in wres]
                cont["max_rel_dE0"] = max(cont["max_rel_dE0"], max(rel))
                cont["mismatch"] += bad
        if warm and plain is None:
            # plain single-problem solves of the same Hamiltonians (each autotuned by a first solve)
            for i, H in enumerate(hs):
                H.lanczos(lam=args.lam, seed=seeds[i], tol=1e-10, check_every=args.check_every, want_psi=False)
            tp = time.perf_counter()
            pres = [H.lanczos(lam=args.lam, seed=seeds[i], tol=1e-10, check_every=args.check_every, want_psi=False)
                    for i, H in enumerate(hs)]
            tp = time.perf_counter() - tp
            pok = sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in pres)
            dmax = max(abs(a["E0"] - b["E0"]) / abs(b["E0"]) for a, b in zip(res, pres))
            plain = {"solves": len(pres), "seconds": tp, "rate": pok / tp, "max_rel_dE0_vs_batched": dmax}
            log(f"plain: {len(pres)} single-problem solves in {tp:.3f} s ({pok / tp:.1f}/s); "
                f"max |E0 batched - plain|/|E0| = {dmax:.1e}")
        if not warm:
            solved += ok
            solves += len(res)
            t_solve += dt
            t_build += tb
            iters += [r["iterations"] for r in res]
        if k == len(order) - 1:
            # roofline pass: one more solve with HIP events around every batched K1 launch
            blqh_amd.profile_reset()
            blqh_amd.profile_enable(True)
            batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False)
            blqh_amd.profile_enable(args.events_all)
            launches, kms, kbytes = blqh_amd.profile_read()
            nnz_total = sum(H.nnz for H in hs)
        del hs
    # multi-point leg: the timed points' problems (every sector at every
    # timed point) in one batch, so that no point's last steps run alone
    multi = None
    if args.multi_point and len(order) - nwarm > 1 and len(set(order[nwarm:])) == len(order) - nwarm:
        pts_t = order[nwarm:]
        hs_all = []
        for d, ds in pts_t:
            hs_all += build(d, ds)[0]
        mb = blqh_amd.Batch(hs_all)
        mseeds = seeds * len(pts_t)
        mb.lanczos(lam=args.lam, seeds=mseeds, tol=1e-10, check_every=bchk, want_psi=False)  # warm-up
        t = time.perf_counter()
        mres = mb.lanczos(lam=args.lam, seeds=mseeds, tol=1e-10, check_every=bchk, want_psi=False)
        dtm = time.perf_counter() - t
        mok = sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in mres)
        multi = {"points": len(pts_t), "solves": len(mres), "converged": mok, "seconds": dtm, "rate": mok / dtm}
        log(f"multi-point batch: {len(pts_t)} points, {len(mres)} problems in one batch: {dtm:.3f} s, "
            f"{mok}/{len(mres)} converged ({mok / dtm:.1f}/s)")
        del mb, hs_all
    # pipelined leg: the timed points again, each point's Hamiltonians built on
    # a second host thread (its own library stream) while the previous point
    # solves, so the builds overlap the batched solves on the GPU
    pipe = None
    if args.pipeline and len(order) > nwarm:
        pts_t = order[nwarm:]
        ex = ThreadPoolExecutor(1)
        t = time.perf_counter()
        fut = ex.submit(build, *pts_t[0])
        psolved = 0
        for k, (d, ds) in enumerate(pts_t):
            hs, _ = fut.result()
            if k + 1 < len(pts_t):
                fut = ex.submit(build, *pts_t[k + 1])
            prev = batch.hams
            batch.rebind(hs)
            res = batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False)
            del prev, hs
            psolved += sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in res)
        wall = time.perf_counter() - t
        ex.shutdown()
        pipe = {"points": len(pts_t), "converged": psolved, "seconds": wall, "rate": psolved / wall}
        log(f"pipelined builds and solves: {len(pts_t)} points, {psolved} solves in {wall:.3f} s "
            f"({psolved / wall:.1f}/s end to end)")
    avg_ms = kms / max(launches, 1)
    achieved = (kbytes / max(launches, 1)) / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    value = solved / t_solve if t_solve > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "K-sector ground states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": nwarm,
        "ms_per_step": t_solve / max(1, args.steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": sw["label"],
            "step": "all sectors at one parameter point, batched solve with H resident",
            "sectors": len(bases),
            "points_timed": args.steps,
            "solves_timed": solves,
            "converged": solved,
            "lanczos_cap": args.lam,
            "tolerance": 1e-10,
            "check_every": args.check_every,
            "batch_check_every": bchk,
            "iterations_mean": float(np.mean(iters)) if iters else None,
            "batched_rate": value,
            "plain_rate": plain["rate"] if plain else None,
            "batched_over_plain": value / plain["rate"] if plain and plain["rate"] else None,
            "plain": plain,
            "end_to_end_rate_incl_build": solved / (t_solve + t_build) if t_solve + t_build > 0 else None,
            "end_to_end_rate_pipelined": pipe["rate"] if pipe else None,
            "pipelined": pipe,
            "batched_multipoint_rate": multi["rate"] if multi else None,
            "multipoint": multi,
            "continuation_rate": cont["solved"] / cont["seconds"] if cont["seconds"] > 0 else None,
            "continuation_iterations_mean": float(np.mean(cont["iters"])) if cont["iters"] else None,
            "continuation_max_rel_dE0_vs_cold": cont["max_rel_dE0"] if cont["iters"] else None,
            "continuation_mismatches": cont["mismatch"] if cont["iters"] else None,
            "continuation_vs_plain": (cont["solved"] / cont["seconds"]) / plain["rate"]
            if cont["seconds"] > 0 and plain else None,
            "build_s_per_point": t_build / max(1, args.steps),  # build + teardown of the point's Hamiltonians
            "build_scheme": "H(d,0) once per d (every third point), H(d,DSAS) = H(d,0) + DSAS*T by merge, "
                            "T once per sector (timed in the first DSAS>0 point)",
            "basis_s_once": round(t_basis, 2),
            "batched_spmv_lanes": batch.lanes,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (MEASURED_SWEEP_TRAFFIC.get((args.workload, "per-point")) or (None,))[0],
            "traffic_source": (MEASURED_SWEEP_TRAFFIC.get((args.workload, "per-point")) or (None, None))[1],
            "kernel": ("k_b1j<LZ=true> (batched fused SpMV + Lanczos K1, step-major panels, two wavefronts per "
                       "64-row slice)" if batch.lanes == 64 else
                       f"k_b1<{batch.lanes}, LZ=true> (batched fused SpMV + Lanczos K1, row-major panels)"),
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "algorithmic_bytes_per_launch": kbytes / max(launches, 1),
            "batch_nnz": nnz_total,
            "measured": "separate profiled solve of the last point (plain launches with HIP events)",
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on one sector)")
        try:
            out["cpu_baseline"] = sweep_cpu_baseline(sw, iters_ref=args.lam)
        except Exception as e:  # report, never hide, a failed baseline
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    return out


def run_sweep_shared(args):
    """A sweep whose d values each carry several DSAS values: one step is one
    d, all sectors at all its DSAS values in one shared-scale batch
    (Batch.scaled: each sector's H(d,0) + T entries are read once per Lanczos
    step for all of its DSAS problems).  Plain rate: single-problem solves of
    the warm-up d's merged Hamiltonians H(d,0) + DSAS T, every sector and DSAS."""
    import blqh_amd
    if args.events_all:
        blqh_amd.profile_enable(True)  # every batched launch event-timed: no graph replays
    sw = SWEEPS[args.workload]
    S, nphi, nel = sw["S"], sw["nphi"], sw["nel"]
    log(f"workload {args.workload} (shared-scale batches): {sw['label']}")
    t0 = time.time()
    bases = [blqh_amd.Basis(nphi, nel, sw["nu"], J, K, nspecies=S, nel_su=sw["nsu"]) for J, K in sw["sectors"]]
    t_basis = time.time() - t0
    log(f"{len(bases)} bases, N {min(b.n for b in bases)}..{max(b.n for b in bases)}, {t_basis:.1f} s")
    d_list, ds_list = [], []
    for d, ds in sw["points"]:
        if d not in d_list:
            d_list.append(d)
        if ds not in ds_list:
            ds_list.append(ds)
    G = len(ds_list)
    nwarm = max(1, args.warmup)
    order = d_list[:nwarm] + d_list[nwarm:nwarm + args.steps]
    seeds = list(range(1, len(bases) * G + 1))
    bchk = args.batch_check_every or args.check_every

    def par(d):
        return blqh_amd.params(nphi, d=d, delta_sas=0.0, delta_v=sw["dv"], delta_z=sw["dz"])

    tun = None  # T does not depend on d without theta_B phases: built once
    batch = None
    cont = {"solved": 0, "seconds": 0.0, "iters": [], "max_rel_dE0": 0.0, "mismatch": 0}
    plain = None
    solved = solves = 0
    t_solve = t_build = 0.0
    iters = []
    for k, d in enumerate(order):
        warm = k < nwarm
        t = time.time()
        base = build_map(lambda B: blqh_amd.Hamiltonian(B, par(d), terms=1), bases)
        if tun is None:
            tun = build_map(lambda B: blqh_amd.Hamiltonian(B, par(d), terms=2), bases)
        if batch is None:
            batch = blqh_amd.Batch.scaled(base, tun, ds_list)
        else:
            batch.rebind_scaled(base, tun)
        tb = time.time() - t
        wres = None
        if args.continuation and k > 0:
            # every problem starts from its ground state at the previous d (same
            # sector and DSAS, kept on the device), checked against the cold
            # solve of this d below
            t = time.perf_counter()
            wres = batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False,
                                 start_mix=args.start_mix, continue_prev=True,
                                 excited_mix=args.excited_mix)
            dtw = time.perf_counter() - t
        t = time.perf_counter()
        res = batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False)
        dt = time.perf_counter() - t
        ok = sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in res)
        log(f"{'warmup' if warm else 'step'} d={d} DSAS={ds_list}: build {tb:.2f} s, shared-scale batched solve "
            f"{dt:.3f} s, {ok}/{len(res)} converged, iterations {min(r['iterations'] for r in res)}.."
            f"{max(r['iterations'] for r in res)}")
        if wres is not None:
            wok = [r["converged"] == 1 and r["dE"] < 1e-10 for r in wres]
            rel = [abs(a["E0"] - b["E0"]) / abs(b["E0"]) for a, b in zip(wres, res)]
            bad = sum(int(x > 1e-10) for x, w in zip(rel, wok) if w)
            log(f"  continuation: batched solve {dtw:.3f} s, {sum(wok)}/{len(wres)} converged, iterations "
                f"{min(r['iterations'] for r in wres)}..{max(r['iterations'] for r in wres)}, "
                f"{bad} above the cold E0 by more than 1e-10")
            if not warm:
                cont["solved"] += sum(int(w and x <= 1e-10) for w, x in zip(wok, rel))
                cont["seconds"] += dtw
                cont["iters"] += [r["iterations"] for r in wres]
                cont["max_rel_dE0"] = max(cont["max_rel_dE0"], max(rel))
                cont["mismatch"] += bad
        if warm and plain is None:
            # plain single-problem solves of the same problems (each autotuned by a first solve)
            hs = [blqh_amd.Hamiltonian.combine(b, tn, x) for b, tn in zip(base, tun) for x in ds_list]
            for i, H in enumerate(hs):
                H.lanczos(lam=args.lam, seed=seeds[i], tol=1e-10, check_every=args.check_every, want_psi=False)
            tp = time.perf_counter()
            pres = [H.lanczos(lam=args.lam, seed=seeds[i], tol=1e-10, check_every=args.check_every, want_psi=False)
                    for i, H in enumerate(hs)]
            tp = time.perf_counter() - tp
            pok = sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in pres)
            dmax = max(abs(a["E0"] - b["E0"]) / abs(b["E0"]) for a, b in zip(res, pres))
            plain = {"solves": len(pres), "seconds": tp, "rate": pok / tp, "max_rel_dE0_vs_batched": dmax}
            log(f"plain: {len(pres)} single-problem solves in {tp:.3f} s ({pok / tp:.1f}/s); "
                f"max |E0 batched - plain|/|E0| = {dmax:.1e}")
            del hs, pres
        if not warm:
            solved += ok
            solves += len(res)
            t_solve += dt
            t_build += tb
            iters += [r["iterations"] for r in res]
        if k == len(order) - 1:
            blqh_amd.profile_reset()
            blqh_amd.profile_enable(True)
            batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False)
            blqh_amd.profile_enable(args.events_all)
            launches, kms, kbytes = blqh_amd.profile_read()
            nnz_total = sum(H.nnz for H in base) + sum(H.nnz for H in tun)
        del base  # the batch keeps its own copy of H(d,0) + T
    # multi-point leg: the timed d values' groups in one shared-scale batch
    multi = None
    if args.multi_point and len(order) - nwarm > 1:
        pts_t = order[nwarm:]
        base_all = []
        for d in pts_t:
            base_all += build_map(lambda B, d=d: blqh_amd.Hamiltonian(B, par(d), terms=1), bases)
        mb = blqh_amd.Batch.scaled(base_all, tun * len(pts_t), ds_list)
        del base_all
        mseeds = seeds * len(pts_t)
        mb.lanczos(lam=args.lam, seeds=mseeds, tol=1e-10, check_every=bchk, want_psi=False)  # warm-up
        t = time.perf_counter()
        mres = mb.lanczos(lam=args.lam, seeds=mseeds, tol=1e-10, check_every=bchk, want_psi=False)
        dtm = time.perf_counter() - t
        mok = sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in mres)
        multi = {"points": len(pts_t) * G, "solves": len(mres), "converged": mok, "seconds": dtm, "rate": mok / dtm}
        log(f"multi-point batch: {len(pts_t)} d values, {len(mres)} problems in one batch: {dtm:.3f} s, "
            f"{mok}/{len(mres)} converged ({mok / dtm:.1f}/s)")
        del mb
    # pipelined leg: the timed d values again, each d's H(d,0) built on a
    # second host thread (its own library stream) while the previous d solves
    pipe = None
    if args.pipeline and len(order) > nwarm:
        pts_t = order[nwarm:]
        ex = ThreadPoolExecutor(1)

        def build_base(d):
            return build_map(lambda B: blqh_amd.Hamiltonian(B, par(d), terms=1), bases)

        t = time.perf_counter()
        fut = ex.submit(build_base, pts_t[0])
        psolved = 0
        for k, d in enumerate(pts_t):
            base = fut.result()
            if k + 1 < len(pts_t):
                fut = ex.submit(build_base, pts_t[k + 1])
            batch.rebind_scaled(base, tun)
            del base
            res = batch.lanczos(lam=args.lam, seeds=seeds, tol=1e-10, check_every=bchk, want_psi=False)
            psolved += sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in res)
        wall = time.perf_counter() - t
        ex.shutdown()
        pipe = {"points": len(pts_t) * G, "converged": psolved, "seconds": wall, "rate": psolved / wall}
        log(f"pipelined builds and solves: {len(pts_t)} d values, {psolved} solves in {wall:.3f} s "
            f"({psolved / wall:.1f}/s end to end)")
    avg_ms = kms / max(launches, 1)
    achieved = (kbytes / max(launches, 1)) / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    value = solved / t_solve if t_solve > 0 else 0.0
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "K-sector ground states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": nwarm,
        "ms_per_step": t_solve / max(1, args.steps) * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": sw["label"],
            "step": f"one d: all sectors at the {G} DSAS values, one shared-scale batched solve, H resident",
            "sectors": len(bases),
            "dsas_values": ds_list,
            "points_timed": args.steps * G,
            "solves_timed": solves,
            "converged": solved,
            "lanczos_cap": args.lam,
            "tolerance": 1e-10,
            "check_every": args.check_every,
            "batch_check_every": bchk,
            "iterations_mean": float(np.mean(iters)) if iters else None,
            "batched_rate": value,
            "plain_rate": plain["rate"] if plain else None,
            "batched_over_plain": value / plain["rate"] if plain and plain["rate"] else None,
            "plain": plain,
            "end_to_end_rate_incl_build": solved / (t_solve + t_build) if t_solve + t_build > 0 else None,
            "end_to_end_rate_pipelined": pipe["rate"] if pipe else None,
            "pipelined": pipe,
            "batched_multipoint_rate": multi["rate"] if multi else None,
            "multipoint": multi,
            "continuation_rate": cont["solved"] / cont["seconds"] if cont["seconds"] > 0 else None,
            "continuation_iterations_mean": float(np.mean(cont["iters"])) if cont["iters"] else None,
            "continuation_max_rel_dE0_vs_cold": cont["max_rel_dE0"] if cont["iters"] else None,
            "continuation_mismatches": cont["mismatch"] if cont["iters"] else None,
            "continuation_vs_plain": (cont["solved"] / cont["seconds"]) / plain["rate"]
            if cont["seconds"] > 0 and plain else None,
            "build_s_per_step": t_build / max(1, args.steps),
            "build_scheme": "H(d,0) per d; T once per sector; the batch merges H(d,0) + T once per d and "
                            "flags T's entries",
            "basis_s_once": round(t_basis, 2),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": (MEASURED_SWEEP_TRAFFIC.get((args.workload, "shared")) or (None,))[0],
            "traffic_source": (MEASURED_SWEEP_TRAFFIC.get((args.workload, "shared")) or (None, None))[1],
            "kernel": f"k_b1s<LZ=true, G={G}> (shared-scale batched fused SpMV + Lanczos K1: each sector's "
                      f"entries read once for its {G} DSAS problems)",
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "algorithmic_bytes_per_launch": kbytes / max(launches, 1),
            "algorithmic_bytes_definition": "per problem and step 20*nnz/G (the sector's entries, shared by "
                                            "its G problems) + 56*N + 8",
            "batch_nnz": nnz_total,
            "measured": "separate profiled solve of the last d (plain launches with HIP events)",
        },
    }
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on one sector)")
        try:
            out["cpu_baseline"] = sweep_cpu_baseline(sw, iters_ref=args.lam)
        except Exception as e:  # report, never hide, a failed baseline
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    return out


# Learned warm start: train on ground states at coarse d, solve at other d.
WARM = {
    "warm_nphi14": dict(S=2, nphi=14, nel=14, nu=-1, nsu=-1, dv=0.0, dz=0.0, sectors=[(0, 0), (7, 7)], dsas=0.05,
               train_d=[0.5 * i for i in range(7)],
               eval_d=[round(0.05 * i, 2) for i in range(61)],
               label="spinless bilayer Nphi=14 Ne=14 all densities, sectors (0,0) and (7,7), DSAS=0.05, "
                     "network trained at d in {0,0.5,..,3.0}, solves at d in {0,0.05,..,3.0}"),
    # the network's sign output fits real sectors (K = 0, nphi/2): complex
    # sectors keep the seeded start
    "warm_spin_nphi8": dict(S=4, nphi=8, nel=8, nu=4, nsu=4, dv=0.02, dz=0.01,
               sectors=[(J, K) for J in range(8) for K in range(8)], dsas=0.05,
               train_d=[0.25, 0.75, 1.25, 1.75, 2.25],
               eval_d=[0.5, 1.0, 1.5, 2.0],
               label="spinful bilayer Nphi=8 Ne=8 nel_up=4 nel_su=4, all 64 (J,K) sectors, DZ=0.01 DSAS=0.05 "
                     "DV=0.02, network (32 occupation bits) trained on the 16 real sectors at d in "
                     "{0.25,0.75,..,2.25}, solves at d in {0.5,1.0,1.5,2.0}; complex sectors start seeded"),
}


def run_warm(args):
    import blqh_amd
    wc = WARM[args.workload]
    S, nphi, nel, dsas = wc["S"], wc["nphi"], wc["nel"], wc["dsas"]
    log(f"workload {args.workload}: {wc['label']}")
    bases = {s: blqh_amd.Basis(nphi, nel, wc["nu"], *s, nspecies=S, nel_su=wc["nsu"]) for s in wc["sectors"]}
    log(f"{len(bases)} bases, N {min(B.n for B in bases.values())}..{max(B.n for B in bases.values())}")
    chk = args.check_every

    def real_sector(s):
        return s[1] in (0, nphi // 2)

    def ham(B, d):
        return blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d, delta_sas=dsas, delta_v=wc["dv"],
                                                       delta_z=wc["dz"]))

    # 1. training data: plain solves at the coarse grid (real sectors)
    t0 = time.time()
    sets = []
    for s, B in bases.items():
        if not real_sector(s):
            continue
        for d in wc["train_d"]:
            H = ham(B, d)
            r = H.lanczos(lam=args.lam, seed=1, tol=1e-10, check_every=chk)
            sets.append((B, d, dsas, r["psi"]))
            del H
    t_data = time.time() - t0
    log(f"training data: {len(sets)} ground states in {t_data:.1f} s")
    # 2. train on the GPU
    net = blqh_amd.Mlp(S * nphi, seed=7)
    info = net.train(sets, steps=args.train_steps, batch=8192, lr=3e-3, lr_final=1e-4, seed=11)
    log(f"trained: loss {info['loss_first']:.3f} -> {info['loss_last']:.3f} in {info['seconds']:.2f} s "
        f"({info['steps']} steps, {info['samples']} states)")
    del sets
    # 3. evaluation points: the fine grid minus the training points, evenly picked
    cand = [d for d in wc["eval_d"] if all(abs(d - t) > 1e-9 for t in wc["train_d"])]
    k = max(1, min(args.steps, len(cand)))
    pts = [cand[int(round(i * (len(cand) - 1) / max(1, k - 1)))] for i in range(k)]
    hs, keys = [], []
    for d in pts:
        for s, B in bases.items():
            hs.append(ham(B, d))
            keys.append((s, d))
    for i, H in enumerate(hs):  # autotune every matrix outside the timed regions
        H.lanczos(lam=args.lam, seed=1, tol=1e-10, check_every=chk, want_psi=False)
    # plain: seeded random start, one problem at a time
    t = time.perf_counter()
    plain = [H.lanczos(lam=args.lam, seed=1, tol=1e-10, check_every=chk, want_psi=False) for H in hs]
    t_plain = time.perf_counter() - t
    # warm: network start (inference included), one problem at a time
    t = time.perf_counter()
    warm = []
    for i, (H, (s, d)) in enumerate(zip(hs, keys)):
        if real_sector(s):
            v = net.start_vector(bases[s], d, dsas)
            warm.append(H.lanczos(lam=args.lam, start=v, tol=1e-10, check_every=chk, want_psi=False))
        else:
            warm.append(H.lanczos(lam=args.lam, seed=1, tol=1e-10, check_every=chk, want_psi=False))
    t_warm = time.perf_counter() - t
    # batched plain and batched + warm
    batch = blqh_amd.Batch(hs)
    t = time.perf_counter()
    bplain = batch.lanczos(lam=args.lam, seeds=1, tol=1e-10, check_every=chk, want_psi=False)
    t_bplain = time.perf_counter() - t
    t = time.perf_counter()
    starts = [net.start_vector(bases[s], d, dsas) if real_sector(s) else None for s, d in keys]
    bwarm = batch.lanczos(lam=args.lam, seeds=1, starts=starts, tol=1e-10, check_every=chk, want_psi=False)
    t_bwarm = time.perf_counter() - t

    def ok(rs):
        return sum(int(r["converged"] == 1 and r["dE"] < 1e-10) for r in rs)

    # energy parity with the plain solve, over pairs where both converged
    def conv(r):
        return r["converged"] == 1 and r["dE"] < 1e-10

    dev = {name: max([abs(a["E0"] - p["E0"]) / abs(p["E0"]) for a, p in zip(rs, plain) if conv(a) and conv(p)],
                     default=0.0)
           for name, rs in (("warm", warm), ("batched", bplain), ("batched_warm", bwarm))}
    unconverged = {name: sum(int(not conv(r)) for r in rs)
                   for name, rs in (("plain", plain), ("warm", warm), ("batched", bplain), ("batched_warm", bwarm))}
    dmax = max(dev.values())
    for name, rs in (("warm", warm), ("batched", bplain), ("batched_warm", bwarm)):
        for (s, d), a, p in zip(keys, rs, plain):
            if abs(a["E0"] - p["E0"]) > 1e-10 * abs(p["E0"]):
                log(f"{name} {s} d={d}: E0 {a['E0']:.15f} vs plain {p['E0']:.15f} "
                    f"(iterations {a['iterations']} vs {p['iterations']}, converged {a['converged']} vs "
                    f"{p['converged']})")
    it = {name: float(np.mean([r["iterations"] for r in rs]))
          for name, rs in (("plain", plain), ("warm", warm), ("batched", bplain), ("batched_warm", bwarm))}
    rates = {"plain": ok(plain) / t_plain, "warm": ok(warm) / t_warm, "batched": ok(bplain) / t_bplain,
             "batched_warm": ok(bwarm) / t_bwarm}
    log(f"iterations {it}; rates {rates}; max |E0 - E0_plain|/|E0| = {dmax:.1e}")
    out = {
        "metric": METRIC,
        "value": rates["batched_warm"],
        "unit": "K-sector ground states/s",
        "n_gpus": 1,
        "steps": k,
        "warmup": 1,
        "ms_per_step": t_bwarm / k * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": wc["label"],
            "step": "both sectors at one d point; batched solve of all points with network start vectors "
                    "(inference included), H resident",
            "d_points": pts,
            "solves": len(hs),
            "nHilbert": {str(s): B.n for s, B in bases.items()},
            "lanczos_cap": args.lam,
            "tolerance": 1e-10,
            "check_every": chk,
            "rates": rates,
            "batched_warm_over_plain": rates["batched_warm"] / rates["plain"],
            "warm_over_plain": rates["warm"] / rates["plain"],
            "iterations_mean": it,
            "iterations_mean_warm_sectors": {
                name: float(np.mean([r["iterations"] for r, (sct, _) in zip(rs, keys) if real_sector(sct)]))
                for name, rs in (("plain", plain), ("warm", warm), ("batched", bplain), ("batched_warm", bwarm))},
            "max_rel_E0_diff_vs_plain_both_converged": dev,
            "not_converged_at_cap": unconverged,
            "training": {"ground_states": sum(int(real_sector(x)) for x in wc["sectors"]) * len(wc["train_d"]), "data_s": round(t_data, 2),
                         "train_s": round(info["seconds"], 3), "steps": info["steps"], "batch": 8192,
                         "loss_first": info["loss_first"], "loss_last": info["loss_last"]},
        },
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="nphi16_all", choices=sorted(WORKLOADS) + sorted(SWEEPS) + sorted(WARM))
    ap.add_argument("--train-steps", type=int, default=16000,
                    help="warm workloads: Adam steps of the network (4000: 0.46 s, 16000: 1.7 s, 40000: 4.2 s)")
    ap.add_argument("--check-every", type=int, default=5)
    ap.add_argument("--lam", type=int, default=200)
    ap.add_argument("--batch-check-every", type=int, default=2,
                    help="steps between convergence checks of batched solves (0: --check-every)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--sweep-points", type=int, default=2,
                    help="default workload: also time the sweep workloads at this many parameter points (0: skip)")
    ap.add_argument("--start-mix", type=float, default=0.1,
                    help="weight of the seeded random unit vector added to warm starts (guards level crossings)")
    ap.add_argument("--events-all", action="store_true",
                    help="sweeps: time every batched K1 launch with HIP events (plain launches, no graph "
                         "replays), e.g. under a hardware-counter profiler")
    ap.add_argument("--shared", type=int, default=1,
                    help="sweeps with several DSAS values per d: one shared-scale batch per d (1) or one batch "
                         "per parameter point (0)")
    ap.add_argument("--multi-point", type=int, default=1,
                    help="sweeps: also solve the timed points' problems in one batch (1) or not (0)")
    ap.add_argument("--pipeline", type=int, default=1,
                    help="sweeps: also time the timed points with each point's builds on a second host thread, "
                         "overlapping the previous point's solve (1) or not (0)")
    ap.add_argument("--excited-mix", type=float, default=1.0,
                    help="continuation solves: weight of the previous solve's second and third Ritz vectors "
                         "added to the start (guards level crossings)")
    ap.add_argument("--continuation", type=int, default=0,
                    help="sweeps: also solve every point from the previous point's ground states (1) or not (0)")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (hardware counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration_mib:
        import blqh_amd
        blqh_amd.copy_calibration(args.copy_calibration_mib << 20)
        print(json.dumps({"copy_calibration_bytes": args.copy_calibration_mib << 20, "reps": 5}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}, WORLD_SIZE={world}); "
              "sectors and parameter points would shard across GPUs without communication, one large "
              "sector by row blocks with an all-gather of x per Lanczos step; neither is built",
              file=sys.stderr)
        sys.exit(2)

    if args.workload in SWEEPS:
        shared = args.shared and len({ds for _, ds in SWEEPS[args.workload]["points"]}) > 1
        print(json.dumps(run_sweep_shared(args) if shared else run_sweep(args)), flush=True)
        return
    if args.workload in WARM:
        run_warm(args)
        return

    import blqh_amd

    wl = WORKLOADS[args.workload]
    S, nphi, nel, nu, J, K, d, dsas, label = wl
    log(f"workload {args.workload}: {label}")
    t0 = time.time()
    B = blqh_amd.Basis(nphi, nel, nu, J, K, nspecies=S)
    t_basis = time.time() - t0
    log(f"basis: {B.n} states in {t_basis:.1f} s")
    t0 = time.time()
    H = blqh_amd.Hamiltonian(B, blqh_amd.params(nphi, d=d, delta_sas=dsas))
    t_build = time.time() - t0
    log(f"hamiltonian: nnz {H.nnz} ({H.nnz / H.n:.1f}/row) built in {t_build:.2f} s")

    for w in range(args.warmup):
        r = H.lanczos(lam=args.lam, seed=1 + w, tol=1e-10, want_psi=False)
        log(f"warmup {w}: E0/Ne={r['E0'] / nel:.12f} it={r['iterations']} {r['seconds']:.3f} s")

    blqh_amd.profile_reset()
    blqh_amd.profile_enable(True)
    solved, iters, energies = 0, [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        r = H.lanczos(lam=args.lam, seed=100 + k, tol=1e-10, want_psi=False)
        solved += int(r["converged"] == 1 and r["dE"] < 1e-10)
        iters.append(r["iterations"])
        energies.append(r["E0"])
        log(f"step {k}: E0/Ne={r['E0'] / nel:.12f} it={r['iterations']} dE={r['dE']:.1e} {r['seconds']:.3f} s")
    elapsed = time.perf_counter() - t0
    blqh_amd.profile_enable(False)
    launches, kms, kbytes = blqh_amd.profile_read()
    avg_ms = kms / max(launches, 1)
    achieved = (kbytes / max(launches, 1)) / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    spread = (max(energies) - min(energies)) / abs(np.mean(energies))
    traffic = MEASURED_TRAFFIC.get((args.workload, tuple(H.get_spmv())))

    out = {
        "metric": METRIC,
        "value": solved / elapsed,
        "unit": "K-sector ground states/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": label,
            "nHilbert": int(H.n),
            "nnz": int(H.nnz),
            "lanczos_cap": args.lam,
            "tolerance": 1e-10,
            "iterations": iters,
            "E0_per_electron": energies[-1] / nel,
            "E0_relative_spread_over_seeds": spread,
            "basis_s": round(t_basis, 2),
            "hamiltonian_build_s": round(t_build, 3),
            "spmv_variant": dict(zip(("kind", "lanes", "threads", "wgs_per_cu", "rows"), H.get_spmv())),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic[0] if traffic else None,
            "traffic_source": traffic[1] if traffic else None,
            "kernel": KERNEL_NAMES.get(H.get_spmv()[0], "?") + " (fused complex-FP64 SpMV + Lanczos K1 epilogue)",
            "avg_launch_ms": avg_ms,
            "launches": launches,
            "algorithmic_bytes_per_launch": kbytes / max(launches, 1),
        },
    }
    # end to end: the Hamiltonian build plus one solve (the per-J basis tables are one-time)
    out["config"]["end_to_end_rate_incl_build"] = 1.0 / (t_build + elapsed / args.steps)
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle CSR SpMV over the full matrix)")
        try:
            out["cpu_baseline"] = cpu_baseline_full(H, iters_ref=args.lam)
        except Exception as e:  # report, never hide, a failed baseline
            out["cpu_baseline"] = {"value": None, "error": repr(e)}
    del H, B
    blqh_amd.release_cached()
    if args.workload == "nphi16_all" and args.sweep_points > 0:
        # the headline's second half in the same run: all K-sectors of the
        # sweep workloads, batched vs plain single-problem solves
        out["sweeps"] = {}
        for name in ("sweep_spin_nphi8", "sweep_nphi12"):
            sa = argparse.Namespace(**vars(args))
            # continuation: the batch also solves every point from the previous
            # point's ground states (the warm leg of "batched + warm")
            sa.workload, sa.steps, sa.warmup, sa.continuation = name, args.sweep_points, 1, 1
            sa.no_cpu_baseline = args.no_cpu_baseline
            shared = args.shared and len({ds for _, ds in SWEEPS[name]["points"]}) > 1
            r = run_sweep_shared(sa) if shared else run_sweep(sa)
            c = r["config"]
            out["sweeps"][name] = {
                "workload": c["workload"], "sectors": c["sectors"], "points_timed": c["points_timed"],
                "solves_timed": c["solves_timed"], "converged": c["converged"],
                "batched_rate": c["batched_rate"], "plain_rate": c["plain_rate"],
                "batched_over_plain": c["batched_over_plain"], "iterations_mean": c["iterations_mean"],
                "batched_warm_rate": c["continuation_rate"], "batched_warm_over_plain": c["continuation_vs_plain"],
                "warm_iterations_mean": c["continuation_iterations_mean"],
                "warm_start": (f"previous parameter point's ground state + {sa.start_mix} x seeded unit vector + "
                               f"{sa.excited_mix} x its second and third Ritz vectors (normalised sum), kept on the device"),
                "warm_solves_above_cold_E0": c["continuation_mismatches"],
                "end_to_end_rate_incl_build": c["end_to_end_rate_incl_build"],
                "end_to_end_rate_pipelined": c["end_to_end_rate_pipelined"],
                "batched_multipoint_rate": c.get("batched_multipoint_rate"),
                "max_rel_dE0_batched_vs_plain": c["plain"]["max_rel_dE0_vs_batched"] if c["plain"] else None,
                "roofline": {k: r["roofline"][k] for k in ("achieved", "peak", "unit", "frac", "traffic", "kernel",
                                                           "avg_launch_ms", "algorithmic_bytes_per_launch")},
                "cpu_baseline": r.get("cpu_baseline"),
            }
            blqh_amd.release_cached()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
