#!/usr/bin/env python3
"""Throughput of the WaterGAP 2 day loop on one MI355X: cell-days/s.

Workload: the synthetic global 0.5-degree grid of 67,420 land cells (seed 1),
FP64 storages, 100 elevation sub-cells per cell.  One "step" is one simulated
year (365 days: calcNewDay for every cell, routing in upstream-to-downstream
order, land-area update) run as hipGraph-captured days; the default times 10
steps = 10 years.  Forcing for all timed days is resident in HBM before the
timed region starts.  The library runs the deepest routing groups as a second
partition beside the others (config.launch.partitions; WG_PARTITIONS=1 in the
environment keeps one, for comparison).

Besides the JSON line's value, the run reports
  * roofline: algorithmic bytes per launch of the dominant kernel (the one with
    the larger share of the day; with two partitions, the routing walk, which
    sets the deep partition's day) divided by its
    average duration from HIP events on the model's stream (eager launches of
    the same kernels over the whole grid, one after the other, in a profiling
    window after the timed region), against
    the 8 TB/s HBM peak, with the HBM traffic per launch when a rocprof counter
    summary is present under profiles/;
  * deep: the same measurement on a global grid of 352 routing levels (60 ocean
    outlets), where the routing chain is three times as deep;
  * water_use: the global grid with water use (subtract_use 2, use_alloc 0,
    delayed satisfaction) from seeded monthly net-use tables, with the same
    loop without water use timed beside it (--water-use runs this leg alone);
  * ensemble: 256 perturbed members of the global grid for one year;
  * calibration: basins calibrated per hour at the 1 % criterion with the
    reference's gamma search (120 synthetic basins, 30 years), and 64 candidates
    per basin in one run;
  * surrogate: basins per hour with the trained network's proposals (weights
    in watergap2_amd/data) against the plain search on 500 unseen basins;
  * cpu_baseline: the CPU oracle built with the C library's exp/pow, vertical
    balance on OpenMP threads and serial routing, as the reference runs it, on a
    bounded sample of the same workload on this host, at 1, 8 (the reference's
    setting) and all host threads.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def cell_kernel_bytes(n_cells, n_members, level0_cells):
    """Algorithmic HBM bytes of one cell-kernel launch (each array element the
    kernel must read or write once; statics once per launch, shared by members)."""
    per_member_cell = (
        16            # forcing T, P, SW, LW (float)
        + 2 * 800     # 100 elevation sub-cell snow storages, read + write
        + 2 * 8 * 5   # canopy, soil, G_snow, storage transfer, LAI precip sum (r+w)
        + 2 * 4 * 2   # LAI day counter and growing status (r+w)
        + 2 * 8 * 2   # groundwater (r+w) -- humid cells write here
        + 2 * 8 * 2   # local lake and wetland storage (r+w)
        + 2 * 8 * 2   # their reduction factors (r+w)
        + 8 * 4       # global lake, wetland, reservoir, river storage (read)
        + 2 * 8 * 4   # reduction factors global lake/wetland/reservoir/river (r+w)
        + 2 * 8 * 2   # river area fraction and its change (r+w)
        + 2 * 8 * 3   # fswb fraction, land area fraction, previous land area fraction
        + 8 * 2 + 4   # gamma, CFA, Smax
    )
    chain_record = 21 * 8  # routing record prefix written for cells with upstream inflow
    statics = (202         # elevations int16 x 101
               + 8 * 36    # FP64 statics (geometry, parameters, exp terms)
               + 4 * 3     # float statics
               + 2 * 2 + 1 * 4 + 4 * 2 + 4)  # short/char statics, CSR offsets, station slot
    n_chain = n_cells - level0_cells
    return (n_members * (n_cells * per_member_cell + n_chain * chain_record + level0_cells * 8)
            + n_cells * statics)


def chain_kernel_bytes(n_cells, n_members, level0_cells):
    """Algorithmic HBM bytes of one chain-kernel launch: per routed cell the routing
    record prefix (21 FP64 terms), the links and chain constants (80 B) and the
    river-storage and transported-volume stores; the transported volumes of
    headwater cells are read once.  The record suffix of lake/reservoir cells and
    the storage stores of lake/reservoir/wetland cells are left out (a lower
    bound)."""
    n_chain = n_cells - level0_cells
    return n_members * (n_chain * (21 * 8 + 16) + level0_cells * 8) + n_chain * 80


def ensemble_run(wg, world, forcing, members, serial_members):
    """Batched ensemble: `members` members of the grid for one year (365 days) in
    one handle, each with its own Philox forcing perturbation (precipitation x
    lognormal sigma 0.2, temperature + N(0, 1)) and gamma and Smax scaled by
    U(0.8, 1.2); against the same members run one by one on the same GPU."""
    import numpy as np
    n = world.n
    rng = np.random.default_rng(2024)
    gamma = np.tile(world.params[wg.PARAM_INDEX["P_GAMRUN_C"]], (members, 1)) * rng.uniform(0.8, 1.2, (members, n))
    smax = (np.tile(world.smax, (members, 1)) * rng.uniform(0.8, 1.2, (members, n))).astype(np.float32)
    seeds = np.arange(1000, 1000 + members, dtype=np.uint64)
    f365 = np.ascontiguousarray(forcing[:365])
    m = wg.WaterBalanceModel(world, n_members=members)
    m.set_member_params(gamma=gamma, smax=smax)
    m.set_forcing_perturbation(seeds, 0.2, 1.0)
    m.set_stations(world.basin_outlet)
    m.load_forcing(f365)
    m.step_days(0, 1, 365)  # warm-up year (captures the graph)
    m.synchronize()
    t0 = time.perf_counter()
    m.step_days(0, 1, 365)
    m.synchronize()
    batched_s = time.perf_counter() - t0
    m.close()
    serial_s = 0.0
    for k in range(serial_members):
        s1 = wg.WaterBalanceModel(world, n_members=1)
        s1.set_member_params(gamma=gamma[k:k + 1], smax=smax[k:k + 1])
        s1.set_forcing_perturbation(seeds[k:k + 1], 0.2, 1.0)
        s1.load_forcing(f365)
        s1.step_days(0, 1, 365)
        s1.synchronize()
        t0 = time.perf_counter()
        s1.step_days(0, 1, 365)
        s1.synchronize()
        serial_s += time.perf_counter() - t0
        s1.close()
    serial_per_member = serial_s / serial_members
    cd = n * 365.0
    return {"members": members, "days": 365, "cell_days_per_s": round(members * cd / batched_s, 1),
            "per_member_cell_days_per_s": round(cd / batched_s, 1), "seconds": round(batched_s, 3),
            "serial_cell_days_per_s": round(cd / serial_per_member, 1),
            "serial_sample": f"{serial_members} members run one by one, {serial_per_member:.4f}s per member-year",
            "speedup_vs_serial": round(serial_per_member * members / batched_s, 2)}


def ensemble_counter_run(args):
    """The batched ensemble (members of the global grid with their own forcing
    perturbation, gamma and Smax) launched eagerly for a few days, with per-kernel
    HIP-event durations: the workload of the ensemble's counter profiles."""
    import numpy as np
    import watergap2_amd as wg
    world = wg.World(wg.global_spec(seed=1))
    n, members, days = world.n, args.ensemble_members, args.ensemble_counters
    rng = np.random.default_rng(2024)
    gamma = np.tile(world.params[wg.PARAM_INDEX["P_GAMRUN_C"]], (members, 1)) * rng.uniform(0.8, 1.2, (members, n))
    smax = (np.tile(world.smax, (members, 1)) * rng.uniform(0.8, 1.2, (members, n))).astype(np.float32)
    m = wg.WaterBalanceModel(world, n_members=members)
    m.set_member_params(gamma=gamma, smax=smax)
    m.set_forcing_perturbation(np.arange(1000, 1000 + members, dtype=np.uint64), 0.2, 1.0)
    m.load_forcing(world.forcing(seed=1, day0=0, n_days=days))
    info = m.schedule_info()
    m.step_days(0, 1, days, profiled=True)
    ms, launches = m.timing()
    cell_ms, chain_ms = ms[0] / launches[0], ms[1] / launches[1]
    l0 = info["level0_cells"]
    out = {"members": members, "days": days}
    for name, kms, kb in (("wg_cell_kernel", cell_ms, cell_kernel_bytes(n, members, l0)),
                          ("wg_chain_kernel", chain_ms, chain_kernel_bytes(n, members, l0))):
        out[name] = {"avg_launch_us": round(kms * 1e3, 2), "algorithmic_bytes_per_launch": kb,
                     "achieved_GBs": round(kb / (kms * 1e-3) / 1e9, 1)}
    m.close()
    return out


def calibration_run(args):
    """Metric 2: basins calibrated per hour at the 1 % criterion, plain search.

    120 synthetic basins (log-uniform 50-5,000 cells), targets from a hidden
    seeded gamma per basin, 30 years; the reference's gamma search for all basins
    side by side (one model run of the whole set per search step); and the batched
    evaluation of 64 (gamma, CFA) candidates per basin in one run."""
    import numpy as np
    import watergap2_amd as wg
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search
    t0 = time.time()
    bs = make_basins(args.cal_basins, seed=2024)
    rng = np.random.default_rng(77)
    hidden = rng.uniform(0.1, 5.0, bs.n_basins)
    years = args.cal_years
    runner = BasinRunner(bs, n_members=1, forcing_years=3)
    print(f"# calibration set: {bs.n_basins} basins, {bs.n} cells (largest {np.diff(bs.offsets).max()}), "
          f"setup {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    measured = runner.run(hidden[None, :], n_years=years)[0].astype(np.float32)
    target_s = time.perf_counter() - t1
    t1 = time.perf_counter()
    res = plain_search(runner, measured, n_years=years)
    plain_s = time.perf_counter() - t1
    ok = int((res["status"] == 1).sum())
    print(f"# plain search: {res['model_runs']} model runs in {plain_s:.1f}s, {ok}/{bs.n_basins} basins "
          f"within 1 %", file=sys.stderr, flush=True)
    runner.close()
    cand = None
    if args.cal_candidates > 0:
        K = args.cal_candidates
        cr = BasinRunner(bs, n_members=K, forcing_years=3)
        g = rng.uniform(0.1, 5.0, (K, bs.n_basins))
        c = rng.uniform(0.5, 1.5, (K, bs.n_basins))
        cr.run(g, c, n_years=1)  # warm-up (graph capture)
        t1 = time.perf_counter()
        cr.run(g, c, n_years=years)
        cand_s = time.perf_counter() - t1
        cr.close()
        cand = {"candidates": K, "seconds": round(cand_s, 2),
                "cell_days_per_s": round(bs.n * K * 365.0 * years / cand_s, 1),
                "basin_candidate_runs_per_hour": round(bs.n_basins * K / cand_s * 3600, 1)}
        print(f"# {K} candidates per basin in one run: {cand_s:.1f}s, "
              f"{cand['cell_days_per_s']:.4g} cell-days/s", file=sys.stderr, flush=True)
    return {"metric": "basins calibrated/hr at the 1% criterion (plain gamma search)",
            "value": round(ok / plain_s * 3600.0, 1), "unit": "basins/hr", "higher_is_better": True,
            "config": {"workload": f"{bs.n_basins} synthetic basins, log-uniform 50-5,000 cells, "
                                   f"{bs.n} cells, {years} years, targets from hidden gamma",
                       "basins": bs.n_basins, "cells": bs.n, "years": years},
            "plain": {"seconds": round(plain_s, 2), "model_runs": res["model_runs"],
                      "runs_per_basin_mean": round(float(res["runs"].mean()), 2),
                      "within_1pc": ok, "status_counts": {int(k): int(v) for k, v in
                                                          zip(*np.unique(res["status"], return_counts=True))},
                      "target_run_seconds": round(target_s, 2)},
            "candidates": cand}


def surrogate_run(args):
    """Metric 2 with the learned part: basins calibrated per hour with surrogate
    proposals (verified by the plain model, plain search as the fallback) against
    the plain search alone, on unseen basins, at the same 1 % criterion.  The
    network comes from --sur-weights (written by watergap2_amd.train_surrogate)
    unless --sur-data asks for training here."""
    from watergap2_amd.calibration import BasinRunner, make_basins, plain_search
    from watergap2_amd import surrogate as sg

    def log(msg):
        print("# " + msg, file=sys.stderr, flush=True)

    out = {}
    if not args.sur_data and os.path.exists(args.sur_weights):
        model, winfo = sg.Surrogate.load(args.sur_weights)
        out["training"] = {"weights": os.path.relpath(args.sur_weights, ROOT)}
        out["training"].update({k: (v.item() if np.ndim(v) == 0 else v.tolist()) for k, v in winfo.items()})
        spread = float(winfo.get("residual_sd", 0.024))
    else:
        t0 = time.perf_counter()
        if args.sur_data and all(os.path.exists(p) for p in args.sur_data.split(",")):
            parts = [np.load(p) for p in args.sur_data.split(",")]
            x = np.concatenate([d["x"] for d in parts])
            t = np.concatenate([d["t"] for d in parts])
            log(f"training set loaded from {args.sur_data}: {len(t)} samples")
            gen_s = None
        else:
            x, t = sg.generate_training_set(args.sur_train_basins, args.sur_candidates, seed=args.sur_seed, log=log)
            gen_s = time.perf_counter() - t0
            if args.sur_data:
                os.makedirs(os.path.dirname(os.path.abspath(args.sur_data)), exist_ok=True)
                np.savez(args.sur_data, x=x, t=t)
        # runs with a non-finite discharge are dropped (see train_surrogate)
        finite = np.isfinite(t) & np.isfinite(x).all(axis=1)
        n_dropped = int((~finite).sum())
        x, t = x[finite], t[finite]
        out["training_set"] = {"samples": int(len(t)), "dropped_non_finite": n_dropped,
                               "basins": args.sur_train_basins, "candidates_per_basin": args.sur_candidates,
                               "generation_seconds": None if gen_s is None else round(gen_s, 1)}
        if args.sur_generate_only:
            return out
        perm = np.random.default_rng(5).permutation(len(t))
        n_val = len(t) // 20
        val, tr = perm[:n_val], perm[n_val:]
        t1 = time.perf_counter()
        model, losses = sg.Surrogate.fit(x[tr], t[tr], epochs=(args.sur_epochs, args.sur_epochs // 2), log=log)
        train_s = time.perf_counter() - t1
        res_v = model.predict(x[val]) - t[val]
        spread = float(np.std(res_v))
        out["training"] = {"seconds": round(train_s, 2), "final_loss": float(losses[-1]),
                           "validation_rmse_log_runoff": round(float(np.sqrt(np.mean(res_v ** 2))), 5)}
        log(f"trained in {train_s:.1f}s, validation RMSE of log runoff {out['training']['validation_rmse_log_runoff']}")

    # unseen basins, targets from a hidden gamma
    bs = make_basins(args.sur_eval_basins, seed=424242)
    forcing = bs.forcing(seed=11, day0=0, n_days=365 * 3)
    feats = sg.basin_features(bs, forcing)
    hidden = np.random.default_rng(99).uniform(0.1, 5.0, bs.n_basins)
    years = args.cal_years
    runner = BasinRunner(bs, n_members=1, forcing_years=3, forcing=forcing)
    measured = runner.run(hidden[None, :], n_years=years)[0].astype(np.float32)
    log(f"evaluation set: {bs.n_basins} basins, {bs.n} cells")

    t1 = time.perf_counter()
    res = plain_search(runner, measured, n_years=years)
    plain_s = time.perf_counter() - t1
    plain_ok = int((res["status"] == 1).sum())
    log(f"plain search: {res['model_runs']} runs in {plain_s:.1f}s, {plain_ok} within 1 %")

    t1 = time.perf_counter()
    rounds, round_members, bracket_rounds, bracket_members = (int(v) for v in args.sur_rounds.split(","))
    sres = sg.calibrate_with_surrogate(model, runner, feats, measured, first_members=args.sur_first_members,
                                       first_spread=spread, rounds=rounds, round_members=round_members,
                                       bracket_rounds=bracket_rounds, bracket_members=bracket_members, log=log)
    sur_s = time.perf_counter() - t1
    sur_ok = int(sres["within_1pc"].sum())
    runner.close()
    by_round = np.bincount(sres["round"], minlength=7)[1:].tolist()
    fb = by_round[5]
    plain_rate = plain_ok / plain_s * 3600.0
    sur_rate = sur_ok / sur_s * 3600.0
    out.update({
        "metric": "basins calibrated/hr at the 1% criterion, surrogate proposals vs plain search",
        "value": round(sur_rate, 1), "unit": "basins/hr", "higher_is_better": True,
        "plain_basins_per_hr": round(plain_rate, 1), "gain": round(sur_rate / max(plain_rate, 1e-9), 2),
        "gain_ceiling": {"lock_step_runs_of_plain_search": res["model_runs"],
                         "plain_runs_per_basin": round(float(res["runs"].mean()), 2)},
        "config": {"workload": f"{bs.n_basins} unseen synthetic basins (log-uniform 50-5,000 cells), {bs.n} "
                               f"cells, {years} years, targets from a hidden gamma U(0.1, 5)",
                   "basins": bs.n_basins, "cells": bs.n, "years": years,
                   "first_round_proposals_per_basin": args.sur_first_members, "later_rounds": args.sur_rounds},
        "plain": {"seconds": round(plain_s, 2), "model_runs": res["model_runs"], "within_1pc": plain_ok,
                  "runs_per_basin_mean": round(float(res["runs"].mean()), 2)},
        "surrogate": {"seconds": round(sur_s, 2), "first_round_hits": by_round[0],
                      "first_round_hit_rate": round(by_round[0] / bs.n_basins, 3),
                      "hits_per_round": by_round[:5], "fallback_basins": fb,
                      "fallback_runs": sres["fallback_runs"], "within_1pc": sur_ok,
                      "runs_per_basin_mean": round(float(sres["runs"].mean()), 2),
                      "median_abs_gamma_error": round(float(np.median(np.abs(sres["gamma"] - hidden))), 4)},
    })
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline_run(world, forcing, args):
    """The reference's arithmetic on the host (the oracle with glibc's exp/pow):
    vertical balance over cells on OpenMP threads (integrateWGHM.cpp:767), serial
    routing, station sums through an index map; timed on bounded samples at 1, 8
    and all host threads, plus the two scans of the reference as footnotes."""
    from oracle import wg_oracle
    n = world.n
    nproc = host_threads()

    def rate(threads, days, station_scan=False, aggr_scan=False):
        days = max(1, min(days, forcing.shape[0]))
        o = wg_oracle.Oracle(world, variant="glibc")
        o.set_stations(world.basin_outlet)
        o.set_threads(threads)
        o.set_reference_scans(station_scan, aggr_scan)
        t1 = time.perf_counter()
        o.run(forcing[:days], 1)
        sec = time.perf_counter() - t1
        return n * days / sec, days, sec

    by = {}
    for th, days in ((1, args.cpu_days_single), (8, args.cpu_days), (nproc, args.cpu_days)):
        v, d, sec = rate(th, days)
        by[str(th)] = {"value": round(v, 1), "days": d, "seconds": round(sec, 2)}
        print(f"# cpu oracle {v:.4g} cell-days/s on {th} threads ({d} days)", file=sys.stderr, flush=True)
    scan_v, scan_d, _ = rate(nproc, args.cpu_days_footnote, station_scan=True)
    aggr_v, aggr_d, _ = rate(nproc, args.cpu_days_footnote, station_scan=True, aggr_scan=True)
    best = by[str(nproc)]
    return {"value": best["value"], "unit": "cell-days/s", "cores": nproc, "kind": "port",
            "sample": (f"CPU oracle (glibc libm, gcc -O2 -ffp-contract=off), calcNewDay over cells on {nproc} "
                       f"OpenMP threads, serial routing; {n} cells x {best['days']} days of the same forcing in "
                       f"{best['seconds']}s"),
            "cpu_model": cpu_model(), "nproc": nproc, "by_threads": by,
            "footnote": {"reference_station_scan": {"value": round(scan_v, 1), "days": scan_d,
                                                    "note": "O(N x stations) station search of routing.cpp:4232-4238"},
                         "reference_default_options": {"value": round(aggr_v, 1), "days": aggr_d,
                                                       "note": "plus the O(N) riparian scan per lake/reservoir cell "
                                                               "of aggrNUsGloLakResOpt=1 (routing.cpp:3083-3160)"}}}


def day_loop_leg(wg, world, n_years, warmup_years, profile_days, autotune):
    """Graph-replayed years of one member after warm-up; per-kernel HIP-event
    durations over an eager window afterwards."""
    n = world.n
    dps = 365
    n_days = max((warmup_years + n_years) * dps, profile_days, 1)
    forcing = world.forcing(seed=1, day0=0, n_days=n_days)
    model = wg.WaterBalanceModel(world)
    model.set_stations(world.basin_outlet)
    model.load_forcing(forcing)
    info = model.schedule_info()
    info["partitions"] = model.partitions()
    if autotune:
        tuned, tuned_ms = model.autotune(n_days=30)
    else:
        tuned, tuned_ms = (256, 512, 5, 3), 0.0
    for s in range(warmup_years):
        model.step_days(s * dps, 1, dps)
    model.synchronize()
    t_start = time.perf_counter()
    for s in range(warmup_years, warmup_years + n_years):
        model.step_days(s * dps, 1, dps)
    model.synchronize()
    elapsed = time.perf_counter() - t_start
    pd = max(1, min(profile_days, n_days, 366))
    model.step_days(0, 1, pd, profiled=True)
    ms, launches = model.timing()
    model.close()
    return {"forcing": forcing, "info": info, "tuned": tuned, "tuned_ms": tuned_ms, "elapsed": elapsed,
            "cell_ms": ms[0] / max(launches[0], 1), "chain_ms": ms[1] / max(launches[1], 1)}


def deep_run(wg, args):
    """The day loop on a deep routing network (about 350 levels, as continental
    river systems at 0.5 degrees reach): the same 67,420-cell mask with 60 ocean
    outlets, 352 routing levels."""
    world = wg.World(wg.global_spec(seed=1, n_outlets=60))
    leg = day_loop_leg(wg, world, args.deep_years, 1, 30, autotune=not args.no_autotune)
    n = world.n
    days = 365 * args.deep_years
    l0 = leg["info"]["level0_cells"]
    chain_b = chain_kernel_bytes(n, 1, l0)
    return {"metric": "cell-days/s at global 0.5deg (FP64), deep routing network",
            "value": round(n * days / leg["elapsed"], 1), "unit": "cell-days/s",
            "config": {"workload": f"global 0.5deg synthetic land mask, {n} cells, seed 1, 60 ocean outlets, "
                                   f"{args.deep_years} years of daily steps", "cells": n, "basins": world.n_basins,
                       "routing_levels": leg["info"]["max_depth"],
                       "largest_basin_cells": int(np.bincount(world.basin).max())},
            "us_per_day": round(leg["elapsed"] / days * 1e6, 2),
            "wg_cell_kernel_us": round(leg["cell_ms"] * 1e3, 2),
            "wg_chain_kernel_us": round(leg["chain_ms"] * 1e3, 2),
            "chain_kernel_achieved_GBs": round(chain_b / (leg["chain_ms"] * 1e-3) / 1e9, 1)}


def water_use_run(wg, world, forcing, n_years=1):
    """The day loop with water use (subtract_use 2, use_alloc 0, delayed
    satisfaction) on the global grid: seeded monthly net-use tables, with return
    flows and a tenth of the cells using twenty times more, one warm-up year
    (graph capture), then n_years timed; the same loop without water use beside it."""
    n = world.n
    rng = np.random.default_rng(11)
    days = np.array([31, 28, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31], dtype=np.float64)
    scale = np.asarray(world.cell_area, dtype=np.float64) * 2e-6
    years = 1 + n_years
    nus = rng.uniform(-0.3, 1.5, (years, n, 12)) * scale[None, :, None]
    nus[:, rng.random(n) < 0.1, :] *= 20.0
    nug = rng.uniform(0.0, 0.5, (years, n, 12)) * scale[None, :, None]
    wu = rng.uniform(0.0, 2.0, (years, n, 12)) * scale[None, :, None]
    cu = wu * rng.uniform(0.2, 0.9, (years, n, 12))
    m3 = 1e9 * days[None, None, :]
    out = {}
    for label, opts in (("with_use", {"subtract_use": 2}), ("without_use", None)):
        model = wg.WaterBalanceModel(world, options=opts)
        model.set_stations(world.basin_outlet)
        if opts:
            model.set_water_use(nus * m3, nug * m3, wu * m3, cu * m3, rng.uniform(0.0, 1.0, n), delayed=1)
        model.load_forcing(forcing[: 365 * years])
        model.step_days(0, 1, 365)
        model.synchronize()
        t0 = time.perf_counter()
        for y in range(1, years):
            model.step_days(365 * y, 1, 365)
        model.synchronize()
        out[label] = time.perf_counter() - t0
        model.close()
    d = 365 * n_years
    return {"metric": "cell-days/s at global 0.5deg (FP64), water use without allocation",
            "value": round(n * d / out["with_use"], 1), "unit": "cell-days/s",
            "us_per_day": round(out["with_use"] / d * 1e6, 2),
            "us_per_day_without_use": round(out["without_use"] / d * 1e6, 2),
            "config": {"workload": f"global 0.5deg synthetic land mask, {n} cells, seed 1, {n_years} year(s) of "
                                   "daily steps, seeded monthly net-use tables, delayed satisfaction",
                       "cells": n}}


def traffic_from_profiles(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprof counter summary
    (FETCH_SIZE + WRITE_SIZE scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_global67k.json")
    if not os.path.exists(path):
        return None
    try:
        return json.load(open(path))["kernels"][kernel]["hbm_bytes_per_launch"]
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed simulated years")
    ap.add_argument("--warmup", type=int, default=1, help="untimed simulated years")
    ap.add_argument("--days-per-step", type=int, default=365)
    ap.add_argument("--profile-days", type=int, default=60)
    ap.add_argument("--cpu-days", type=int, default=150,
                    help="days of the multi-threaded CPU-oracle samples (bounded CPU baseline)")
    ap.add_argument("--cpu-days-single", type=int, default=60,
                    help="days of the single-thread CPU-oracle sample")
    ap.add_argument("--cpu-days-footnote", type=int, default=30,
                    help="days of the samples with the reference's scans")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the day loop (no deep network, ensemble, calibration or surrogate legs)")
    ap.add_argument("--deep-years", type=int, default=1, help="timed years of the deep-network leg (0 = skip)")
    ap.add_argument("--sur-weights", default=os.path.join(ROOT, "watergap2_amd", "data", "surrogate_weights.npz"),
                    help="trained surrogate (python -m watergap2_amd.train_surrogate)")
    ap.add_argument("--sur-first-members", type=int, default=1,
                    help="proposals per basin verified in the surrogate's first run")
    ap.add_argument("--sur-rounds", default="3,3,2,3",
                    help="surrogate's later rounds: correction rounds, proposals per basin in each, "
                         "bracket rounds, points per basin in each")
    ap.add_argument("--water-use", action="store_true",
                    help="only the water-use leg (global grid with seeded net-use tables)")
    ap.add_argument("--wu-years", type=int, default=1, help="timed years of the water-use leg")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch configuration (profiling runs)")
    ap.add_argument("--ensemble-members", type=int, default=256,
                    help="members of the batched-ensemble line (0 = skip)")
    ap.add_argument("--serial-members", type=int, default=4,
                    help="members timed one by one for the ensemble comparison")
    ap.add_argument("--calibration", action="store_true",
                    help="measure the calibration metric (basins/hr) instead of the day loop")
    ap.add_argument("--cal-basins", type=int, default=120, help="(0 = skip the calibration leg)")
    ap.add_argument("--cal-years", type=int, default=30)
    ap.add_argument("--cal-candidates", type=int, default=64)
    ap.add_argument("--surrogate", action="store_true",
                    help="measure basins/hr with surrogate proposals against the plain search")
    ap.add_argument("--sur-data", default="", help="training-set .npz to load, or to write when absent")
    ap.add_argument("--sur-generate-only", action="store_true")
    ap.add_argument("--sur-train-basins", type=int, default=5000)
    ap.add_argument("--sur-seed", type=int, default=9001, help="seed of the training basins")
    ap.add_argument("--sur-candidates", type=int, default=40)
    ap.add_argument("--sur-epochs", type=int, default=30)
    ap.add_argument("--sur-eval-basins", type=int, default=500, help="(0 = skip the surrogate leg)")
    ap.add_argument("--ensemble-counters", type=int, default=0,
                    help="only run this many eagerly launched days of the batched ensemble (counter runs) "
                         "and exit")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run a plain device copy of this size (counter calibration) and exit")
    args = ap.parse_args()

    if args.copy_probe_mib:
        import torch
        a = torch.empty(args.copy_probe_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": a.numel() * 4, "launches": 3}))
        return

    if args.ensemble_counters:
        print(json.dumps(ensemble_counter_run(args)))
        return
    if args.calibration:
        print(json.dumps(calibration_run(args)))
        return
    if args.surrogate:
        print(json.dumps(surrogate_run(args)))
        return
    if args.water_use:
        import watergap2_amd as wg
        w = wg.World(wg.global_spec(seed=1))
        print(json.dumps(water_use_run(wg, w, w.forcing(seed=1, day0=0, n_days=365 * (1 + args.wu_years)),
                                       args.wu_years)))
        return

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; the path would shard by basin with no "
                          "per-day exchange, but only --gpus 1 is measured"}))
        sys.exit(2)

    import watergap2_amd as wg

    t0 = time.time()
    world = wg.World(wg.global_spec(seed=1))
    n = world.n
    dps = 365
    print(f"# setup {time.time() - t0:.1f}s: {n} cells, {world.n_basins} basins", file=sys.stderr, flush=True)

    # timed region: K steps (simulated years) of graph-replayed days, bracketed by
    # device synchronisation, after W warm-up years (which also capture the graphs)
    leg = day_loop_leg(wg, world, args.steps, args.warmup, args.profile_days, autotune=not args.no_autotune)
    forcing, info, tuned = leg["forcing"], leg["info"], leg["tuned"]
    elapsed = leg["elapsed"]
    timed_days = args.steps * dps
    value = n * timed_days / elapsed if timed_days else 0.0
    ms_per_step = 1000.0 * elapsed / args.steps if args.steps else 0.0
    print(f"# autotune: cell tile {tuned[0]}, chain workgroup {tuned[1]}, snow chunk {tuned[2]}, "
          f"chain mode {tuned[3]} ({leg['tuned_ms'] * 1e3:.1f} us/day eager)", file=sys.stderr, flush=True)
    print(f"# timed {timed_days} days in {elapsed:.3f}s -> {value:.4g} cell-days/s; depth {info['max_depth']} "
          f"levels", file=sys.stderr, flush=True)

    cell_ms, chain_ms = leg["cell_ms"], leg["chain_ms"]
    l0 = info["level0_cells"]
    kern = {
        "wg_cell_kernel": (cell_ms, cell_kernel_bytes(n, 1, l0)),
        "wg_chain_kernel": (chain_ms, chain_kernel_bytes(n, 1, l0)),
    }
    per_kernel = {}
    for name, (kms, kb) in kern.items():
        gbs = kb / (kms * 1e-3) / 1e9
        per_kernel[name] = {"avg_launch_us": round(kms * 1e3, 2), "algorithmic_bytes_per_launch": kb,
                            "achieved_GBs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4),
                            "share_of_day_time": round(kms / (cell_ms + chain_ms), 3)}
        print(f"# {name} {kms * 1e3:.1f} us/launch, {kb / 1e6:.1f} MB algorithmic -> {gbs:.0f} GB/s",
              file=sys.stderr, flush=True)
    day_bytes = sum(kb for _, kb in kern.values())
    # The kernel that sets the day: the routing walk.  With two partitions it is
    # most of the deep partition's day (the critical path of the graph-replayed
    # days); with one it has the larger share.  The eager window below runs the
    # whole grid's partition-major tables, where the cell kernel is slower than in
    # either partition, so its share there does not decide.  Both kernels' figures
    # are under per_kernel.
    dominant = "wg_chain_kernel" if info.get("partitions") and len(info["partitions"]) > 1 \
        else max(kern, key=lambda k: kern[k][0])
    dom_ms, dom_bytes = kern[dominant]
    dom_gbs = dom_bytes / (dom_ms * 1e-3) / 1e9
    roofline = {
        "kernel": dominant,
        "bound": "hbm",
        "achieved": round(dom_gbs, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
        "traffic": traffic_from_profiles(dominant),
        "per_kernel": per_kernel,
        # the whole simulated day against the algorithmic bytes of both kernels
        "day": {"bytes_per_cell_day": round(day_bytes / n, 1), "us_per_day": round((cell_ms + chain_ms) * 1e3, 2),
                "achieved_GBs": round(day_bytes / ((cell_ms + chain_ms) * 1e-3) / 1e9, 1),
                "frac": round(day_bytes / ((cell_ms + chain_ms) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                "graph_us_per_day": round(elapsed / max(timed_days, 1) * 1e6, 2)},
        "note": (f"wg_chain_kernel (the wave walk) routes each routing group level by level in "
                 f"batches of 64 cells, {info['max_depth']} dependent levels at this grid; its time is "
                 f"latency, not bandwidth"),
    }

    cpu = None if args.no_cpu else cpu_baseline_run(world, forcing, args)

    deep = None
    if not args.no_extras and args.deep_years > 0:
        deep = deep_run(wg, args)
        print(f"# deep network: {deep['value']:.4g} cell-days/s, {deep['config']['routing_levels']} levels, "
              f"chain {deep['wg_chain_kernel_us']} us/day", file=sys.stderr, flush=True)

    use = None
    if not args.no_extras and args.wu_years > 0 and forcing.shape[0] >= 365 * (1 + args.wu_years):
        use = water_use_run(wg, world, forcing, args.wu_years)
        print(f"# water use: {use['value']:.4g} cell-days/s, {use['us_per_day']} us/day "
              f"({use['us_per_day_without_use']} without)", file=sys.stderr, flush=True)

    ens = None
    if not args.no_extras and args.ensemble_members > 0:
        ens = ensemble_run(wg, world, forcing if forcing.shape[0] >= 365 else world.forcing(seed=1, day0=0, n_days=365),
                           args.ensemble_members, max(1, args.serial_members))
        print(f"# ensemble {ens['members']} members x 1 year: {ens['cell_days_per_s']:.4g} cell-days/s, "
              f"{ens['speedup_vs_serial']}x the serial member runs", file=sys.stderr, flush=True)

    cal = None
    if not args.no_extras and args.cal_basins > 0:
        cal = calibration_run(args)
        print(f"# calibration: {cal['value']:.4g} basins/hr (plain search)", file=sys.stderr, flush=True)

    sur = None
    if not args.no_extras and args.sur_eval_basins > 0 and os.path.exists(args.sur_weights):
        sur = surrogate_run(args)
        print(f"# surrogate: {sur['value']:.4g} basins/hr, {sur['gain']}x the plain search", file=sys.stderr,
              flush=True)

    line = {
        "metric": "cell-days/s at global 0.5deg (FP64)",
        "value": round(value, 1),
        "unit": "cell-days/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": (f"global 0.5deg synthetic land mask, {n} cells, seed 1, 100 elevation "
                                f"sub-cells, {args.steps} years of daily steps, hipGraph-captured days"),
                   "cells": n, "days_per_step": dps, "members": 1, "basins": world.n_basins,
                   "routing_levels": info["max_depth"],
                   "launch": {"cell_tile": tuned[0], "chain_workgroup": tuned[1], "snow_chunk": tuned[2],
                              "chain_kernel": ["level walk, 1 level ahead", "dataflow walk",
                                               "level walk, 2 levels ahead",
                                               "wave walk in 64-cell batches"][tuned[3]],
                              "partitions": [{"cells": c, "deepest_group_levels": d}
                                             for c, d in info["partitions"]]}},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "deep": deep,
        "water_use": use,
        "ensemble": ens,
        "calibration": cal,
        "surrogate": sur,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
