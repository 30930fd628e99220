#!/usr/bin/env python3
"""Benchmark of the short-range force path on the 64k-atom LJ liquid-vapour slab.

One step = thermal vibration of every atom about its site in the initial
configuration (x = x0 + 0.01 sigma N(0,1) per component, seeded, drawn afresh
each step, so no two atoms ever come closer than the set-up allows) ->
position refresh (the perturbation is applied inside it) + per-wavefront
skin test -> LJ force, energy and virial with the per-atom cutoff rc(z) and
the per-atom tail correction.  The neighbor list is rebuilt every
--rebuild-every steps (10, the rebuild period of the large-slab workload);
the displacement recorded on the device is checked at every rebuild, so a run
in which any atom moved more than skin/2 between rebuilds fails instead of
reporting.  The energy and virial are reduced and read back every 100 steps,
like the reference driver's thermo output.  The segments between two
read-backs, each of its steps and its closing rebuild, are replayed as
hipGraphs of several segments each.

Prints one JSON line:  value = pair evaluations per second (list entries
processed by the force kernel per second of whole-job wall time, rebuilds
included), ms_per_step, the FP64 form measured the same way, the force
kernel's HBM roofline fraction, force/energy error of the mixed path against
the FP64 form, the CPU path (the FP64 oracle, OpenMP) timed on the host, and
under "md" the device MD loop (velocity Verlet + Nose-Hoover chain, the same
force path) timed on the same slab, and under "arc" the adaptive-cutoff
producer (density grid -> error estimate -> rc lattice) on the slab's box.
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "md.error_amd"))
sys.path.insert(0, REPO)

METRIC = "pair evals/s + ms/step, 64k LJ slab: mixed-prec vs FP64, max force err<=1e-5"
HBM_PEAK_GBS = 8000.0
PERTURB_SIGMA = 0.01
SEED = 7
THERMO_EVERY = 100   # energy/virial read back like the reference's thermo output
                     # (H/app/lj.nhc-xtc.cu:198-224)
WARM_SECONDS = 0.25  # least wall time of the untimed warm-up of each form


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def run_form(w, precision, steps, warmup, every, torch, Engine, autotune=False):
    """Times `steps` perturb+update+force steps; returns a dict of figures."""
    stream = torch.cuda.Stream()          # default priority
    e = Engine(n_types=w.n_types, skin=w.skin, precision=precision,
               stream=stream.cuda_stream)
    w.setup(e, tail=True)
    pos = torch.from_numpy(w.xyz.copy()).to("cuda")
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.rebuild()
    tune_s = None
    if autotune:
        t0 = time.perf_counter()
        e.autotune(stats=True, tail=True)
        tune_s = time.perf_counter() - t0
    e.set_perturbation(PERTURB_SIGMA, SEED, 0, anchored=True)

    def check_skin(disp):
        # every atom must have stayed within skin/2 of its position at the
        # rebuild before its segment
        if disp > 0.5 * w.skin:
            raise RuntimeError(f"skin exceeded between rebuilds ({disp:.4f})")

    state = {"step": 0}     # steps run since the first rebuild (warmup included)

    def segment_loop(n_steps):
        """Runs n_steps steps in segments of `every` steps, each ended by a
        rebuild.  The energy and virial are reduced and read back at every
        multiple of THERMO_EVERY steps (counted from the start), like the
        reference driver's thermo output; between two read-backs the segments
        and their rebuilds are queued with one synchronisation
        (sre_run_segments, which checks every rebuild's list capacity and
        displacement at the end).  A block at a read-back ends with the
        read-back of its last step's energy and virial, after the queued
        segments (their closing rebuild leaves those values), or, when the
        block does not end on a segment, after its partial segment and
        before a rebuild; a run that ends between two read-backs ends with
        the queued segments' own check (or the skin-flag check of a trailing
        partial segment)."""
        done = 0
        pairs = 0
        rebuilds = 0
        nl, _, _ = e.neighbor_count(in_rc=False)   # known since the last rebuild
        while done < n_steps:
            to_thermo = THERMO_EVERY - state["step"] % THERMO_EVERY
            block = min(to_thermo, n_steps - done)
            thermo = block == to_thermo
            nfull, rem = divmod(block, every)
            if nfull > 0:
                nls, disp = e.run_segments(nfull, every, stats=True, tail=True,
                                           perturb=True)
                check_skin(disp)
                for k in range(nfull):
                    pairs += nl * every
                    nl = int(nls[k])
                rebuilds += nfull
                if thermo and rem == 0:
                    # the energy and virial of the block's last step (the
                    # segment's closing rebuild leaves them); raises on a
                    # force overflow
                    e.stats()
            if rem > 0:
                e.graph_run(rem, stats=True, tail=True, perturb=True)
                pairs += nl * rem
                if thermo:
                    e.stats()             # reduce + read back; raises on a force overflow
                    e.rebuild()
                    rebuilds += 1
                    check_skin(e.last_rebuild_displacement())
                    nl, _, _ = e.neighbor_count(in_rc=False)
                else:
                    e.check_skin_flag()   # synchronises; raises past skin/2
            done += block
            state["step"] += block
        return pairs, rebuilds

    # untimed warm-up: at least one whole read-back block, rounded up to
    # whole blocks, so that every graph the timed steps replay is captured,
    # every buffer is sized and the timed steps start just after a read-back;
    # then further whole blocks until the GPU has worked for WARM_SECONDS, so
    # that a short timed run does not start on clocks still ramping up from
    # the idle host-side set-up (measured: the first 20-step window after
    # the set-up runs 5-10 % slower than later ones)
    t_warm = time.perf_counter()
    warm = -(-max(warmup, THERMO_EVERY) // THERMO_EVERY) * THERMO_EVERY
    segment_loop(warm)
    torch.cuda.synchronize()
    while time.perf_counter() - t_warm < WARM_SECONDS:
        segment_loop(THERMO_EVERY)
        warm += THERMO_EVERY
        torch.cuda.synchronize()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    ev0.record(stream)
    pairs, rebuilds = segment_loop(steps)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    # force kernel alone, back-to-back launches on the same stream (HIP events)
    e.time_compute(20, stats=True, tail=True)      # warm the clocks and caches
    k_ms = e.time_compute(200, stats=True, tail=True)
    k_piped_ms = None
    if precision == "fp64":
        # the FP64 form in the mixed kernel's pipelined structure, same sums
        plain = e.tuning()
        e.set_tuning(force_unroll=1)
        e.time_compute(20, stats=True, tail=True)
        k_piped_ms = e.time_compute(200, stats=True, tail=True)
        e.set_tuning(**plain)
    k_cells_ms = None
    if precision == "mixed":
        # the cell-window form without a list (LDS-staged candidates, bitwise
        # the same forces): needs a rebuild for its tile map, then one back
        plain = e.tuning()
        e.set_tuning(force_mode=1)
        e.rebuild()
        e.time_compute(5, stats=True, tail=True)
        k_cells_ms = e.time_compute(20, stats=True, tail=True)
        e.set_tuning(**plain)
        e.rebuild()
    dist = force_time_distribution(e, torch, stream)
    rb_ms = e.time_rebuild(10)
    e.compute(stats=True, tail=True)
    e.stats()    # raises if any pair force since the warm-up left the fixed-point range
                 # (the flag is kept until it is read)
    nl, nin, mx = e.neighbor_count()
    return dict(engine=e, pos=pos, wall_s=wall, gpu_ms=gpu_ms, pairs=pairs, warm=warm,
                rebuilds=rebuilds, force_ms=k_ms, n_list=nl, n_in_rc=nin, max_list=mx,
                force_us_dist=dist, rebuild_ms=rb_ms, autotune_s=tune_s,
                force_piped_ms=k_piped_ms, force_cells_ms=k_cells_ms,
                pair_evals_per_s=pairs / wall, ms_per_step=1e3 * wall / steps)


def _errors(F, S, F_ref, S_ref):
    import numpy as np
    rms = np.sqrt((F_ref ** 2).sum(1).mean())
    d = np.sqrt(((F - F_ref) ** 2).sum(1))
    floor = np.maximum(np.sqrt((F_ref ** 2).sum(1)), rms)
    return {"max_rel_force_err": float(d.max() / rms),
            "rms_rel_force_err": float(np.sqrt((d ** 2).mean()) / rms),
            "max_per_atom_rel_err_floored": float((d / floor).max()),
            "energy_rel_err": float(abs(S[0] - S_ref[0]) / abs(S_ref[0])),
            "virial_rel_err": float(np.max(np.abs(S[1:] - S_ref[1:]) / np.abs(S_ref[1:])))}


def accuracy(w, torch, Engine):
    """Mixed path and FP64 GPU form at the initial positions against the FP64
    oracle (C++ on the host: same rc assignment, list and tail), and the mixed
    path against the FP64 GPU form.  Errors: max_i |F_i - F_i^ref| / RMS_i
    |F_i^ref|, the RMS of the same, per atom with the floor max(|F_i^ref|, RMS),
    and the relative energy and virial errors.  Runs outside every timed
    region."""
    import numpy as np
    from oracle import oracle as O
    res = {}
    for prec in ("mixed", "fp64"):
        e = Engine(n_types=w.n_types, skin=w.skin, precision=prec)
        w.setup(e, tail=True)
        e.rebuild()
        e.compute(stats=True, tail=True)
        res[prec] = (e.forces(), e.stats())
        e.close()
    ref_F, ref_S = [], []
    for r in range(len(w.n_atoms)):
        sl = w.replica_slice(r)
        x, box, types = w.xyz[sl], w.boxes[r], w.types[sl]
        rc, _ = O.assign_rc(x, box, w.rc_tables[r])
        off, j = O.build_pairs(x, box, rc, w.skin)
        tail = None
        if w.rho and w.rho[r] is not None:
            tail = O.tail_atoms(x, box, types, w.eps, w.sigma, w.rc_tables[r].ravel(),
                                w.rho[r])
        F, E, W, _ = O.forces(x, box, types, w.eps, w.sigma, rc, off, j, tail=tail)
        ref_F.append(F)
        ref_S.append(np.concatenate([[E.sum()], W.sum(0)]))
    F64 = np.concatenate(ref_F)
    out = {}
    for prec in ("mixed", "fp64"):
        F, S = res[prec]
        per = [_errors(F[w.replica_slice(r)], S[r], ref_F[r], ref_S[r])
               for r in range(len(w.n_atoms))]
        whole = _errors(F, S[0], F64, ref_S[0])
        # max over replicas for the energy/virial figures
        for k in ("energy_rel_err", "virial_rel_err"):
            whole[k] = max(p[k] for p in per)
        out[prec] = whole
    mixed_vs_fp64 = _errors(res["mixed"][0], res["mixed"][1][0], res["fp64"][0], res["fp64"][1][0])
    return out, mixed_vs_fp64


def measured_traffic(workload):
    """HBM bytes per launch of the mixed force kernel from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summary (counter values scaled by a copy
    kernel of known size), or None when there is none for this workload."""
    path = os.path.join(REPO, "profiles", f"k_force_mixed_{workload}_hbm.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    return None


def _cpu_rate(w, budget_s, max_evals=2000):
    """FP64 oracle on the host: one list build, then force evaluations for
    about budget_s.  Returns (pair evals/s, ms per force, ms per build, L, n)."""
    from oracle import oracle as O
    box = w.boxes[0]
    rc, _ = O.assign_rc(w.xyz, box, w.rc_tables[0])
    tail = None
    if w.rho and w.rho[0] is not None:
        tail = O.tail_atoms(w.xyz, box, w.types, w.eps, w.sigma, w.rc_tables[0].ravel(),
                            w.rho[0])
    t0 = time.perf_counter()
    off, j = O.build_pairs(w.xyz, box, rc, w.skin)
    t_build = time.perf_counter() - t0
    n_eval = 0
    t0 = time.perf_counter()
    while True:
        O.forces(w.xyz, box, w.types, w.eps, w.sigma, rc, off, j, tail=tail)
        n_eval += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n_eval >= max_evals:
            break
    t_force = el / n_eval
    L = int(off[-1])
    return L / t_force, t_force * 1e3, t_build * 1e3, L, n_eval


def time_adjoint(e, w, torch, reps=5):
    """The adjoint of E and of a force-matching loss (dLdE = 1 per replica,
    dLdF = seeded Gaussian device tensor) w.r.t. eps, sigma and the rc
    lattice: median wall ms per call (includes the host readback of the
    gradients, which the call returns)."""
    import numpy as np
    g = torch.Generator(device="cpu").manual_seed(SEED)
    dLdF = torch.randn(w.n, 3, generator=g, dtype=torch.float64).to("cuda")
    n_rc = int(sum(t.size for t in w.rc_tables))
    torch.cuda.synchronize()
    e.adjoint(dLdE=1.0, dLdF=dLdF, tail=True, n_rc=n_rc)      # warm-up
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        e.adjoint(dLdE=1.0, dLdF=dLdF, tail=True, n_rc=n_rc)
        t.append(time.perf_counter() - t0)
    return {"ms": round(1e3 * float(np.median(t)), 3), "n_rc_bins": n_rc,
            "what": "gradients of E and of sum_i dLdF_i . F_i w.r.t. eps, sigma "
                    "(n_types^2 each) and every rc bin, pair terms recomputed"}


def cpu_baseline(w, budget_s, every, configs):
    """The FP64 oracle (C++/OpenMP) on the host: all threads and one thread,
    on the benchmark workload and on the 4k FCC plain-method case."""
    from oracle import oracle as O
    threads = O.num_threads()
    rate, ms_f, ms_b, L, n_eval = _cpu_rate(w, budget_s)
    out = {"value": rate, "unit": "pair evals/s", "cores": threads, "kind": "port",
           "sample": f"{n_eval} FP64 force evaluations of the {w.name} workload ({L} list "
                     f"entries, {ms_f:.2f} ms each) + 1 list build ({ms_b:.1f} ms); "
                     f"ms/step at the same rebuild period {ms_f + ms_b / every:.2f}",
           "ms_force": ms_f, "ms_build": ms_b,
           "ms_per_step": ms_f + ms_b / every}
    try:
        O.set_threads(1)
        r1, f1, b1, _, n1 = _cpu_rate(w, budget_s / 2, max_evals=50)
        out["single_thread"] = {"value": r1, "ms_force": f1, "ms_build": b1,
                                "ms_per_step": f1 + b1 / every, "evals": n1}
        fcc = configs.plain_fcc()
        for t, key in ((threads, "fcc_4k_all_threads"), (1, "fcc_4k_single_thread")):
            O.set_threads(t)
            rf, ff, bf, Lf, nf = _cpu_rate(fcc, 1.0, max_evals=500)
            out[key] = {"value": rf, "ms_force": ff, "ms_build": bf, "list_entries": Lf,
                        "ms_per_step": ff + bf / every, "threads": t}
    finally:
        O.set_threads(threads)
    return out


def sweep_launch(e, n=200):
    """Back-to-back force kernel time (us, HIP events) of the flat launch (one
    wavefront per tile) and of persistent launches of k wavefronts per SIMD
    (each taking a contiguous chunk of tiles); the fixed-point forces must
    not change."""
    import numpy as np
    base = e.tuning()
    out = {}
    f0 = None
    for k in (0, 1, 2, 4, 8):
        e.set_tuning(persistent=k)
        e.time_compute(20, stats=True, tail=True)
        out[str(k)] = round(e.time_compute(n, stats=True, tail=True) * 1e3, 3)
        e.compute(stats=True, tail=True)
        f = e.forces_fixed()
        if f0 is None:
            f0 = f
        elif not np.array_equal(f, f0):
            raise RuntimeError(f"persistent launch {k} changed the forces")
    e.set_tuning(**base)
    return {"force_kernel_us_by_persistent_waves_per_simd": out}


def force_time_distribution(e, torch, stream, n=200):
    """Per-launch force kernel durations (HIP events around each launch on the
    engine stream): median, p10, p90 in microseconds."""
    import numpy as np
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(n)]
    for a, b in evs:
        a.record(stream)
        e.compute(stats=True, tail=True)
        b.record(stream)
    torch.cuda.synchronize()
    t = np.array([a.elapsed_time(b) * 1e3 for a, b in evs])
    return {"p10": round(float(np.percentile(t, 10)), 3),
            "median": round(float(np.median(t)), 3),
            "p90": round(float(np.percentile(t, 90)), 3)}


MD_DT, MD_REF_T, MD_TAU_T = 0.005, 1.10, 1.0   # H/app/lj.nhc-xtc.cu:42-46


def run_md(w, steps, warmup, torch, Engine):
    """Times the device MD loop: velocity Verlet + two-link Nose-Hoover chain
    + COM removal, the reference driver's step (H/app/lj.nhc-xtc.cu:158-196,
    dt 0.005, T 1.1, tau 1), with the force path of the metric (rc(z) + tail)
    and adaptive rebuild stretches; thermo output (energy, virial, kinetic
    energy and chain state read back) every 100 steps.  The generated slab
    holds pairs down to dmin = 0.85, whose release would heat it far above
    1.1, so the warmup first relaxes it under a tight chain (tau 0.05), then
    draws fresh velocities at 1.1 and restarts the chain with tau 1."""
    import numpy as np
    stream = torch.cuda.Stream()
    e = Engine(n_types=w.n_types, skin=w.skin, precision="mixed", stream=stream.cuda_stream)
    w.setup(e, tail=True)
    pos = torch.from_numpy(w.xyz.copy()).to("cuda")
    torch.cuda.synchronize()
    e.bind_positions(pos)
    e.set_masses(np.ones(w.n_types, np.float32))
    rng = np.random.default_rng(SEED)

    def draw():
        e.set_velocities(rng.normal(0.0, np.sqrt(MD_REF_T), (w.n, 3)).astype(np.float32))

    def loop(n_steps):
        done, pairs, rebuilds = 0, 0, 0
        while done < n_steps:
            n = min(THERMO_EVERY, n_steps - done)
            rebuilds += e.md_run(n, rebuild_every="adaptive", stats=True, tail=True)
            e.stats()
            e.md_state()
            pairs += e.neighbor_count(in_rc=False)[0] * n
            done += n
        return pairs, rebuilds

    draw()
    e.set_integrator(MD_DT, "nhc2", MD_REF_T, 0.05, remove_com=True)
    loop(warmup)
    draw()
    e.set_integrator(MD_DT, "nhc2", MD_REF_T, MD_TAU_T, remove_com=True)
    loop(100)
    r0 = e.md_state()["rollbacks"]
    t_start = 2.0 * e.md_state()["kinetic"] / (3.0 * w.n)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    pairs, rebuilds = loop(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = e.md_state()
    return {"ms_per_step": round(1e3 * wall / steps, 5),
            "steps_per_s": round(steps / wall, 1),
            "pair_evals_per_s": round(pairs / wall, 1),
            "steps": steps, "warmup": warmup, "rebuilds": rebuilds,
            "mean_steps_per_rebuild": round(steps / max(rebuilds, 1), 2),
            "rollbacks": st["rollbacks"] - r0,
            "temperature_start": round(t_start, 4),
            "temperature_end": round(2.0 * st["kinetic"] / (3.0 * w.n), 4),
            "mean_list_entries": round(pairs / steps),
            "dt": MD_DT, "thermostat": f"NHC2 T={MD_REF_T} tau={MD_TAU_T}",
            "pairs_def": "list entries after each 100-step call x 100",
            "rebuild_policy": "adaptive stretches (snapshot + rollback), skin/2 rule"}


def time_arc(w, torch, cpu=True):
    """Times the adaptive-cutoff producer on the workload's box (grid of ~1 sigma
    cells, rc candidates rc_min..rc_max step 0.25): the kernel tables (once per
    grid), then one update = deposit + error estimate of every candidate +
    choice + two widenings + read-back of the lattice (the reference's
    every-rcutUpdateFeq work).  cpu: numpy's FFT route with the same kernel
    tables (the oracle, one thread) on the same grid."""
    import numpy as np
    from sre import Arc
    box = [float(v) for v in w.boxes[0]]
    rt = w.rc_tables[0]
    rcmin, rcmax = float(rt.min()), float(rt.max())
    arc = Arc(box, 1.0)
    t0 = time.perf_counter()
    rcs = arc.set_candidates(rcmin, rcmax, 0.25)
    t_tab = time.perf_counter() - t0
    pos = torch.from_numpy(w.xyz.copy()).to("cuda")
    torch.cuda.synchronize()

    def update():
        arc.clear()
        arc.deposit(pos)
        arc.cal_error()
        arc.cal_rcut(0.0045)
        arc.refine(2)
        return arc.result()

    update()
    ts = []
    for _ in range(5):
        t0 = time.perf_counter()
        update()
        ts.append(time.perf_counter() - t0)
    out = {"grid": list(arc.dims), "cells": arc.nele, "candidates": len(rcs),
           "kernel_tables_ms": round(1e3 * t_tab, 2),
           "update_ms": round(1e3 * float(np.median(ts)), 3)}
    if cpu:
        from oracle import arc as A
        S1, G = arc.kernels()
        rho = arc.density()
        kv = A.k_grid(box, arc.dims)
        t0 = time.perf_counter()
        A.cal_error(rho, box, S1, G, kv)
        out["cpu_numpy_fft_error_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
    arc.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--rebuild-every", type=int, default=10,
                    help="steps per neighbor-list rebuild (the large-slab workload's period)")
    ap.add_argument("--cpu-seconds", type=float, default=5.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mixed-only", action="store_true")
    ap.add_argument("--workload", default="slab_64k",
                    help="slab_64k (the metric's workload), binary_slab_64k, large_slab_2M, "
                         "replicas_256, plain_fcc_4k")
    ap.add_argument("--autotune", action="store_true",
                    help="pick tile/block/cell sizes with sre_autotune before timing")
    ap.add_argument("--adjoint", action="store_true",
                    help="also time the adjoint (parameter gradients) of the mixed form")
    ap.add_argument("--no-arc", action="store_true",
                    help="skip timing the adaptive-cutoff producer")
    ap.add_argument("--no-md", action="store_true",
                    help="skip timing the device MD loop (VV + NHC2 + COM removal)")
    ap.add_argument("--sweep-launch", action="store_true",
                    help="also time the mixed force kernel as a persistent launch of "
                         "1, 2, 4, 8 wavefronts per SIMD against the flat launch")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run device copies of MIB MiB (HBM counter calibration)")
    ap.add_argument("--force-kernel-only", action="store_true",
                    help="only time the mixed force kernel back to back (the launches "
                         "the roofline figure averages), for a profiler run")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the "
              "multi-GPU layout of this path is described, not run.", file=sys.stderr)
        sys.exit(2)

    import numpy as np  # noqa: F401
    import torch
    from sre import Engine, configs

    if not torch.cuda.is_available():
        log("bench.py needs a GPU")
        sys.exit(1)
    if args.calibrate_copy:
        # known traffic: each copy reads and writes MIB MiB once
        n = args.calibrate_copy * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 5 copies of {args.calibrate_copy} MiB")
        return
    w = configs.WORKLOADS[args.workload]()
    log(f"[bench] workload {w.name}: N={w.n}, {w.notes}")
    if args.force_kernel_only:
        e = Engine(n_types=w.n_types, skin=w.skin)
        w.setup(e, tail=True)
        e.rebuild()
        e.time_compute(20, stats=True, tail=True)
        k_ms = e.time_compute(200, stats=True, tail=True)
        nl, _, _ = e.neighbor_count(in_rc=False)
        b = 4.0 * nl + 84.0 * w.n
        print(json.dumps({"workload": w.name, "force_kernel_us": round(k_ms * 1e3, 3),
                          "launches": 220, "bytes_per_launch": b,
                          "achieved_GBs": round(b / (k_ms * 1e-3) / 1e9, 1)}))
        e.close()
        return
    mixed = run_form(w, "mixed", args.steps, args.warmup, args.rebuild_every, torch, Engine,
                     autotune=args.autotune)
    log(f"[bench] mixed: {mixed['pair_evals_per_s']:.4g} pair evals/s, "
        f"{mixed['ms_per_step']:.4f} ms/step, force kernel {mixed['force_ms'] * 1e3:.2f} us")
    fp64 = None
    if not args.mixed_only:
        fp64 = run_form(w, "fp64", args.steps, args.warmup, args.rebuild_every, torch, Engine,
                        autotune=args.autotune)
        log(f"[bench] fp64:  {fp64['pair_evals_per_s']:.4g} pair evals/s, "
            f"{fp64['ms_per_step']:.4f} ms/step, force kernel {fp64['force_ms'] * 1e3:.2f} us")
    acc_oracle, acc = accuracy(w, torch, Engine)
    acc["reference"] = "FP64 GPU form"
    acc_oracle["reference"] = ("FP64 oracle (C++ on the host) at the initial positions; "
                               "mixed and fp64 are the two GPU forms")
    log(f"[bench] accuracy vs oracle: {acc_oracle}")

    L, N = mixed["n_list"], w.n
    b_force = 4.0 * L + 84.0 * N
    achieved = b_force / (mixed["force_ms"] * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(w.name),
                "kernel": "k_force_mixed<T,stats,tail>",
                "bytes_per_launch": b_force,
                "bytes_def": "4*L + 84*N (int32 list entry per pair; per atom: pos 16 + rc 4 "
                             "+ count 4 + tail index 4 read, 3x int64 force 24 + FP64 "
                             "energy/virial 32 written)",
                "avg_launch_us": round(mixed["force_ms"] * 1e3, 3)}
    cpu = None
    if not args.no_cpu and len(w.n_atoms) == 1:
        cpu = cpu_baseline(w, args.cpu_seconds, args.rebuild_every, configs)
        log(f"[bench] cpu baseline: {cpu['value']:.4g} pair evals/s on {cpu['cores']} threads")
    result = {
        "metric": METRIC,
        "value": round(mixed["pair_evals_per_s"], 1),
        "unit": "pair evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "warmup_steps_run": mixed["warm"],
        "ms_per_step": round(mixed["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair math, i64 fixed-point forces, f64 energy/virial",
        "data": "synthetic (seeded random-sequential-addition slab)",
        "config": {"workload": w.name, "n_atoms": N,
                   "box": [float(v) for v in w.boxes[0]], "n_types": w.n_types,
                   "n_replicas": len(w.n_atoms),
                   "rho_liquid_vapour": [0.73, 0.02], "geometry": w.notes,
                   "skin": w.skin, "rebuild_every": args.rebuild_every,
                   "motion": "vibration about the initial sites", "perturbation_sigma": PERTURB_SIGMA,
                   "tail": "per-atom, bin mode",
                   "list_entries": L, "pairs_in_rc": mixed["n_in_rc"],
                   "tuning": mixed["engine"].tuning()},
        "force_kernel_us": round(mixed["force_ms"] * 1e3, 3),
        "force_kernel_us_per_launch": mixed["force_us_dist"],
        # the same forces (bitwise) from the cell-window kernel without a list
        "force_kernel_cells_us": round(mixed["force_cells_ms"] * 1e3, 3),
        "rebuild_us": round(mixed["rebuild_ms"] * 1e3, 1),
        "rebuild_us_per_step": round(mixed["rebuild_ms"] * 1e3 / args.rebuild_every, 2),
        "force_pair_evals_per_s": round(L / (mixed["force_ms"] * 1e-3), 1),
        "roofline": roofline,
        "accuracy_vs_oracle": acc_oracle,
        "accuracy_vs_fp64": acc,
        "cpu_baseline": cpu,
    }
    if args.autotune:
        result["autotune_s"] = round(mixed["autotune_s"], 2)
    if args.adjoint:
        result["adjoint"] = time_adjoint(mixed["engine"], w, torch)
    if args.sweep_launch:
        result["launch_sweep"] = sweep_launch(mixed["engine"])
    if not args.no_arc and len(w.n_atoms) == 1:
        try:
            result["arc"] = time_arc(w, torch, cpu=not args.no_cpu)
        except Exception as ex:          # reported in the line, never silent
            result["arc"] = {"error": f"{type(ex).__name__}: {ex}"}
        log(f"[bench] arc: {result['arc']}")
    if not args.no_md and len(w.n_atoms) == 1:
        try:
            result["md"] = run_md(w, args.steps, max(args.warmup, 300), torch, Engine)
        except Exception as ex:          # reported in the line, never silent
            result["md"] = {"error": f"{type(ex).__name__}: {ex}"}
        log(f"[bench] md: {result['md']}")
    if fp64 is not None:
        result["fp64_form"] = {
            "value": round(fp64["pair_evals_per_s"], 1), "ms_per_step": round(fp64["ms_per_step"], 5),
            "force_kernel_us": round(fp64["force_ms"] * 1e3, 3),
            "force_kernel_us_per_launch": fp64["force_us_dist"],
            "rebuild_us": round(fp64["rebuild_ms"] * 1e3, 1),
            "force_kernel_us_pipelined": round(fp64["force_piped_ms"] * 1e3, 3)}
        result["speedup_mixed_vs_fp64"] = {
            "whole_step": round(mixed["pair_evals_per_s"] / fp64["pair_evals_per_s"], 3),
            "force_kernel": round(fp64["force_ms"] / mixed["force_ms"], 3),
            "force_kernel_vs_pipelined_fp64": round(fp64["force_piped_ms"] / mixed["force_ms"], 3)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
