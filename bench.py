"""Benchmark of the multi-system relaxation path on one MI355X.

Metric: relaxed systems/s at Fmax < 1e-3 eV/A.

--workload uniform (default): 4096 replicas of the 38-atom PTCDA-like
  molecule (C24H8O6), each with a seeded in-plane rotation, xy offset
  U(0,24) A and height U(3,5) A, on a 24 x 24 x 20 A rock-salt GridFF at
  0.1 A (240 x 240 x 200 nodes, 3 FP32 channels), FIRE with dt0 = 0.1 fs,
  dt_max = 1 fs, at most 2000 steps.
--workload uneven: 2048 random fused-ring molecules with atom counts
  log-uniform in [20, 400], xy U(0,48) A, height U(3,5) A, on a 48 x 48 x 20 A
  GridFF at 0.1 A (480 x 480 x 200 nodes); same FIRE settings.
--workload layer: one coupled monolayer of 512 PTCDA-like molecules (16 x 32
  square lattice, 13 A by default) with intermolecular pairs within 12 A
  (device cell list) and one FIRE state, on the 6 x 6 x 20 A rock-salt cell at
  0.1 A; `value` counts relaxed layers/s.  The molecule spans 11.3 x 6.9 A
  between atom centres, so an 11 A lattice leaves 2.05 A contacts in its best
  flat arrangement: molecules are squeezed out of the layer and FIRE does not
  reach Fmax < 1e-3 (Fmax 0.71 eV/A after 2000 steps in FP64 and on the
  device alike, 3.6 after 100k device steps); at 13 A the layer relaxes in
  about 6500-6700 steps (the count depends on the last bits of the forces).  --layer-spacing 11 measures the step rate at 11 A.
--workload serve: the serving loop -- relax requests with Poisson arrivals
  (rates --rates, req/s), sizes lognormal (median 60 atoms, ln-sigma 0.5,
  clipped to [20, 600]), on the 48 x 48 x 20 A GridFF; reports p50/p99
  latency and served throughput per rate (one JSON line).

One "step" of this script is one complete batched relaxation of every
system of the workload (positions restored on the device before each).
`value` counts only systems that reached Fmax < 1e-3 eV/A.  Also reported:
the plain method (one launch sequence per FIRE step, host FIRE decision) on a
sample of systems, and the FP64 CPU restatement (single thread) on a bounded
sample of the same systems.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_PEAK_TFLOPS = 157.3     # MI355X FP32 (vector) peak
HBM_PEAK_GBS = 8000.0


def falg_per_system_step(nb, na, nd, ni, npairs, n):
    """Algorithmic FLOP of one system-step (FMA = 2, div/sqrt/exp = 1)."""
    return 28 * nb + 80 * na + 120 * nd + 120 * ni + 41 * npairs + 740 * n + 40 * n


def measured_traffic(args):
    """HBM bytes of the k_relax launches of one relaxation from the committed
    rocprofv3 counter runs (FETCH_SIZE and WRITE_SIZE, copy-calibrated), for
    the workload they were measured on; (None, None) otherwise.  Counters need their own profiler
    run, so this figure cannot be taken inside the timed loop."""
    if args.workload != "uniform" or args.nsys != 4096 or args.max_steps != 2000:
        return None, None
    path = os.path.join(ROOT, "profiles", "k_relax_uniform4096_ptcda.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        rec = json.load(f)
    return float(rec["traffic_bytes_per_relaxation"]), rec["source"]


def grid_for(ctx, S, W, L, zlen=20.0, step=0.1, alpha=1.5, rdamp=1.0, npbc=2):
    """Rock-salt GridFF of an L x L x zlen cell at `step` (device FP64 build + fit)."""
    nxy, nz = int(round(L / step)), int(round(zlen / step))
    el, spos, sq = W.rocksalt_slab(L, L)
    sreq = S.substrate_reqs(el, sq)
    spec = S.GridSpec(nxy, nxy, nz, (0.0, 0.0, -1.0), (step, step, step), alpha, rdamp)
    t0 = time.time()
    grid = S.GridFF.build(ctx, spec, spos, sreq, npbc=npbc)
    return grid, spec, time.time() - t0


def workload_uniform(args, S, W):
    mol = W.ptcda_like(W.SEED_BASE)
    ang, xy, h = W.uniform_batch_poses(args.nsys, L=24.0, hmin=3.0, hmax=5.0, seed=W.SEED_BASE + 1)
    X = [W.place_on_surface(mol, ang[s], xy[s], h[s]) for s in range(args.nsys)]
    desc = ("uniform batch: %d PTCDA-like C24H8O6 replicas (38 atoms), rotation U(0,2pi), xy U(0,24) A, height "
            "U(3,5) A, on a 24x24x20 A rock-salt GridFF at 0.1 A (240x240x200 nodes); UFF + intra-molecular "
            "LJ/Coulomb + GridFF B-spline; FIRE dt0 0.1 fs, dt_max 1 fs, <= 2000 steps" % args.nsys)
    return [mol], np.zeros(args.nsys, dtype=np.int32), X, 24.0, desc


def workload_layer(args, S, W):
    sp = args.layer_spacing
    mol, X = W.layer_lattice(16, 32, spacing=sp, seed=W.SEED_BASE + 3)
    desc = (f"coupled layer: 512 PTCDA-like molecules (19,456 atoms) on a 16x32 lattice at {sp:g} A (0/90 degree "
            "checkerboard orientations, +-2 degree jitter, height U(3,4) A); UFF + intra-molecular LJ/Coulomb + "
            "intermolecular LJ/Coulomb within 12 A (energy switched off over 10-12 A; Verlet lists with a 1 A skin in "
            "one cooperative launch, or the device cell list of the step launches) + GridFF B-spline on the 6x6x20 A rock-salt "
            "cell at 0.1 A (60x60x200 nodes, images to +-60 A); one FIRE state, dt0 0.1 fs, dt_max 1 fs, "
            f"<= {args.max_steps} steps")
    return [mol], np.zeros(len(X), dtype=np.int32), X, 6.0, desc


def workload_uneven(args, S, W):
    nsys = args.nsys if args.nsys_set else 2048
    mols = W.uneven_batch(nsys, 20, 400, seed=W.SEED_BASE + 2)
    ang, xy, h = W.uniform_batch_poses(nsys, L=48.0, hmin=3.0, hmax=5.0, seed=W.SEED_BASE + 3)
    X = [W.place_on_surface(m, ang[s], xy[s], h[s]) for s, m in enumerate(mols)]
    desc = ("uneven batch: %d random fused-6-ring C/N/O/H molecules, atom count log-uniform [20,400], rotation "
            "U(0,2pi), xy U(0,48) A, height U(3,5) A, on a 48x48x20 A rock-salt GridFF at 0.1 A (480x480x200 "
            "nodes); UFF + intra-molecular LJ/Coulomb + GridFF B-spline; FIRE dt0 0.1 fs, dt_max 1 fs, "
            "<= 2000 steps; launch shapes chosen by the autotuner per size class" % nsys)
    return mols, np.arange(nsys, dtype=np.int32), X, 48.0, desc


def raff_bench(args, S, W, nsys, nsteps=100, warmup=10, cpu_seconds=0.0, serve=True):
    """The reference's RAFF benchmark (crates/libs/oclff/src/bin/
    raff_ocl_benchmark.rs): xylitol (22 atoms, 10 nodes) packed into 32-atom
    groups, port stiffness 200, its LCG perturbations, StepConfig dt 0.1,
    relaxation 0.5, no collisions; `nsteps` steps of every replica timed after
    `warmup` steps.  The unit is the reference report's: real atom-steps/s."""
    from surfmol_amd.raff import RaffMolecule, RaffReplicas, StepConfig
    el, x, bonds = W.xylitol_raff_molecule(os.path.join(ROOT, "tests", "golden", "xylitol.xyz"))
    m = RaffMolecule.build(el, x, bonds, group_size=32)
    X = W.raff_perturbed_replicas(m.pos3, [e == "H" for e in m.elems], m.is_padding, nsys)
    ctx = S.Context(0)
    rep = RaffReplicas.from_molecule(ctx, m, X, nsys)
    cfg = StepConfig(dt=0.1, k_coll=0.0, relaxation=0.5)
    rep.step(cfg, warmup)
    t0 = time.perf_counter()
    rep.step(cfg, nsteps)                      # one persistent launch (synchronises)
    T = time.perf_counter() - t0
    # the reference's shape: one call (and launch) per step
    t0 = time.perf_counter()
    for _ in range(20):
        rep.step(cfg, 1)
    T1 = (time.perf_counter() - t0) / 20
    nports = int((m.neighs[:m.nnode] >= 0).sum())
    flop_step = 100.0 * nports + 30.0 * m.nreal + 60.0 * m.nnode    # per replica-step (stated in DESIGN)
    ach = flop_step * nsys * nsteps / T / 1e12
    out = {"metric": "RAFF replica relaxation throughput (real atom-steps/s)",
           "value": round(m.nreal * nsys * nsteps / T, 1), "unit": "atom-steps/s", "n_gpus": 1,
           "steps": nsteps, "warmup": warmup, "ms_per_step": round(1e3 * T / nsteps, 5), "higher_is_better": True,
           "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"RAFF multi-replica: {nsys} xylitol replicas (22 atoms, 10 nodes, 21 ports, "
                                  f"packed to 32), K 200, LCG perturbations 0.2 A (H 0.3 A), dt 0.1, relaxation 0.5, "
                                  f"no collisions; {nsteps} steps in one persistent launch",
                      "nsys": nsys, "natoms_real": m.nreal},
           "one_launch_per_step": {"ms_per_step": round(1e3 * T1, 4),
                                   "value": round(m.nreal * nsys / T1, 1), "unit": "atom-steps/s"},
           "roofline": {"bound": "valu", "achieved": round(ach, 3), "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": round(ach / FP32_PEAK_TFLOPS, 4), "flop_per_replica_step": flop_step,
                        "traffic": None, "note": "state stays in registers/LDS during the launch"},
           "reference_published": {"value": 5.6e8, "unit": "atom-steps/s", "hardware": "NVIDIA GTX 1650, 5000 "
                                   "replicas", "source": "notes/reports/2026-08-29_gpu_raff_benchmark.md:24"}}
    # serving: Poisson arrivals of single-replica relax requests (tol 2e-4,
    # <= 3000 steps) into a 4096-slot server, 64-step chunks
    if serve:
        from surfmol_amd.raff import RaffServer
        nreq, rate = min(nsys, 4000), 20000.0
        tmpl = RaffReplicas.from_molecule(ctx, m, X[:1], 1)
        srv = RaffServer(tmpl, capacity=4096, chunk_steps=64, max_steps=3000, tol=2e-4, cfg=cfg)
        arr = W.poisson_arrivals(rate, nreq, W.SEED_BASE + 77)
        ids = []
        t0 = time.perf_counter()
        for k in range(nreq):
            wait = arr[k] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(X[k]))
        srv.drain()
        res = [srv.poll(i) for i in ids]
        sst = srv.stats()
        srv.close()
        tmpl.close()
        lat = np.array([r["latency"] for r in res])
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        out["serve"] = {"rate_req_s": rate, "requests": nreq,
                        "converged": int(sum(1 for r in res if r["status"] == 1)),
                        "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
                        "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
                        "served_per_s": round(nreq / span, 1),
                        "steps_mean": round(float(np.mean([r["steps"] for r in res])), 1),
                        "chunks": int(sst["chunks"]), "peak_resident": int(sst["peak_resident"]),
                        "config": "4096 slots, 64-step chunks, tol 2e-4 on the step's largest correction, <= 3000 steps"}
    if cpu_seconds > 0:
        from oracle import oracle as O
        O.build()
        t = O.raff_pack(el, x, bonds, group_size=32)
        orc = O.RaffReplica(t)
        n, tc = 0, 0.0
        while tc < cpu_seconds:
            t0 = time.perf_counter()
            orc.run(X[n % nsys], nsteps=100)
            tc += time.perf_counter() - t0
            n += 1
        out["cpu_baseline"] = {"value": round(m.nreal * 100 * n / tc, 1), "unit": "atom-steps/s", "cores": 1,
                               "kind": "port", "sample": f"{n} replicas x 100 steps on the FP64 C restatement "
                                                         f"(oracle/oracle.c orc_raff_step), one thread, {tc:.1f} s",
                               **host_info()}
    rep.close()
    return out


def host_info():
    """Model name and usable core count of this host (the CPU baseline's)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count()
    return {"cpu_model": model or "unknown", "cores_usable": cores, "logical_cpus": os.cpu_count()}


def cpu_relax_rate(O, osys_list, X_of, nsys, ofire, max_steps, fconv, seconds, threads):
    """Converged systems/s of the FP64 C restatement (oracle/oracle.c) relaxing
    systems 0, 1, 2, ... of the batch until `seconds` have passed, on
    `threads` host threads (the C code runs outside the GIL; each thread owns
    one oracle system whose positions are replaced per system)."""
    import threading
    lock = threading.Lock()
    state = {"next": 0, "done": 0, "conv": 0}
    t_end = time.perf_counter() + seconds

    def worker(osys):
        while True:
            with lock:
                if state["next"] >= nsys or time.perf_counter() > t_end:
                    return
                s = state["next"]
                state["next"] += 1
            r = O.relax_fire(osys, ofire, max_steps, fconv, apos=X_of(s))
            with lock:
                state["done"] += 1
                state["conv"] += int(r["converged"])

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker, args=(osys_list[k],)) for k in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    T = time.perf_counter() - t0
    return state["conv"] / T, state["done"], T


def run_workload(args, S, W, workload, steps, warmup, cpu_seconds, want_plain, want_cpu, nsys=None):
    """One batched-relaxation workload: `steps` timed relaxations of every
    system (positions restored on the device before each) and the result
    dictionary of the bench line."""
    alpha, rdamp = 1.5, 1.0
    layer = workload == "layer"
    max_steps = args.max_steps if args.max_steps is not None else (20000 if layer else 2000)
    wargs = argparse.Namespace(**vars(args))
    wargs.max_steps = max_steps
    wargs.nsys = nsys if nsys is not None else (args.nsys if args.nsys is not None else 4096)
    wargs.nsys_set = nsys is not None or args.nsys is not None
    t0 = time.time()
    wl = {"uniform": workload_uniform, "uneven": workload_uneven, "layer": workload_layer}[workload]
    mols, sys_topo, Xs, L, desc = wl(wargs, S, W)
    rc = 12.0 if layer else 0.0
    nsys = len(sys_topo)
    ctx = S.Context(0)
    grid, spec, t_grid = grid_for(ctx, S, W, L, alpha=alpha, rdamp=rdamp, npbc=10 if layer else 2)
    if layer and getattr(args, "layer_grid", "cell") == "full":
        # the configuration's 180 x 360 x 20 A grid: the surface is periodic
        # with the 6 A cell, so its B-spline coefficients are the cell's,
        # tiled (30 x 60 cells)
        t0 = time.time()
        C = grid.coefficients()
        big = np.tile(C, (30, 60, 1, 1))
        del C
        spec = S.GridSpec(big.shape[0], big.shape[1], big.shape[2], spec.p0, spec.step, spec.alpha_morse, spec.rdamp)
        grid = S.GridFF.upload(ctx, spec, big)
        del big
        t_grid += time.time() - t0
        desc = desc.replace("on the 6x6x20 A rock-salt cell at 0.1 A (60x60x200 nodes, images to +-60 A)",
                            "on the 180x360x20 A rock-salt grid at 0.1 A (1800x3600x200 nodes, 20.7 GB of float4: "
                            "the 6x6x20 A cell's coefficients, images to +-60 A, tiled)")
    hms = [S.Molecule.build(m.elems, Xs[list(sys_topo).index(t)], m.charges, bonds=m.bonds) for t, m in enumerate(mols)]
    X = np.concatenate(Xs)
    batch = S.Batch(ctx, hms, sys_topo, X, grid=grid, rdamp=rdamp, inter_cutoff=rc, inter_switch=2.0 if layer else 0.0)
    fire = S.FireParams(dt0=W.fs_to_internal(0.1), dt_max=W.fs_to_internal(1.0))
    fconv = 1e-3
    if args.no_graph:
        batch.set_graph(False)
    if layer and args.layer_step_launches:
        batch.set_layer_persistent(False)
    tuned = None
    if args.autotune or workload == "uneven":
        t0 = time.time()
        tuned = batch.autotune(fire, 50)
        t_tune = time.time() - t0

    for _ in range(warmup):
        batch.reset()
        batch.relax_batched(fire, max_steps, fconv)
    times, kms, results = [], [], None
    for _ in range(steps):
        batch.reset()
        t0 = time.perf_counter()
        results = batch.relax_batched(fire, max_steps, fconv)   # synchronises the stream
        times.append(time.perf_counter() - t0)
        kms.append(batch.last_launch()["kernel_ms"])
    T = sum(times)
    X_dev = batch.positions()
    # the same relaxation with the boundary's host buffers: FP64 positions in
    # (smr_batch_set_positions, over PCIe), relaxed, positions out
    batch.reset()
    t0 = time.perf_counter()
    batch.set_positions(X)
    r_io = batch.relax_batched(fire, max_steps, fconv)
    _ = batch.positions()
    t_io = time.perf_counter() - t0
    nconv_io = sum(1 for r in r_io if r["status"] == S.SMR_SYS_CONVERGED)
    nconv = sum(1 for r in results if r["status"] == S.SMR_SYS_CONVERGED)
    steps_arr = np.array([r["steps"] for r in results])
    if layer:
        nconv = 1 if nconv == nsys else 0     # the layer is one relaxed system
    value = nconv * steps / T

    # algorithmic FLOP and grid-lookup bytes of one relaxation (sum over systems of per-step work x steps taken)
    counts = [hm.counts for hm in hms]
    flop_launch, bytes_launch, natoms_sys = 0.0, 0.0, []
    for s in range(nsys):
        n, nb, na, nd, ni = counts[sys_topo[s]]
        npairs = n * (n - 1) // 2 - nb
        flop_launch += falg_per_system_step(nb, na, nd, ni, npairs, n) * float(steps_arr[s])
        bytes_launch += 768.0 * n * float(steps_arr[s])
        natoms_sys.append(n)
    if layer:
        # intermolecular pairs within the cutoff at the start (unique pairs), per step
        from scipy.spatial import cKDTree
        mol_of = np.repeat(np.arange(nsys), [counts[t][0] for t in sys_topo])
        pairs = cKDTree(X).query_pairs(rc, output_type="ndarray")
        n_inter = int((mol_of[pairs[:, 0]] != mol_of[pairs[:, 1]]).sum())
        flop_launch += 41.0 * n_inter * float(steps_arr[0])
    kavg_s = (sum(kms) / len(kms)) / 1e3
    achieved_tf = flop_launch / kavg_s / 1e12
    traffic, traffic_src = measured_traffic(wargs) if workload == "uniform" else (None, None)
    natoms_sys = np.array(natoms_sys)

    cfg = {
        "workload": desc,
        "nsys": nsys,
        "relaxed_unit": "layer (all systems, one FIRE state)" if layer else "system",
        "natoms_per_system": int(natoms_sys[0]) if workload == "uniform" else
        {"min": int(natoms_sys.min()), "mean": round(float(natoms_sys.mean()), 1), "max": int(natoms_sys.max())},
        "converged": nconv,
        "fire_steps_mean": round(float(steps_arr.mean()), 1),
        "fire_steps_max": int(steps_arr.max()),
        "grid_build_fit_s": round(t_grid, 3),
        "plan": batch.plan(),
        "precision": "FP32 forces with Kahan per-atom sums; FP64 energies, FIRE reductions and convergence test",
    }
    if tuned is not None:
        cfg["autotune_s"] = round(t_tune, 2)
    out = {
        "metric": "relaxed systems/sec at Fmax<1e-3 eV/A (batched)",
        "value": round(value, 4 if layer else 2),
        "unit": "layers/s" if layer else "systems/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(1e3 * T / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": cfg,
        "roofline": {
            "bound": "valu",
            "note": "compute roof = FP32 vector (VALU) peak: no step is a dense contraction, so MFMA is unused",
            "achieved": round(achieved_tf, 3),
            "peak": FP32_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tf / FP32_PEAK_TFLOPS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": "k_relax",
            "kernel_ms_note": "HIP-event time of the relaxation's k_relax launches (two when the schedule is re-ranked)",
            "flop_per_relaxation": flop_launch,
            "kernel_ms_avg": round(1e3 * kavg_s, 3),
            "grid_lookup_GBs": round(bytes_launch / kavg_s / 1e9, 1),
        },
    }

    out["host_buffers_inclusive"] = {
        "value": round((1 if layer and nconv_io == nsys else (0 if layer else nconv_io)) / t_io, 4 if layer else 2),
        "unit": out["unit"], "note": "one relaxation including the FP64 position upload and download over PCIe "
        "(smr_batch_set_positions / smr_batch_get_positions); not `value`"}
    if layer:
        cfg["intermolecular_pairs_within_cutoff"] = n_inter
        cfg["fmax_final"] = round(float(results[0]["fmax"]), 6)
        # the 11 A square lattice keeps rearranging for more than 20k FIRE steps,
        # so the step rate of the coupled layer is reported beside `value`
        out["layer_fire_steps_per_s"] = round(float(steps_arr[0]) * steps / T, 1)
        out["atom_steps_per_s"] = round(float(steps_arr[0]) * steps / T * len(X), 1)
        path = batch.layer_path()
        cfg["layer_path"] = ("one cooperative launch (Verlet lists, grid barriers)" if path["persistent"] else
                             "step launches (force launch, decision, cell list)" +
                             (f"; one launch not taken: {path['fallback']}" if path["fallback"] else ""))
        cfg["us_per_fire_step"] = round(1e6 * T / (float(steps_arr[0]) * steps), 2)
        cfg["energy_eV_per_atom"] = round(float(sum(r["energy"] for r in results)) / len(X), 9)
        if path["persistent"]:
            cfg["neighbour_list_builds"] = path["rebuilds"]
        if path["persistent"] and not args.no_extras:
            # the other path on the same layer, one relaxation
            batch.set_layer_persistent(False)
            batch.reset()
            t0 = time.perf_counter()
            r_s = batch.relax_batched(fire, max_steps, fconv)
            t_s = time.perf_counter() - t0
            batch.set_layer_persistent(True)
            ok_s = all(r["status"] == S.SMR_SYS_CONVERGED for r in r_s)
            out["step_launch_path"] = {"value": round((1.0 if ok_s else 0.0) / t_s, 4), "unit": "layers/s",
                                       "fire_steps": int(r_s[0]["steps"]),
                                       "us_per_fire_step": round(1e6 * t_s / float(r_s[0]["steps"]), 2),
                                       "energy_eV_per_atom": round(float(sum(r["energy"] for r in r_s)) / len(X), 9),
                                       "fmax_final": round(float(r_s[0]["fmax"]), 6)}
    if want_plain and not layer:
        pts = []
        for s in range(min(args.plain_sample, nsys)):
            batch.reset()
            t0 = time.perf_counter()
            r = batch.relax_plain(s, fire, max_steps, fconv)
            dt = time.perf_counter() - t0
            pts.append((dt, r["status"] == S.SMR_SYS_CONVERGED, r["steps"]))
        tp = sum(p[0] for p in pts)
        plain_rate = sum(1 for p in pts if p[1]) / tp
        out["plain_gpu"] = {"value": round(plain_rate, 3), "unit": "systems/s", "sample": f"first {len(pts)} "
                            f"systems, one launch sequence + host FIRE decision per step", "ms_per_fire_step":
                            round(1e3 * tp / sum(p[2] for p in pts), 4)}
        out["speedup_vs_plain"] = round(value / plain_rate, 1) if plain_rate > 0 else None

    if want_cpu:
        # the oracle here is the checker and the timed CPU baseline only
        from oracle import oracle as O
        O.build()
        P = O.Params(S.DATA_DIR)
        C = grid.coefficients().astype(np.float64)
        ofire = O.FireParams(dt0=fire.dt0, dt_max=fire.dt_max)
        host = host_info()
        gkw = dict(rdamp=rdamp, grid=C, grid_ns=(spec.nx, spec.ny, spec.nz), grid_p0=spec.p0, grid_step=spec.step,
                   alpha_morse=alpha)
        if layer:
            osys = O.layer_from_molecules([mols[0]] * nsys, Xs, P, rc, layer_cells=True, rswitch_inter=2.0, **gkw)
            osys.cstruct()
            nst, t_cpu = 0, 0.0
            x, v = None, None
            while t_cpu < cpu_seconds:
                t0 = time.perf_counter()
                r = O.relax_fire(osys, ofire, 10, 1e-3, apos=x, vapos=v)
                t_cpu += time.perf_counter() - t0
                x, v = r["apos"], r["vapos"]
                nst += 10
            # the FIRE state restarts every 10 steps in this timing loop; the cost per step is what is measured
            t_step = t_cpu / nst
            cpu_rate = 1.0 / (t_step * float(steps_arr[0]))
            out["cpu_baseline"] = {"value": round(cpu_rate, 6), "unit": "layers/s", "cores": 1, "kind": "port",
                                   "sample": f"{nst} FIRE steps of the same 19,456-atom layer on the FP64 C "
                                             f"restatement (oracle/oracle.c, cell-grid pair search), single thread, "
                                             f"{t_cpu:.1f} s; {1e3 * t_step:.1f} ms/step x the {int(steps_arr[0])} "
                                             f"steps the device took", "fire_steps_per_s": round(1.0 / t_step, 2),
                                   **host}
            out["speedup_vs_cpu"] = round(out["layer_fire_steps_per_s"] * t_step, 1)
        else:
            def make_osys(s):
                m = mols[sys_topo[s]]
                o = O.system_from_geometry(m.elems, Xs[s], P, charges=m.charges, bonds=m.bonds, nb_mode=O.NB_GPU, **gkw)
                o.cstruct()
                return o
            uniform = all(t == sys_topo[0] for t in sys_topo)
            ncores = host["cores_usable"]
            if uniform:
                # one oracle system per thread (replicas: only the positions change)
                pool = [make_osys(0) for _ in range(ncores)]
                r1, d1, T1 = cpu_relax_rate(O, pool[:1], lambda s: Xs[s], nsys, ofire, max_steps, fconv, cpu_seconds, 1)
                rN, dN, TN = cpu_relax_rate(O, pool, lambda s: Xs[s], nsys, ofire, max_steps, fconv, cpu_seconds,
                                            ncores)
                what = ("FP64 C restatement of the reference CPU path (oracle/oracle.c: UFF, O(N^2) LJ+Coulomb, "
                        "tricubic GridFF, the reference FIRE loop), whole systems per thread")
                out["cpu_baseline"] = {"value": round(rN, 3), "unit": "systems/s", "cores": ncores, "kind": "port",
                                       "sample": f"systems 0..{dN - 1} of the same batch on {ncores} threads, "
                                                 f"{TN:.1f} s; {what}", **host}
                out["cpu_baseline_1thread"] = {"value": round(r1, 3), "unit": "systems/s", "cores": 1,
                                               "sample": f"systems 0..{d1 - 1}, one thread, {T1:.1f} s (the "
                                                         f"reference CPU path is single-threaded)"}
                out["speedup_vs_cpu"] = round(value / rN, 1) if rN > 0 else None
                out["speedup_vs_cpu_1thread"] = round(value / r1, 1) if r1 > 0 else None
            else:
                done, conv, t_cpu = 0, 0, 0.0
                # a sample that follows the workload's size mix: systems in batch order
                while t_cpu < cpu_seconds and done < nsys:
                    osys = make_osys(done)   # set-up outside the timed region
                    t0 = time.perf_counter()
                    r = O.relax_fire(osys, ofire, max_steps, fconv)
                    t_cpu += time.perf_counter() - t0
                    conv += int(r["converged"])
                    done += 1
                cpu_rate = conv / t_cpu
                out["cpu_baseline"] = {"value": round(cpu_rate, 3), "unit": "systems/s", "cores": 1, "kind": "port",
                                       "sample": f"first {done} systems of the same batch, FP64 C restatement of the "
                                                 f"reference CPU path (oracle/oracle.c), single thread, {t_cpu:.1f} s",
                                       **host}
                out["speedup_vs_cpu"] = round(value / cpu_rate, 1) if cpu_rate > 0 else None
            # error against FP64 on a sample of the timed relaxation's systems
            out["fp64_error"] = fp64_error(O, S, batch, results, X_dev, make_osys, mols, sys_topo, ofire, fconv,
                                           min(8, nsys))
    return out


def fp64_error(O, S, batch, results, X_dev, make_osys, mols, sys_topo, ofire, fconv, nsample):
    """Errors of the device against the FP64 oracle on `nsample` systems:
    forces at the start (max_i|dF_i| / max_i|F_i|) and at the relaxed
    positions (per atom, against the sum of the magnitudes of the atom's force
    contributions); the reported energy against the FP64 energy of the
    returned positions; and how far FP64 FIRE restarted from the device's
    minimum moves the energy (eV/atom)."""
    batch.reset()
    F0, _ = batch.eval_forces()
    X0 = batch.positions()
    off = batch.offsets
    e_start, e_final, e_energy, e_basin = 0.0, 0.0, 0.0, 0.0
    for s in range(nsample):
        sl = slice(off[s], off[s + 1])
        o = make_osys(s)
        n = off[s + 1] - off[s]
        fo, _ = O.eval_forces(o, X0[sl])
        e_start = max(e_start, float(np.linalg.norm(F0[sl] - fo, axis=1).max() / np.linalg.norm(fo, axis=1).max()))
        ff, eo = O.eval_forces(o, X_dev[sl])
        e_energy = max(e_energy, abs(float(eo.sum()) - float(results[s]["energy"])) / n)
        rf = O.relax_fire(o, ofire, 3000, fconv, apos=X_dev[sl])
        e_basin = max(e_basin, abs(rf["E"] - float(results[s]["energy"])) / n)
    # forces at the relaxed positions
    batch.set_positions(X_dev)
    Ff, _ = batch.eval_forces()
    for s in range(nsample):
        sl = slice(off[s], off[s + 1])
        o = make_osys(s)
        ff, _ = O.eval_forces(o, X_dev[sl])
        sc = O.force_scale(o, X_dev[sl])
        e_final = max(e_final, float((np.linalg.norm(Ff[sl] - ff, axis=1) / sc).max()))
    batch.reset()
    return {"systems": nsample, "force_rel_start": float("%.3g" % e_start),
            "force_rel_to_contributions_relaxed": float("%.3g" % e_final),
            "energy_eV_per_atom": float("%.3g" % e_energy),
            "fp64_restart_dE_eV_per_atom": float("%.3g" % e_basin),
            "note": "energies are FP64 evaluations of the device's FP32 positions; the restart term is the energy "
                    "change when FP64 FIRE continues from the device's relaxed state"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["uniform", "uneven", "layer", "serve", "raff"], default="uniform")
    ap.add_argument("--rates", type=str, default="200,1000,5000", help="serve: arrival rates [req/s]")
    ap.add_argument("--requests", type=int, default=1000, help="serve: requests per rate")
    ap.add_argument("--chunk-steps", type=int, default=64, help="serve: FIRE steps per persistent launch (<= 64 atoms)")
    ap.add_argument("--nsys", type=int, default=None)
    ap.add_argument("--max-steps", type=int, default=None, help="FIRE step cap (2000; 20000 for the layer)")
    ap.add_argument("--layer-step-launches", action="store_true",
                    help="layer: relax by step launches (force launch, decision, cell list) instead of one "
                         "cooperative launch")
    ap.add_argument("--layer-spacing", type=float, default=13.0,
                    help="layer: square lattice spacing [A]; at 11 A the molecule has no flat arrangement with "
                         "contacts above 2.05 A and the layer does not relax (FP64 oracle and device agree)")
    ap.add_argument("--layer-grid", choices=["cell", "full"], default="cell",
                    help="layer: the 6x6x20 A rock-salt cell's grid (60x60x200 nodes, 11.5 MB, wrapped laterally), "
                         "or the same coefficients tiled to the 180x360x20 A grid of the configuration "
                         "(1800x3600x200 nodes, 20.7 GB of float4)")
    ap.add_argument("--autotune", action="store_true", help="time every launch shape per size class first "
                    "(always on for the uneven workload)")
    ap.add_argument("--no-graph", action="store_true",
                    help="issue the launches directly instead of replaying captured HIP graphs (same results)")
    ap.add_argument("--plain-sample", type=int, default=4)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="uniform workload: skip the short uneven, layer and serving measurements added to the line")
    ap.add_argument("--copy-probe", type=int, default=0,
                    help="run only a plain device copy of this many bytes (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = args.copy_probe // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n}))
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); the systems are independent, "
              f"so N GPUs would each relax their own share with no exchange.", file=sys.stderr)
        sys.exit(2)

    import surfmol_amd as S
    from surfmol_amd import workloads as W

    if args.workload == "serve":
        print(json.dumps(serve(args, S, W, 1.5, 1.0, [float(r) for r in args.rates.split(",") if r],
                               args.requests)))
        return
    if args.workload == "raff":
        print(json.dumps(raff_bench(args, S, W, args.nsys or 5000, nsteps=max(args.steps, 100), warmup=10,
                                    serve=not args.no_extras,
                                    cpu_seconds=0.0 if args.no_cpu else args.cpu_seconds)))
        return
    out = run_workload(args, S, W, args.workload, args.steps, args.warmup, args.cpu_seconds, not args.no_plain,
                       not args.no_cpu)
    if args.workload == "uniform" and not args.no_extras:
        # the other configurations, each measured once in the same run (driver-timed with the headline)
        extras = {}
        t0 = time.time()
        un = run_workload(args, S, W, "uneven", 1, 1, 0.0, False, False, nsys=2048)
        extras["uneven"] = {"value": un["value"], "unit": un["unit"], "ms_per_relaxation": un["ms_per_step"],
                            "converged": un["config"]["converged"], "nsys": un["config"]["nsys"],
                            "roofline_frac": un["roofline"]["frac"], "kernel_ms": un["roofline"]["kernel_ms_avg"],
                            "workload": un["config"]["workload"], "wall_s": round(time.time() - t0, 1)}
        print(json.dumps({"progress": "uneven", **extras["uneven"]}), file=sys.stderr, flush=True)
        t0 = time.time()
        largs = argparse.Namespace(**vars(args))
        largs.max_steps = None
        la = run_workload(largs, S, W, "layer", 1, 0, 0.0, False, False)
        extras["layer"] = {"value": la["value"], "unit": la["unit"], "ms_per_relaxation": la["ms_per_step"],
                           "fire_steps": la["config"]["fire_steps_max"],
                           "layer_fire_steps_per_s": la["layer_fire_steps_per_s"],
                           "us_per_fire_step": la["config"]["us_per_fire_step"], "path": la["config"]["layer_path"],
                           "step_launch_path": la.get("step_launch_path"),
                           "roofline_frac": la["roofline"]["frac"], "workload": la["config"]["workload"],
                           "wall_s": round(time.time() - t0, 1)}
        print(json.dumps({"progress": "layer", **extras["layer"]}), file=sys.stderr, flush=True)
        t0 = time.time()
        sargs = argparse.Namespace(**vars(args))
        sargs.max_steps = None
        sv = serve(sargs, S, W, 1.5, 1.0, [1000.0], 500)
        r = sv["per_rate"][0]
        extras["serve"] = {"rate_req_s": r["rate_req_s"], "requests": r["requests"], "p50_ms": r["p50_ms"],
                           "p99_ms": r["p99_ms"], "served_per_s": r["served_per_s"], "converged": r["converged"],
                           "served_vs_plain": r["served_vs_plain"], "workload": sv["config"]["workload"],
                           "wall_s": round(time.time() - t0, 1)}
        t0 = time.time()
        ra = raff_bench(args, S, W, 5000)
        extras["raff"] = {"value": ra["value"], "unit": ra["unit"], "ms_per_step": ra["ms_per_step"],
                          "one_launch_per_step_ms": ra["one_launch_per_step"]["ms_per_step"],
                          "roofline_frac": ra["roofline"]["frac"], "reference_published": ra["reference_published"],
                          "serve": ra["serve"],
                          "workload": ra["config"]["workload"], "wall_s": round(time.time() - t0, 1)}
        out["other_configs"] = extras
    print(json.dumps(out))


def serve(args, S, W, alpha, rdamp, rates, nreq):
    """Serving loop under Poisson arrivals: the result dictionary (one JSON
    line) with p50/p99 latency and served throughput per rate."""
    max_steps = args.max_steps if args.max_steps is not None else 2000
    ctx = S.Context(0)
    grid, spec, t_grid = grid_for(ctx, S, W, 48.0, alpha=alpha, rdamp=rdamp)
    fire = S.FireParams(dt0=W.fs_to_internal(0.1), dt_max=W.fs_to_internal(1.0))
    per_rate = []
    for k, rate in enumerate(rates):
        seed = W.SEED_BASE + 4 + 16 * k
        sizes = W.lognormal_sizes(nreq, median=60.0, sigma=0.5, lo=20, hi=600, seed=seed)
        rng = np.random.Generator(np.random.PCG64(seed + 1))
        mols = [W.random_polyhex(rng, int(n)) for n in sizes]
        ang, xy, h = W.uniform_batch_poses(nreq, L=48.0, hmin=3.0, hmax=5.0, seed=seed + 2)
        X = [W.place_on_surface(m, ang[s], xy[s], h[s]) for s, m in enumerate(mols)]
        hm = [S.Molecule.build(m.elems, X[s], m.charges, bonds=m.bonds) for s, m in enumerate(mols)]
        arrivals = W.poisson_arrivals(rate, nreq, seed + 3)
        # the plain method on the same requests: one request at a time, one
        # launch sequence + host FIRE decision per step; its service times on
        # a sample feed a FIFO single-server queue over the same arrivals
        plain_t = []
        for s in range(min(args.plain_sample, nreq)):
            pb = S.Batch(ctx, [hm[s]], [0], X[s], grid=grid, rdamp=rdamp)
            tq = time.perf_counter()
            pb.relax_plain(0, fire, max_steps, 1e-3)
            plain_t.append(time.perf_counter() - tq)
            pb.close()
        srv = S.Server(ctx, grid, fire, max_resident=4096, max_atoms=600, hbm_fraction=0.25,
                       chunk_steps=args.chunk_steps, max_steps=max_steps, fconv=1e-3, rdamp=rdamp)
        ids = []
        t0 = time.perf_counter()
        for s in range(nreq):
            wait = arrivals[s] - (time.perf_counter() - t0)
            if wait > 0:
                time.sleep(wait)
            ids.append(srv.submit(hm[s], X[s]))
        srv.drain()
        res = [srv.poll(i, positions=False) for i in ids]
        st = srv.stats()
        srv.close()
        lat = np.array([r["latency"] for r in res])
        conv = sum(1 for r in res if r["status"] == S.SMR_SYS_CONVERGED)
        span = max(r["t_done"] for r in res) - min(r["t_submit"] for r in res)
        steps = np.array([r["steps"] for r in res])
        free, plat = 0.0, []
        for s in range(nreq):
            start = max(free, arrivals[s])
            free = start + plain_t[s % len(plain_t)]
            plat.append(free - arrivals[s])
        plat = np.array(plat)
        plain = {"p50_ms": round(1e3 * float(np.percentile(plat, 50)), 1),
                 "p99_ms": round(1e3 * float(np.percentile(plat, 99)), 1),
                 "served_per_s": round(nreq / (free - arrivals[0]), 2),
                 "service_ms_mean": round(1e3 * float(np.mean(plain_t)), 2),
                 "model": f"FIFO single-server queue over the request arrivals with the plain method's service "
                          f"times measured on the first {len(plain_t)} requests"}
        per_rate.append({
            "rate_req_s": rate, "requests": nreq, "converged": conv,
            "p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "mean_ms": round(1e3 * float(lat.mean()), 3),
            "served_per_s": round(conv / span, 1),
            "fire_steps_mean": round(float(steps.mean()), 1),
            "atoms_mean": round(float(np.mean(sizes)), 1),
            "chunks": int(st["chunks"]), "peak_resident": int(st["peak_resident"]),
            "device_busy_frac": round(st["device_s"] / span, 3),
            "plain": plain,
            "served_vs_plain": round((conv / span) / plain["served_per_s"], 1),
        })
        print(json.dumps({"progress": per_rate[-1]}), file=sys.stderr, flush=True)
    out = {
        "metric": "relax request latency p50/p99 (served, continuous batching)",
        "unit": "ms",
        "n_gpus": 1,
        "higher_is_better": False,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "serving: Poisson arrivals, request sizes lognormal (median 60 atoms, ln-sigma 0.5, "
                               "clipped [20,600]) random fused-ring molecules on a 48x48x20 A rock-salt GridFF at "
                               "0.1 A; FIRE dt0 0.1 fs, dt_max 1 fs, Fmax < 1e-3 eV/A, <= %d steps; %d FIRE steps "
                               "per chunk" % (max_steps, args.chunk_steps),
                   "grid_build_fit_s": round(t_grid, 3)},
        "per_rate": per_rate,
    }
    return out


if __name__ == "__main__":
    main()
