#!/usr/bin/env python3
"""Particle hot-path benchmark: particle-steps/s of push + absorption +
deposit + MCC on one MI355X, with the HBM roofline of the dominant kernel and
the CPU oracle timed on the host as a reported baseline.

    python bench.py [--steps K] [--warmup W] [--config c2]

A step is one fused pass over every species (push, absorb, deposit,
null-collision selection), the collision pass over the selected particles,
and the grid epilogue (RhoOne/ChiOne per species, Rho/Chi sums), plus the
cell sort every ``sort_interval`` steps.  Particles are resident in HBM
before the timed region; the field is a fixed synthetic potential (the host
Poisson solve is not part of the particle path).  Prints one JSON line.

``value`` times the step as one captured graph replay (``ipm_step_graph``),
after a pass that launches the same kernels one by one with HIP events
around each (``eager_value``, and the kernel times the roofline uses).  The
cell sort enters every step at its amortised cost: the sort time measured
after a full sort interval of drift, divided by the interval.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "2dmatching-cf4_amd"))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "particle_kernel_c2_traffic.json")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def stream_ceiling(n, reps=5):
    """What a pure streaming pass with the particle pass's access pattern
    (8 arrays of n doubles read, 7 written: 120 B per element) and a plain
    one-array copy reach on this device (csrc/ipm_probe.hip)."""
    import ctypes as C
    import torch
    lib = C.CDLL(os.path.join(ROOT, "2dmatching-cf4_amd", "ipm2d", "_lib", "libipm_probe.so"))
    lib.ipm_probe_stream.argtypes = [C.POINTER(C.c_void_p), C.c_longlong, C.c_int, C.c_void_p,
                                     C.POINTER(C.c_double)]
    lib.ipm_probe_stream.restype = C.c_int
    arrs = [torch.zeros(n, dtype=torch.float64, device="cuda") for _ in range(8)]
    ptrs = (C.c_void_p * 8)(*[a.data_ptr() for a in arrs])
    ms = (C.c_double * 3)()
    torch.cuda.synchronize()
    rc = lib.ipm_probe_stream(ptrs, n, reps, C.c_void_p(torch.cuda.current_stream().cuda_stream), ms)
    torch.cuda.synchronize()
    del arrs
    if rc != 0:
        raise RuntimeError(f"ipm_probe_stream failed with HIP error {rc}")
    plain = 120.0 * n / (ms[0] * 1e-3) / 1e9
    nt = 120.0 * n / (ms[2] * 1e-3) / 1e9
    return {"value": max(plain, nt), "unit": "GB/s", "plain_GBps": plain, "nontemporal_GBps": nt,
            "copy_GBps": 16.0 * n / (ms[1] * 1e-3) / 1e9,
            "what": f"median of {reps} passes of a kernel that reads 8 and writes 7 arrays of {n} doubles "
                    f"(the particle pass's streams, no compute), one element per thread, with plain and with "
                    f"non-temporal accesses (value: the faster); copy_GBps: one-array device copy (read + write)"}


# Products one collision of each reaction type makes (the resultants of
# ElectronIonization, ElectronAttachment, ElectronDissociation and the
# reactive ion channels, R/src/mc/MCCElectron.F90:47-147, R/src/mc/MCCIon.F90:54-74).
PRODUCTS_BY_TYPE = {13: 2, 14: 1, 15: 2, 112: 1, 113: 1}


def collision_counts(hp, tables, steps=4):
    """Selections (null events included) and created particles per step,
    from the event counts of `steps` further steps.  The whole step's
    algorithmic bytes add 48 B per selected particle (its velocity read and
    written around the collision) and 64 B per created particle (8 fields
    written) to the particle pass's 120 B per particle."""
    hp.set_event_counting(True)
    for _ in range(steps):
        hp.one_step()
    hp.synchronize()
    c = hp.take_mcc_counts()
    hp.set_event_counting(False)
    selected = float(c.sum()) / steps
    created = 0.0
    for (s, g), t in tables.items():
        if t is None or getattr(t, "rtype", None) is None:
            continue
        for j, ty in enumerate(t.rtype):
            created += float(c[s, g, j + 1]) * PRODUCTS_BY_TYPE.get(int(ty), 0)
    return selected, created / steps


def live_total(hp, ns):
    return sum(hp.count(s) for s in range(ns))


def sort_phase(steps, sort_interval):
    """Steps after a sort at which a timed region of `steps` steps starts, so
    that the region sits in the middle of the sort cycle: the particle pass
    slows as particles drift out of cell order between sorts, so a region
    that began right after a sort would time the fastest part of the cycle
    only (and one that began late, the slowest)."""
    if not sort_interval:
        return 0
    return (sort_interval - steps % sort_interval) // 2 if steps % sort_interval else 0


def start_phase(hp, phase, graph=False):
    """Sort, then advance `phase` untimed steps."""
    hp.sort(-1)
    for _ in range(phase):
        if graph:
            hp.one_step_graph()
        else:
            hp.one_step()
    hp.synchronize()


def timed_steps(hp, cfg, steps, sort_interval, graph=False, phase=0):
    """K steps with the cell sort every sort_interval steps, the first step
    being step `phase` after a sort.  Returns the wall time, particle-steps,
    populations before/after and the wall time spent in the sorts (each sort
    starts by synchronizing the stream, so its own time is bracketed by two
    synchronizations)."""
    hp.synchronize()
    n0 = live_total(hp, cfg.ns)
    t_sort = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        if graph:
            hp.one_step_graph()
        else:
            hp.one_step()
        if sort_interval and (phase + k + 1) % sort_interval == 0:
            hp.synchronize()
            ts = time.perf_counter()
            hp.sort(-1)
            t_sort += time.perf_counter() - ts
    hp.synchronize()
    t1 = time.perf_counter()
    n1 = live_total(hp, cfg.ns)
    return t1 - t0, 0.5 * (n0 + n1) * steps, n0, n1, t_sort


def sort_after_full_interval(hp, sort_interval):
    """Wall time of one sort of every species after sort_interval steps of
    drift from a sorted state (what each sort of a long run costs)."""
    hp.sort(-1)
    for _ in range(sort_interval):
        hp.one_step_graph()
    hp.synchronize()
    t = time.perf_counter()
    hp.sort(-1)
    return time.perf_counter() - t


def interval_in_graph(hp, sort_interval):
    """Wall time per step of one whole sort interval replayed with the sort
    inside the graph (ipm_set_sort_interval): sort_interval graph steps from
    a sorted state, the last of which ends with the sort.  A cross-check of
    the amortised step: no host synchronization inside."""
    hp.sort(-1)
    hp.set_sort_interval(sort_interval)
    hp.synchronize()
    t = time.perf_counter()
    for _ in range(sort_interval):
        hp.one_step_graph()
    hp.synchronize()
    el = time.perf_counter() - t
    hp.set_sort_interval(0)
    return el / sort_interval


def host_cpus():
    """Host CPU description: logical CPUs, the ones this process may use, the
    cgroup CPU quota (if any) and the model name."""
    info = {"nproc": os.cpu_count()}
    try:
        info["affinity"] = len(os.sched_getaffinity(0))
    except Exception:
        info["affinity"] = None
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = float(q) / float(per)
    except Exception:
        pass
    info["cgroup_cpu_quota"] = quota
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.lower().startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except Exception:
        pass
    info["model"] = model
    return info


def native_oracle():
    """Compile the oracle with -O3 -march=native for this host (same source,
    contraction off) into a temporary directory; None if that fails."""
    import subprocess
    import tempfile
    src = os.path.join(ROOT, "oracle", "ipm_oracle.c")
    out = os.path.join(tempfile.mkdtemp(prefix="ipm_oracle_"), "libipm_oracle_native.so")
    cmd = ["gcc", "-shared", "-fPIC", "-std=c99", "-ffp-contract=off", "-fno-fast-math", "-O3", "-march=native",
           src, "-o", out, "-lm"]
    try:
        subprocess.run(cmd, check=True, capture_output=True, timeout=120)
        return out
    except Exception as e:
        log(f"[bench] -march=native oracle build failed ({e}); using the prebuilt -O3 oracle")
        return None


def plain_steps(hp, cfg, steps):
    """Per-species launches in the reference's call order (OneStep,
    R/src/onestep.F90:141-177: MoveES and MaterialAborptionParticle per
    species, P2CES per species, FieldSource%Sum, MCC), each call synchronous."""
    hp.synchronize()
    n0 = live_total(hp, cfg.ns)
    t0 = time.perf_counter()
    for _ in range(steps):
        for s in range(cfg.ns):
            hp.move_es(s)
            hp.material_absorption(s)
        for s in range(cfg.ns):
            hp.p2ces(s, copy=False)
        hp.field_source_sum(copy=False)
        hp.mcc()
    hp.synchronize()
    t1 = time.perf_counter()
    n1 = live_total(hp, cfg.ns)
    return t1 - t0, 0.5 * (n0 + n1) * steps


_SHARE = None   # (cfg, geom, tables, particles, ...) for forked oracle workers


def _oracle_run(cfg, geom, tables, sample, ez, er, seconds, max_steps):
    """Advance `sample` with the oracle for `seconds` (or max_steps)."""
    from oracle import ipm_oracle as O
    from ipm2d import workload as W
    species = [W.species_params(cfg, s) for s in range(cfg.ns)]
    setup = O.MccSetup(cfg.dx, cfg.dt, [q for q, _ in species], [m for _, m in species], tables)
    st = O.StepState(geom, setup, sample, ez, er, extra=1.0, fast=True)
    psteps, nsteps = 0, 0
    t0 = time.perf_counter()
    while True:
        n = int(st.n.sum())
        st.step(cfg.seed, nsteps)
        psteps += n
        nsteps += 1
        el = time.perf_counter() - t0
        if el >= seconds or nsteps >= max_steps:
            return psteps, nsteps, el


def _oracle_worker(w):
    cfg, geom, tables, particles, ez, er, seconds, workers = _SHARE
    return _oracle_run(cfg, geom, tables, [P[:, w::workers].copy() for P in particles], ez, er, seconds, 1000)


def cpu_baseline(cfg, geom, tables, particles, seconds, workers):
    """The oracle (statement-by-statement restatement of the reference
    routines), compiled -O3 -march=native on this host, timed on the host:
    one core on a ~2M-particle sample, then `workers` processes each owning
    an interleaved 1/workers share of the whole workload (the reference's
    MPI ranks each own their particles; the per-step grid reduction between
    ranks is not included).  Runs before any GPU initialisation, so the
    workers can be forked."""
    global _SHARE
    import multiprocessing as mp
    from oracle import ipm_oracle as O
    from ipm2d import workload as W
    native = native_oracle()
    if native:
        O.use_timed_library(native)
    build = "-O3 -march=native oracle" if native else "-O3 -march=x86-64-v2 oracle"
    total = sum(P.shape[1] for P in particles)
    ez, er = W.fields(cfg, 0.0)
    stride = max(1, total // 2_000_000)
    sample = [P[:, ::stride].copy() for P in particles]
    ps1, n1, el1 = _oracle_run(cfg, geom, tables, sample, ez, er, seconds, 100)
    out = dict(value=ps1 / el1, unit="particle-steps/s", cores=1, kind="port",
               sample=f"every {stride}th particle of the same workload ({sum(P.shape[1] for P in sample)} "
                      f"particles), {n1} steps of push+absorb+deposit+MCC in {el1:.1f} s, {build}, one core")
    if workers > 1:
        _SHARE = (cfg, geom, tables, particles, ez, er, seconds, workers)
        with mp.get_context("fork").Pool(workers) as pool:
            res = pool.map(_oracle_worker, range(workers))
        _SHARE = None
        ps = sum(r[0] for r in res)
        el = max(r[2] for r in res)
        out = dict(value=ps / el, unit="particle-steps/s", cores=workers, kind="port",
                   sample=f"the whole workload ({total} particles) split over {workers} processes, "
                          f"{min(r[1] for r in res)}-{max(r[1] for r in res)} steps each of "
                          f"push+absorb+deposit+MCC in {el:.1f} s, {build}",
                   single_core_value=ps1 / el1, single_core_sample=out["sample"])
    return out


def adjoint_leg(reps=5):
    """The differentiable window on its own workload (129x257 nodes, ~4M
    particles, 200 collisionless steps, checkpoint every 20): HIP-event time
    of the forward replay and of the adjoint, and the adjoint gradient against
    central differences (h = 1e-4 relative) for RF amplitude, phase, Cm1, Cm2.
    Two losses: the ion flux to the powered electrode (4 node planes) and the
    same flux weighted over the whole domain (every ion contributes)."""
    import numpy as np
    import ipm2d
    from ipm2d import circuit as ck
    from ipm2d import workload as W
    from ipm2d.matching import FluxGradient, electrode_capacitance
    cfg = W.make_config("c4")
    geom = W.make_geometry(cfg)
    rng = np.random.default_rng(cfg.seed)
    parts = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
    hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, particles=parts, capacity_factor=1.2)
    imn = ck.IMN(Cm1=500e-12, Cm2=200e-12, Rs=50.0, amplitude=200.0, phase_deg=30.0,
                 eqc=electrode_capacitance(cfg))
    out = {}
    for name, layers in (("electrode_flux", 4), ("whole_domain", cfg.nz - 2)):
        fg = FluxGradient(cfg, hp, geom, parts, imn, T=cfg.steps, interval=20, layers=layers)
        _, g = fg.gradient()
        fd = fg.central_fd(rel=1e-4)
        w, c = fg.window, fg.base_volts
        f_ms, a_ms = [], []
        for _ in range(reps):
            w.forward(c)
            f_ms.append(w.info()["forward_ms"])
            w.backward(c, np.ones(cfg.steps))
            a_ms.append(w.info()["adjoint_ms"])
        fms, ams = float(np.median(f_ms)), float(np.median(a_ms))
        npart = w.info()["particles"]
        out[name] = {"forward_ms": fms, "adjoint_ms": ams, "adjoint_over_forward": ams / fms,
                     "fd_max_rel": float(((g - fd).abs() / fd.abs()).max()),
                     "forward_particle_steps_per_s": npart * cfg.steps / (fms * 1e-3),
                     "plain_equals_window": float(np.abs(fg.base_loss - fg.record_loss).max() /
                                                  np.abs(fg.base_loss).max())}
        log(f"[bench] adjoint {name}: forward {fms:.3f} ms, adjoint {ams:.3f} ms, "
            f"ratio {ams / fms:.2f}, fd rel {out[name]['fd_max_rel']:.2e}")
        fg.close()
    hp.close()
    out["workload"] = (f"c4: {cfg.nz}x{cfg.nr} nodes, {npart} particles (e/CF3+/F-), {cfg.steps} steps, "
                       f"checkpoint every 20, IMN Rs 50 Ohm Cm1 500 pF Cm2 200 pF, 200 V 13.56 MHz phase 30 deg")
    out["model"] = ("reduced window model: the plasma part of the field and the absorption events are frozen "
                    "over the window, the electrode capacitance is the vacuum-plate value with q0 = 0, and there "
                    "is no adjoint through the Poisson solve; the gradients are those of this model")
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2")
    ap.add_argument("--share", type=int, default=1,
                    help="K > 1: one GPU's share of a K-GPU run (every species' per-cell population / K, full grid)")
    ap.add_argument("--sort-interval", type=int, default=-1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-workers", type=int, default=0, help="0: min(16, cores)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--device-load", action="store_true",
                    help="make the particles on the GPU (automatic above 2.5e8 particles)")
    ap.add_argument("--no-extra", action="store_true", help="skip the plain-launch comparison")
    ap.add_argument("--eager", action="store_true",
                    help="report the kernel-by-kernel launches instead of the captured graph replay")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-window leg")
    ap.add_argument("--blocks-per-cu", type=int, default=-1, help="-1: autotuned; 0: one block per tile")
    ap.add_argument("--no-autotune", action="store_true", help="keep the config's launch shape and sort interval")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="for rocprofv3 --pmc runs: a 1 GiB device copy (counter calibration), then 2 warmup and 4 timed steps")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for {args.gpus}); the multi-GPU path is a design "
              f"(particles sharded per species, one all-reduce of Rho and Chi per step) and is not built here.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import ipm2d
    from ipm2d import workload as W

    if args.traffic_probe:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        a.fill_(1.0)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)                      # reads and writes exactly 2**30 bytes
        torch.cuda.synchronize()
        del a, b
        args.steps, args.warmup, args.no_cpu_baseline, args.no_extra, args.no_adjoint = 4, 2, True, True, True

    cfg = W.make_config(args.config)
    if args.share > 1:
        import dataclasses
        cfg.species = [dataclasses.replace(sp, ppc=sp.ppc / args.share) for sp in cfg.species]
    sort_interval = cfg.sort_interval if args.sort_interval < 0 else args.sort_interval
    t = time.perf_counter()
    geom = W.make_geometry(cfg)
    tables = W.build_tables(cfg)
    rng = np.random.default_rng(cfg.seed)
    expected = sum(W.expected_count(cfg, geom, s) for s in range(cfg.ns))
    device_load = args.device_load or expected > 2.5e8
    if device_load:
        # too large to build on the host in reasonable time: the particles are
        # made on the GPU; the CPU baseline runs on a host-built sample of the
        # same distribution (every species' per-cell population / K)
        import dataclasses
        k = max(1, int(round(expected / 2.0e7)))
        sample_cfg = dataclasses.replace(cfg, species=[dataclasses.replace(sp, ppc=sp.ppc / k) for sp in cfg.species])
        particles = [W.load_species(sample_cfg, geom, s, rng) for s in range(cfg.ns)]
        cpu_cfg = sample_cfg
    else:
        particles = [W.load_species(cfg, geom, s, rng) for s in range(cfg.ns)]
        cpu_cfg = cfg
    log(f"[bench] {cfg.name}: {cfg.nz}x{cfg.nr} nodes, {expected:.4g} particles expected, host load "
        f"{time.perf_counter() - t:.1f}s{' (device load)' if device_load else ''}")
    cpu = None
    if not args.no_cpu_baseline:
        host = host_cpus()
        usable = host["affinity"] or host["nproc"] or 1
        if host["cgroup_cpu_quota"]:
            usable = min(usable, max(1, int(host["cgroup_cpu_quota"])))
        # one GPU's share of a shared host is 16 cores
        workers = args.cpu_workers or min(16, usable)
        log(f"[bench] cpu baseline (1 core, then {workers} processes) on {host}")
        cpu = cpu_baseline(cpu_cfg, geom, tables, particles, args.cpu_seconds, workers)
        if device_load:
            cpu["sample"] += f" (a 1/{k} per-cell sample of the workload, built on the host)"
        cpu["host"] = host
        cpu["cores_of_host"] = (f"{cpu['cores']} of {host['nproc']} logical CPUs ({host['affinity']} in this "
                                f"process's affinity, cgroup quota {host['cgroup_cpu_quota']}), {host['model']}")
        log(f"[bench] cpu {cpu['value']:.4g} particle-steps/s on {cpu['cores_of_host']}")
    if device_load:
        del particles
        t = time.perf_counter()
        # capacity for the population's growth over the tuner, warmup and
        # timed steps (ionization outpaces losses in the synthetic field)
        hp, _, _, counts = ipm2d.from_config_device(cfg, geom=geom, tables=tables, capacity_factor=1.5)
        ntot = sum(counts)
        log(f"[bench] device load of {ntot} particles in {time.perf_counter() - t:.1f}s")
    else:
        ntot = sum(P.shape[1] for P in particles)
        # capacity for the populations' growth over the whole run (F- grows
        # about 0.03 % per step in the synthetic field: attachment, no loss):
        # the tuner (~600 steps), warmup, the eager and the graph regions
        steps_total = 600 + args.warmup + 2 * args.steps
        hp, _, _, _ = ipm2d.from_config(cfg, geom=geom, tables=tables, particles=particles,
                                        capacity_factor=max(2.0, 1.0 + 0.0008 * steps_total))
        del particles
    tuned = None
    if not args.no_autotune and not args.traffic_probe and args.sort_interval < 0 and args.blocks_per_cu < 0:
        # the tuner advances the live simulation, so it also warms up
        from ipm2d import autotune
        tuned = autotune.autotune(hp, ntot)
        sort_interval = tuned["sort_interval"]
        log(f"[bench] autotuned: blocks_per_cu {tuned['blocks_per_cu']}, tile shape {tuned['tile_shape']} "
            f"{tuned['step_seconds_by_tile_shape']}, collision blocks per CU "
            f"{tuned['collide_blocks_per_cu']}, sort every {sort_interval} steps")
    elif args.blocks_per_cu > 0:
        hp.set_launch(args.blocks_per_cu)
    log(f"[bench] uploaded; warmup {args.warmup} steps")
    for k in range(args.warmup):
        hp.one_step()
    hp.synchronize()
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass

    # per-kernel launch times: the step launched kernel by kernel with HIP
    # events around each launch on the library's stream
    phase = sort_phase(args.steps, sort_interval) if not args.traffic_probe else 0
    if not args.traffic_probe:
        start_phase(hp, phase)
    hp.set_timing(True)
    el_e, psteps, n0, n1, ts_e = timed_steps(hp, cfg, args.steps, sort_interval, phase=phase)
    tm = hp.take_timing()
    hp.set_timing(False)
    log(f"[bench] eager: {args.steps} steps in {el_e:.3f}s ({ts_e:.4f}s of it sorting); kernels {tm}")
    if args.eager or args.traffic_probe:
        el, ps_g, ts_g = el_e, psteps, ts_e
    else:
        # the headline: the whole step replayed as one captured graph, the
        # capture and instantiation done before the timed region
        hp.one_step_graph()
        hp.synchronize()
        start_phase(hp, phase, graph=True)
        el, ps_g, _, _, ts_g = timed_steps(hp, cfg, args.steps, sort_interval, graph=True, phase=phase)
        log(f"[bench] graph: {args.steps} steps in {el:.3f}s ({ts_g:.4f}s of it sorting)")
    # The cell sort runs every sort_interval steps; a timed region of K steps
    # holds floor(K / interval) of them, which over- or under-weights the
    # sort unless K is a multiple of the interval.  The step time reported is
    # therefore the timed steps without their sorts plus the cost of one sort
    # after a full interval of drift, divided by the interval.
    t_sort = sort_after_full_interval(hp, sort_interval) if (sort_interval and not args.traffic_probe) else 0.0
    amort = t_sort / sort_interval if sort_interval else 0.0
    step_s = (el - ts_g) / args.steps + amort
    value = ps_g / args.steps / step_s
    eager_value = psteps / args.steps / ((el_e - ts_e) / args.steps + amort)
    log(f"[bench] sort after {sort_interval} steps of drift {t_sort * 1e3:.3f} ms; step {step_s * 1e3:.4f} ms "
        f"-> {value:.4g} particle-steps/s (eager {eager_value:.4g})")
    t_int = interval_in_graph(hp, sort_interval) if (sort_interval and not args.traffic_probe and not args.eager) \
        else None
    if t_int:
        log(f"[bench] one sort interval in graph replays: {t_int * 1e3:.4f} ms per step")

    pk_ms, pk_n = tm["particle"]
    avg_s = pk_ms / pk_n / 1e3
    nn = cfg.nz * cfg.nr
    grid_bytes = 2 * (cfg.nz + 1) * (cfg.nr + 1) * 8 + cfg.ns * nn * 8
    bytes_per_launch = 120.0 * (psteps / args.steps) + grid_bytes
    achieved = bytes_per_launch / avg_s / 1e9
    traffic = None
    # the committed PMC figure belongs to the default workload (c2, whole)
    if os.path.exists(PROFILE_TRAFFIC) and args.config == "c2" and args.share == 1:
        try:
            with open(PROFILE_TRAFFIC) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None

    # (not in the traffic probe, whose last launches the PMC figures average)
    sel_step, made_step = collision_counts(hp, tables) if not args.traffic_probe else (0.0, 0.0)
    step_bytes = bytes_per_launch + 48.0 * sel_step + 64.0 * made_step
    log(f"[bench] per step: {sel_step:.4g} selected, {made_step:.4g} created particles")

    extra = {"eager_value": eager_value, "step_form": "eager launches" if el is el_e else "captured graph replay"}
    if not args.no_extra:
        p_steps = max(5, args.steps // 10)
        el_p, ps_p = plain_steps(hp, cfg, p_steps)
        extra["plain_per_species_value"] = ps_p / el_p
        extra["batched_over_plain"] = value / extra["plain_per_species_value"]
        log(f"[bench] plain per-species launches {extra['plain_per_species_value']:.4g}")
    hp_counts = [hp.count(s) for s in range(cfg.ns)]
    capacities = list(hp.capacity)
    log(f"[bench] live particles per species {hp_counts} of capacities {capacities}")
    hp.close()

    ceiling = None
    if not args.traffic_probe:
        ceiling = stream_ceiling(min(ntot, 200_000_000))
        ceiling["frac_of_ceiling"] = achieved / ceiling["value"]
        log(f"[bench] streaming ceiling {ceiling['value']:.0f} GB/s (copy {ceiling['copy_GBps']:.0f}); "
            f"particle kernel at {ceiling['frac_of_ceiling']:.3f} of it")

    adj = None
    if not args.no_adjoint:
        adj = adjoint_leg()

    out = {
        "metric": "particle-steps/s (push+deposit+MCC) & % HBM peak; adjoint/forward time",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": step_s * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"{cfg.name}{f' (one GPU share, 1/{args.share})' if args.share > 1 else ''}: "
                        f"{cfg.nz}x{cfg.nr} nodes r-z, {ntot} particles "
                        f"({'/'.join(sp.name for sp in cfg.species)}), CF4 {cfg.pressure_mtorr} mTorr, "
                        f"dt {cfg.dt:g} s, cell sort every {sort_interval} steps",
            "particles": ntot,
            "live_particles_timed_region": [n0, n1],
            "live_particles_per_species_after": [hp_counts[s] for s in range(cfg.ns)],
            "capacity_per_species": capacities,
            "species": cfg.ns,
            "grid": [cfg.nz, cfg.nr],
            "sort_interval": sort_interval,
            "launch": ("autotuned: " if tuned else "") + (
                "one block per 2048-slot tile" if (tuned["blocks_per_cu"] if tuned else max(args.blocks_per_cu, 0)) == 0
                else f"{tuned['blocks_per_cu'] if tuned else args.blocks_per_cu} persistent blocks per CU") + (
                f", {tuned['tile_shape'][0]} threads per tile ({2048 // tuned['tile_shape'][0]} particles per "
                f"thread), {tuned['tile_shape'][1]}-row deposit window, collision pass "
                f"{tuned['collide_blocks_per_cu']} blocks per CU, {tuned['sort_tile']}-slot sort tiles" if tuned else ""),
            "autotune": ({k: v for k, v in tuned.items() if k.startswith(("step_seconds", "sort_seconds"))}
                         if tuned else None),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": "particle_kernel<push|absorb|deposit|mcc-select>",
            "kernel_avg_ms": avg_s * 1e3,
            "algorithmic_bytes_per_launch": bytes_per_launch,
            "stream_ceiling": ceiling,
        },
        "step_roofline": {
            "achieved": step_bytes / step_s / 1e9,
            "frac": step_bytes / step_s / 1e9 / HBM_PEAK_GBS,
            "unit": "GB/s",
            "algorithmic_bytes_per_step": step_bytes,
            "selected_per_step": sel_step,
            "created_per_step": made_step,
            "frac_particle_pass_bytes_only": bytes_per_launch / step_s / 1e9 / HBM_PEAK_GBS,
            "what": "the whole step (particle pass, collision pass, grid epilogue, launch gaps and the amortised "
                    "sort): its algorithmic bytes (120 B per particle-step plus the grids, 48 B per selected "
                    "particle, 64 B per created particle) over ms_per_step",
        },
        "sort": {"interval": sort_interval, "ms_after_full_interval": t_sort * 1e3,
                 "ms_per_step_amortised": amort * 1e3,
                 "sorts_in_timed_region": ((phase + args.steps) // sort_interval) if sort_interval else 0,
                 "timed_region_starts_steps_after_sort": phase,
                 "timed_region_ms_per_step": el / args.steps * 1e3,
                 "timed_region_sort_ms": ts_g * 1e3,
                 "interval_in_graph_ms_per_step": t_int * 1e3 if t_int else None,
                 "interval_in_graph_what": "one whole interval of graph steps from a sorted state, the sort "
                                           "replayed as a graph after the last (ipm_set_sort_interval): a "
                                           "cross-check of ms_per_step"},
        "kernel_ms_per_step": {k: (v[0] / v[1] if v[1] else None) for k, v in tm.items()},
        "cpu_baseline": cpu,
        "adjoint_over_forward": adj["electrode_flux"]["adjoint_over_forward"] if adj else None,
        "adjoint": adj,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
