"""Benchmark of the P-SV shot gradient on one MI355X.

A step is one shot-gradient of the viscoelastic layered workload
(2000x800 cells, dh 2 m, FDORDER 8, L=2, Q=20, NT=4500, DTINV=3): forward run
with on-device snapshots, residual against observed traces already resident
in HBM, adjoint run with the fused zero-lag correlation, chain rule.  Each
step uses the next source position of the 64-shot line.

Prints one JSON line (metric, value in shot-gradients/s, roofline of the
dominant kernel from HIP-event timing, and a CPU baseline measured with the
reference's own kernels, oracle/_ref, on a bounded sample of the same grid).
"""
import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic.json")


def measured_traffic(kernel_prefix, workload):
    """HBM bytes per launch of a kernel from the committed counter summary
    (rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes, each scaled by what the
    same counter reads for a device-to-device copy of known size)."""
    try:
        import json as _j
        d = _j.load(open(TRAFFIC_FILE))
    except (OSError, ValueError):
        return None
    for k in d.get("kernels", []):
        if k["kernel"].startswith(kernel_prefix) and k["workload"] == workload:
            return k["hbm_bytes_per_launch"]
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_workload(name):
    from ifos_amd import workloads
    if name == "layered_visc":
        return workloads.layered_visc()
    if name == "toy_elastic":
        return workloads.toy_elastic()
    if name == "batched_aperture":
        return workloads.batched_aperture()
    if name == "precision_large":
        return workloads.precision_large()
    raise SystemExit(f"unknown workload {name}")


def host_info():
    """Sockets, cores and threads of the host, from /proc/cpuinfo."""
    sockets, cores, model = set(), set(), ""
    cpus = 0
    phys = core = None
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                cpus += 1
            elif k == "physical id":
                phys = v
                sockets.add(v)
            elif k == "core id":
                core = v
                cores.add((phys, core))
            elif k == "model name" and not model:
                model = v
    except OSError:
        pass
    return {"cpu": model, "sockets": len(sockets) or None, "physical_cores": len(cores) or None,
            "logical_cpus": cpus or os.cpu_count(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS"),
            "OMP_PLACES": os.environ.get("OMP_PLACES"), "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
            "mpi_ranks": "none (one process, one thread)"}


def cpu_baseline(w, steps=300):
    """The reference's own kernels and inline loops (oracle/_ref, compiled
    from its sources) on one host core: one shot gradient of the same grid
    with all receivers -- forward run with the snapshot store, adjoint run
    with residual injection and the zero-lag correlation, chain rule -- over
    a reduced number of time steps (`steps` in all: steps / 2 forward and
    steps / 2 adjoint, rounded to the snapshot stride), scaled to the full
    NT: shot-gradients/s = 1 / (t * NT / NT_sample).  The reference has no
    threads; MPI is not used."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    cfg = w.cfg
    nt = max(cfg.DTINV, (steps // 2) // cfg.DTINV * cfg.DTINV)
    fields = {k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"}
    fields["NT"] = nt
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)], **fields)
    shot = w.shots[0]
    ntr = len(shot.recpos)
    rng = np.random.default_rng(0)
    diff_vx = (1e-3 * rng.standard_normal((ntr, nt))).astype(np.float32)
    diff_vy = (1e-3 * rng.standard_normal((ntr, nt))).astype(np.float32)
    if O.ref_available():
        kind = "reference"
        t0 = time.perf_counter()
        O.ref_gradient(p, w.start, shot.srcpos, shot.recpos, diff_vx, diff_vy)
        t = time.perf_counter() - t0
        what = "the reference's kernels and main's snapshot/correlation/chain-rule loops (oracle/_ref)"
    else:
        kind = "port"
        obs = np.zeros((ntr, nt), np.float32)
        t0 = time.perf_counter()
        O.gradient(p, w.start, shot.srcpos, shot.recpos, obs, obs)
        t = time.perf_counter() - t0
        what = "the oracle's restatement (oracle/liboracle.so)"
    rate = 1.0 / (t * cfg.NT / nt)
    return dict(value=rate, unit="shot-gradients/s", cores=1, kind=kind, host=host_info(),
                sample=f"one shot gradient of the {cfg.NX}x{cfg.NY} L={cfg.L} FDORDER={cfg.FDORDER} grid with all "
                       f"{ntr} receivers over NT={nt} (forward with snapshots every {cfg.DTINV} steps, adjoint with "
                       f"the correlation, chain rule) by {what}: {t:.1f} s, "
                       f"{2 * cfg.NX * cfg.NY * nt / t / 1e6:.2f} Mcell-updates/s, scaled to NT={cfg.NT}")


def serving(args):
    """Poisson arrivals of mixed forward-only and gradient requests on the
    request server (include/ifos2d_serve.h).  Grids, shot counts, orders,
    physics and NT are drawn per request from a seeded generator over a
    reduced version of the serving workload's ranges so the run fits a few
    minutes; observed traces are zero (the residual is then the synthetic
    trace itself, the same work).  Latency = finish - arrival."""
    import ifos_amd
    from ifos_amd import workloads
    from ifos_amd.serve import FORWARD, GRADIENT, Server
    rng = np.random.default_rng(args.seed)
    full = args.preset == "full"
    if full:
        # the serving workload's own ranges: 20 requests/s, grids 300x150 to
        # 3000x1200 with dh 1-10 m, 1-64 shots, FDORDER 4 or 8, elastic or
        # L=2, NT 1000-6000, layered models, a 288 GB device
        args.rate, args.max_nx, args.min_nt, args.max_nt, args.max_shots = 20.0, 3000, 1000, 6000, 64
        n = args.requests or 24
    else:
        n = args.requests or max(1, int(round(args.rate * args.duration)))
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n))

    def survey(k):
        nx = int(rng.integers(300, args.max_nx + 1)) // 8 * 8
        if full:
            ny = max(150, min(1200, int(nx * rng.uniform(0.4, 0.5)) // 8 * 8))
            dh = float(rng.uniform(1.0, 10.0))
            nt = int(rng.integers(args.min_nt, args.max_nt + 1)) // 4 * 4
            L = int(rng.choice([0, 2]))
            w = workloads.layered_visc(seed=1000 + k, NX=nx, NY=ny, DH=dh, NT=nt, DTINV=4, nshots=1,
                                       nrec=max(4, (nx - 60) // 5), FDORDER=int(rng.choice([4, 8])), L=L, fc=10.0)
            if L == 0:
                w.cfg.L = 0
            return w
        ny = int(nx * rng.uniform(0.4, 0.5)) // 8 * 8
        return workloads.small_case(seed=1000 + k, NX=nx, NY=ny,
                                    NT=int(rng.integers(args.min_nt, args.max_nt + 1)) // 4 * 4,
                                    FDORDER=int(rng.choice([4, 8])), L=int(rng.choice([0, 2])), VELOCITY=1, DTINV=4,
                                    FW=10)

    # --surveys K: requests come from K surveys (grid, NT, order, physics and
    # model fixed per survey), so engines and models can stay resident
    # between requests; 0: every request is a survey of its own
    surveys = [survey(k) for k in range(args.surveys)]
    reqs = []
    for k in range(n):
        w = surveys[int(rng.integers(0, len(surveys)))] if surveys else survey(k)
        nshots = int(rng.integers(1, args.max_shots + 1))
        shots = []
        for _ in range(nshots):
            sp = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8).copy()
            sp[:, 0] = rng.integers(w.cfg.FW + 3, w.cfg.NX - w.cfg.FW - 3)
            shots.append((sp, w.shots[0].recpos))
        kind = GRADIENT if rng.random() < 0.5 else FORWARD
        ntr = sum(len(r) for _, r in shots)
        obs = (np.zeros((ntr, w.cfg.NT), np.float32),) * 2 if kind == GRADIENT else None
        reqs.append((arrivals[k], w, kind, shots, obs))
    log(f"[serve] {n} requests over {arrivals[-1]:.1f} s")
    srv = Server(hbm_budget=int(args.budget_gb * (1 << 30)), workers=args.workers, max_slots=args.max_slots)
    t0 = time.perf_counter()
    ids = []
    for at, w, kind, shots, obs in reqs:
        while time.perf_counter() - t0 < at:
            time.sleep(min(0.002, at - (time.perf_counter() - t0)))
        ids.append(srv.submit(kind, w.cfg, w.start, shots, obs=obs, snap=args.snap, model_id=id(w)))
    done = []
    for k, (rid, (at, w, kind, shots, obs)) in enumerate(zip(ids, reqs)):
        while not srv.wait(rid, timeout=30.0):
            log(f"[serve] waiting for request {k + 1}/{n} ({time.perf_counter() - t0:.0f} s)")
        r = srv.result(rid)
        done.append((kind, len(shots), r["times"], 2.0 * w.cfg.NX * w.cfg.NY * w.cfg.NT * len(shots), r["shared"]))
        log(f"[serve] request {k + 1}/{n} done: {'gradient' if kind == GRADIENT else 'forward'}, "
            f"{w.cfg.NX}x{w.cfg.NY} NT {w.cfg.NT}, {len(shots)} shots, latency {r['times'][2] - r['times'][0]:.2f} s")
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    lat = np.array([t[2] - t[0] for _, _, t, _, _ in done])
    wait = np.array([t[1] - t[0] for _, _, t, _, _ in done])
    gshots = sum(ns for kind, ns, _, _, _ in done if kind == GRADIENT)
    fshots = sum(ns for kind, ns, _, _, _ in done if kind == FORWARD)
    cells = sum(c for kind, _, _, c, _ in done if kind == GRADIENT) + \
        sum(c / 2 for kind, _, _, c, _ in done if kind == FORWARD)
    out = {"metric": "shot-gradients/s & Gcell-updates/s under Poisson arrivals (serving)",
           "value": round(gshots / wall, 4), "unit": "shot-gradients/s", "n_gpus": 1, "higher_is_better": True,
           "data": "synthetic (seeded layered/anomaly models, zero observed traces)",
           "config": {"workload": "serving" + ("_full_ranges" if full else ""), "requests": n,
                      "rate_per_s": args.rate, "arrival_span_s": round(float(arrivals[-1]), 3),
                      "surveys": args.surveys or n,
                      "grid_nx": [300, args.max_nx], "nt": [args.min_nt, args.max_nt], "shots": [1, args.max_shots],
                      "workers": args.workers, "max_slots": args.max_slots, "budget_gb": args.budget_gb,
                      "snapshots": args.snap},
           "latency_s": {"p50": round(float(np.percentile(lat, 50)), 4), "p95": round(float(np.percentile(lat, 95)), 4),
                         "p99": round(float(np.percentile(lat, 99)), 4), "max": round(float(lat.max()), 4)},
           "queue_wait_s": {"p50": round(float(np.percentile(wait, 50)), 4),
                            "p99": round(float(np.percentile(wait, 99)), 4)},
           "forward_shots_per_s": round(fshots / wall, 4), "gcell_updates_per_s": round(cells / wall / 1e9, 3),
           "wall_s": round(wall, 2), "requests_in_shared_launches": sum(1 for d in done if d[4] > 0),
           "server": st}
    print(json.dumps(out))
    return 0


def dropin(args):
    """One shot gradient of the workload through the drop-in layer
    (include/ifos2d_dropin.h), made the way the reference's shot loop makes
    its calls (ifos_amd/dropin.py): forward with the snapshot store, the
    residual on the host as the reference's calc_res forms it (LNORM 2,
    VELOCITY 1: the time-reversed difference), adjoint with the zero-lag
    correlation, chain rule.  Observed traces come from an untimed forward
    run on the true model through the same layer."""
    import ifos_amd
    from ifos_amd.dropin import DropinShot
    w = build_workload(args.workload)
    cfg = w.cfg
    if cfg.VELOCITY != 1:
        raise SystemExit("--dropin times VELOCITY 1 workloads (the residual is formed on the host here)")
    lib = ifos_amd.lib()
    shot = w.shots[0]
    src = np.array(shot.srcpos, np.float32).reshape(-1, 8)
    rec = np.array(shot.recpos, np.int32).reshape(-1, 2)
    log(f"[dropin] workload {w.name}: {cfg.NX}x{cfg.NY} L={cfg.L} FDORDER={cfg.FDORDER} NT={cfg.NT}")
    run = DropinShot(lib, cfg, w.true)
    ovx, ovy = run.forward(src, rec, snapshots=False)
    run.close()
    ovx[:, 0] = 0.0
    ovy[:, 0] = 0.0
    log("[dropin] observed data done")
    run = DropinShot(lib, cfg, w.start)
    t0 = time.perf_counter()
    vx, vy = run.forward(src, rec, snapshots=True)
    t1 = time.perf_counter()
    log(f"[dropin] forward with snapshots {t1 - t0:.3f} s")
    dvx = np.ascontiguousarray((vx - ovx)[:, ::-1])
    dvy = np.ascontiguousarray((vy - ovy)[:, ::-1])
    run.adjoint_gradient(rec, dvx, dvy)
    t2 = time.perf_counter()
    log(f"[dropin] adjoint and gradient {t2 - t1:.3f} s")
    run.close()
    steps = 2 * cfg.NT
    out = {"metric": "shot-gradients/s through the drop-in layer (reference call sequence)",
           "value": round(1.0 / (t2 - t0), 4), "unit": "shot-gradients/s", "n_gpus": 1, "steps": 1, "warmup": 0,
           "ms_per_step": round((t2 - t0) * 1e3, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded layered model with faults; observed traces from the true model)",
           "config": {"workload": w.name, "grid": [cfg.NX, cfg.NY], "FDORDER": cfg.FDORDER, "L": cfg.L,
                      "NT": cfg.NT, "DTINV": cfg.DTINV, "path": "drop-in entry points driven from Python ctypes, "
                      "one call per reference function and time step"},
           "forward_s": round(t1 - t0, 3), "adjoint_gradient_s": round(t2 - t1, 3),
           "us_per_time_step": round((t2 - t0) / steps * 1e6, 2)}
    print(json.dumps(out))
    return 0


def fwi_iteration(args):
    """One FWI iteration of the workload (ifos_amd/fwi.py, in the order of
    src/IFOS2D.c:1076-3876): gradient over every shot, PCG direction on the
    device, the step-length search with its test forwards and the parabola,
    the bounded model update.  Observed traces come from untimed forward
    runs on the true model.  Each phase ends with a value copied back to the
    host, so the phase times are wall-clock times of finished device work."""
    from ifos_amd.fwi import FWI, FWISettings
    w = build_workload(args.workload)
    cfg = w.cfg
    s = FWISettings(TESTSHOT_START=1, TESTSHOT_END=1)
    f = FWI(w, s, slots=max(1, args.batch), snap=args.snap)
    f.observe()
    log(f"[fwi] {w.name}: {cfg.NX}x{cfg.NY} NT={cfg.NT}, {len(w.shots)} shots, observed data done")
    times = []
    for it in range(1, args.warmup + args.steps + 1):
        t0 = time.perf_counter()
        grads, L2_1 = f.gradient(f.model)
        t1 = time.perf_counter()
        f.direction(grads, it)
        t2 = time.perf_counter()
        opteps, L2t, epst, forwards = f.step_search(L2_1, it)
        t3 = time.perf_counter()
        f.model, _ = f.trial(opteps, it, 0)
        t4 = time.perf_counter()
        log(f"[fwi] iteration {it}: gradient {t1 - t0:.3f} s, direction {1e3 * (t2 - t1):.2f} ms, step search "
            f"{t3 - t2:.3f} s ({forwards} test forwards, step {opteps:.4g}), update {1e3 * (t4 - t3):.2f} ms, "
            f"misfit {L2_1:.6e}")
        if it > args.warmup:
            times.append((t0, t1, t2, t3, t4, forwards, L2_1, opteps))
    f.close()
    tot = sum(t[4] - t[0] for t in times)
    ph = lambda a, b: round(1e3 * sum(t[b] - t[a] for t in times) / len(times), 3)
    out = {"metric": "FWI iterations/s (gradient, PCG direction, step search, update)",
           "value": round(len(times) / tot, 4), "unit": "iterations/s", "n_gpus": 1, "steps": len(times),
           "warmup": args.warmup, "ms_per_step": round(1e3 * tot / len(times), 2), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "fp32",
           "data": "synthetic (seeded model; observed traces from the true model)",
           "config": {"workload": w.name, "grid": [cfg.NX, cfg.NY], "FDORDER": cfg.FDORDER, "L": cfg.L, "NT": cfg.NT,
                      "shots": len(w.shots), "snapshots": args.snap, "gradient": "PCG", "test_shots": 1},
           "phase_ms": {"gradient": ph(0, 1), "direction": ph(1, 2), "step_search": ph(2, 3), "update": ph(3, 4)},
           "test_forwards": [t[5] for t in times], "misfit": [t[6] for t in times], "step": [t[7] for t in times]}
    print(json.dumps(out))
    return 0


def post_kernels(args):
    """Device times of the gradient post-processing and trace filtering on
    the workload's sizes: median smoothing (GRAD_FILTER, 5x5 window),
    spatial filter (SPATFILTER, half-width 3), whole-gradient taper,
    Butterworth time filter of one shot's traces (order 4, 10 Hz lowpass).
    Each figure is the wall time of `reps` calls ended by a device read."""
    import ifos_amd
    from ifos_amd import opt
    w = build_workload(args.workload)
    cfg = w.cfg
    NX, NY, reps = cfg.NX, cfg.NY, 20
    ntr = len(w.shots[0].recpos)
    rng = np.random.default_rng(1)
    g = rng.standard_normal((NY, NX)).astype(np.float32)
    o = opt.Optimizer()
    buf = ifos_amd.DeviceBuffer.from_array(g)
    coef, zero = opt.taper_coeff(opt.TAPER_HORIZONTAL, NX, NY, gradt=(10, 40, 0, 0))
    tr = ifos_amd.DeviceBuffer.from_array(rng.standard_normal((ntr, cfg.NT)).astype(np.float32))
    out = {}

    def timed(name, fn, nbytes):
        fn()
        o.max(buf, NX * NY)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        o.max(buf, NX * NY)
        ms = (time.perf_counter() - t0) * 1e3 / reps
        out[name] = {"ms": round(ms, 4), "GB_s": round(nbytes / (ms * 1e-3) / 1e9, 1)}
        log(f"[post] {name}: {ms:.4f} ms per call")

    cell = 4 * NX * NY
    timed("smooth_5x5", lambda: o.smooth(NX, NY, 5, buf), 2 * cell)
    timed("spat_filt_3", lambda: o.spat_filt(NX, NY, 3, 1, buf), 2 * cell)
    timed("taper", lambda: o.taper(buf, coef, zero), 3 * cell)
    timed("time_filter", lambda: ifos_amd.time_filter(tr, ntr, cfg.NT, cfg.DT, 10.0, 4), 2 * 4 * ntr * cfg.NT)
    buf.free()
    tr.free()
    o.close()
    res = {"metric": "gradient post-processing and trace filter times", "value": out["smooth_5x5"]["ms"],
           "unit": "ms per smoothing", "n_gpus": 1, "higher_is_better": False, "dtype": "fp32",
           "data": "synthetic (random gradient and traces)",
           "config": {"workload": w.name, "grid": [NX, NY], "traces": [ntr, cfg.NT], "reps": reps}, "kernels": out}
    print(json.dumps(res))
    return 0


def precision_errors(args):
    """Errors of the 16-bit snapshot runs against the FP64 build of the
    scheme (ifos_fp64_*, forward recomputed from checkpoints) at the
    workload's full size.  Observed traces come from FP32 forward runs on the
    true model (untimed); the FP64, BF16 and FP16 runs then compute the
    gradient of the same shots on the start model.  Reported per parameter:
    relative L2 error of the summed gradient, and the largest relative L2
    error of any shot's seismograms (the FP32 forward against the FP64
    forward)."""
    import ifos_amd
    w = build_workload(args.workload)
    cfg = w.cfg
    shots = w.shots[:args.shots] if args.shots else w.shots
    log(f"[precision] workload {w.name}: {cfg.NX}x{cfg.NY} L={cfg.L} FDORDER={cfg.FDORDER} NT={cfg.NT} "
        f"DTINV={cfg.DTINV}, {len(shots)} shots")
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.true)
    obs = []
    for k, s in enumerate(shots):
        eng.set_shot(s.srcpos, s.recpos)
        eng.forward(snapshots=False)
        vx, vy = eng.seismograms()
        vx[:, 0] = 0.0
        vy[:, 0] = 0.0
        obs.append((vx, vy))
        log(f"[precision] observed data {k + 1}/{len(shots)}")
    eng.close()

    def kill_of(s):
        return s.kill if (cfg.TRKILL and s.kill is not None) else None

    e64 = ifos_amd.Fp64Engine(cfg, segment_steps=args.segment)
    e64.set_model(**w.start)
    seis64 = []
    t0 = time.perf_counter()
    for k, s in enumerate(shots):
        e64.set_shot(s.srcpos, s.recpos)
        e64.shot_gradient(obs[k][0], obs[k][1], kill_of(s))
        seis64.append(e64.seismograms())
        log(f"[precision] fp64 shot {k + 1}/{len(shots)} ({time.perf_counter() - t0:.1f} s)")
    t64 = time.perf_counter() - t0
    g64 = e64.gradient(total=True)
    seg = e64.segment_steps
    e64.close()

    def rel(a, b):
        return float(np.linalg.norm((a.astype(np.float64) - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))

    res = {}
    for prec in ("bf16", "fp16"):
        eng = ifos_amd.Engine(cfg)
        eng.set_model(**w.start)
        eng.set_snapshot_precision(prec)
        eng.reset_gradient()
        seis_err = 0.0
        t0 = time.perf_counter()
        for k, s in enumerate(shots):
            eng.set_shot(s.srcpos, s.recpos)
            eng.shot_gradient(obs[k][0], obs[k][1], kill_of(s))
            vx, vy = eng.seismograms()
            seis_err = max(seis_err, rel(vx, seis64[k][0]), rel(vy, seis64[k][1]))
            log(f"[precision] {prec} shot {k + 1}/{len(shots)} ({time.perf_counter() - t0:.1f} s)")
        t = time.perf_counter() - t0
        g = eng.gradient(total=True)
        eng.close()
        res[prec] = {"grad_rel_l2": {n: rel(g[i], g64[i]) for i, n in enumerate(("vp", "vs", "rho"))},
                     "seismogram_rel_l2_max": seis_err, "shot_gradients_per_s": round(len(shots) / t, 4)}
        log(f"[precision] {prec}: {res[prec]}")
    out = {"metric": "relative L2 error vs the FP64 build of the scheme (16-bit snapshots, FP32 stencils)",
           "value": max(res["bf16"]["grad_rel_l2"].values()), "unit": "max relative L2 gradient error (bf16)",
           "n_gpus": 1, "higher_is_better": False, "dtype": "fp32 stencils, 16-bit snapshots; yardstick fp64",
           "data": "synthetic (seeded layered model with faults; observed traces from FP32 forward runs on the "
                   "true model)",
           "config": {"workload": w.name, "grid": [cfg.NX, cfg.NY], "FDORDER": cfg.FDORDER, "L": cfg.L,
                      "NT": cfg.NT, "DTINV": cfg.DTINV, "shots": len(shots), "fp64_segment_steps": seg},
           "errors": res, "fp64_shot_gradients_per_s": round(len(shots) / t64, 4)}
    print(json.dumps(out))
    return 0


def batched_vs_plain(args):
    """The same shots' gradients on the one-shot-at-a-time FP32 path and on
    the batched path (args.batch shots per launch, args.snap snapshots),
    each autotuned, in one JSON
    line with their shot-gradients/s, the ratio, the algorithmic bytes per
    shot-gradient of each, and the batched gradient's error against the
    FP32 path's.  Observed traces come from untimed forward runs on the true
    model."""
    import ifos_amd
    w = build_workload(args.workload)
    cfg = w.cfg
    # shots per launch: as given, or the MALL-sized count the library suggests
    S = args.batch if args.batch >= 2 else max(2, ifos_amd.mall_slots(cfg))
    n = S * max(1, args.steps)
    shots = [(s.srcpos, s.recpos) for s in w.shots[:n]]
    eng = ifos_amd.Engine(cfg, slots=S)
    eng.set_model(**w.true)
    obs = []
    for b in range(0, n, S):
        eng.set_shots(shots[b:b + S])
        eng.forward()
        vx, vy = eng.seismograms()
        vx[:, 0] = 0.0
        vy[:, 0] = 0.0
        obs.append((ifos_amd.DeviceBuffer.from_array(vx), ifos_amd.DeviceBuffer.from_array(vy)))
    eng.close()
    log(f"[batched] {w.name}: {n} shots, observed data done")
    cells = cfg.NX * cfg.NY
    nsnap = cfg.NT // cfg.DTINV

    def algorithmic_bytes(e, snap16):
        step = e.kernel_bytes_per_cell(0) + e.kernel_bytes_per_cell(1)
        store = 10.0 if snap16 else 20.0
        # correlation step: the snapshots, vxp1/vyp1 written and read, the
        # three sums read and written, the three moduli read
        corr = (10.0 if snap16 else 20.0) + 8 + 8 + 24 + 12
        return cells * (2.0 * cfg.NT * step + nsnap * (store + corr))

    # plain: one shot per launch, FP32 snapshots, plain sums
    one = ifos_amd.Engine(cfg)
    one.set_model(**w.start)
    one.set_shot(*shots[0])
    one.autotune(reps=5)
    ntr = [len(r) for _, r in shots]
    hosts = [(o[0].to_array((sum(ntr[b:b + S]), cfg.NT)), o[1].to_array((sum(ntr[b:b + S]), cfg.NT)))
             for b, o in zip(range(0, n, S), obs)]
    one.reset_gradient()
    t0 = time.perf_counter()
    for b in range(0, n, S):
        t = 0
        ovx, ovy = hosts[b // S]
        for k in range(b, min(b + S, n)):
            one.set_shot(*shots[k])
            one.shot_gradient(ovx[t:t + ntr[k]], ovy[t:t + ntr[k]])
            t += ntr[k]
        log(f"[batched] plain path: {min(b + S, n)}/{n} shots")
    t_plain = time.perf_counter() - t0
    g_plain = one.gradient(total=True)
    bytes_plain = algorithmic_bytes(one, False)
    one.close()
    # batched
    bat = ifos_amd.Engine(cfg, slots=S)
    bat.set_model(**w.start)
    bat.set_snapshot_precision(args.snap)
    bat.set_shots(shots[:S])
    spb, grid, ms = bat.autotune(reps=5)
    bat.reset_gradient()
    t0 = time.perf_counter()
    for b in range(0, n, S):
        bat.set_shots(shots[b:b + S])
        bat.shot_gradient(obs[b // S][0], obs[b // S][1])
        log(f"[batched] batched path: {min(b + S, n)}/{n} shots")
    t_bat = time.perf_counter() - t0
    g_bat = bat.gradient(total=True)
    bytes_bat = algorithmic_bytes(bat, args.snap != "fp32")
    shape = dict(spb=spb, grids=list(bat.get_launch_grids()), slot_major=bat.get_work_order(),
                 velocity_rows=bat.get_velocity_rows(), one_kernel_step=bat.get_fused())
    bat.close()
    rate_p, rate_b = n / t_plain, n / t_bat
    rel = [float(np.linalg.norm((g_bat[k].astype(np.float64) - g_plain[k]).ravel()) /
                 np.linalg.norm(g_plain[k].ravel().astype(np.float64))) for k in range(3)]
    out = {"metric": "shot-gradients/s, batched vs one shot at a time", "value": round(rate_b, 4),
           "unit": "shot-gradients/s", "n_gpus": 1, "steps": args.steps, "higher_is_better": True,
           "dtype": f"fp32 stencils, {args.snap} snapshots", "data": "synthetic (seeded layered model with faults)",
           "config": {"workload": w.name, "grid": [cfg.NX, cfg.NY], "NT": cfg.NT, "DTINV": cfg.DTINV, "shots": n,
                      "batch": S, "batch_rule": "given" if args.batch >= 2 else "MALL-sized (ifos_engine_mall_slots)",
                      "snapshots": args.snap, "launch": shape},
           "plain": {"shot_gradients_per_s": round(rate_p, 4), "algorithmic_bytes_per_shot_gradient": bytes_plain,
                     "GB_s": round(bytes_plain * rate_p / 1e9, 1)},
           "batched": {"shot_gradients_per_s": round(rate_b, 4), "algorithmic_bytes_per_shot_gradient": bytes_bat,
                       "GB_s": round(bytes_bat * rate_b / 1e9, 1)},
           "speedup": round(rate_b / rate_p, 3), "batched_vs_plain_rel_l2": {"vp": rel[0], "vs": rel[1], "rho": rel[2]}}
    print(json.dumps(out))
    for o in obs:
        o[0].free()
        o[1].free()
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="layered_visc")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=300)
    ap.add_argument("--batch", type=int, default=1,
                    help="shots advanced together per kernel launch (1: the one-shot-at-a-time path; with "
                         "--batched-vs-plain, 1 picks the MALL-sized count of ifos_engine_mall_slots)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shape (no timing sweep)")
    ap.add_argument("--no-several", action="store_true",
                    help="skip the companion run with several shots per launch (max(2, MALL-sized), at most 8)")
    ap.add_argument("--order", default="auto", choices=["auto", "tile", "slot"],
                    help="work order of several-shot launches (auto: the autotuner's pick)")
    ap.add_argument("--vel-rows", type=int, default=0, choices=[0, 1, 2, 4],
                    help="velocity step cells per thread along y (0: the autotuner's pick, or the engine's "
                         "default without it)")
    ap.add_argument("--vel-lds", action="store_true",
                    help="velocity step with two rows per thread and the tile's stresses staged in LDS")
    ap.add_argument("--snap", default="fp32", choices=["fp32", "bf16", "fp16"],
                    help="snapshot storage (fp32 keeps the gradient bit-identical to the reference)")
    ap.add_argument("--copy-floor", action="store_true",
                    help="with --kernels-only: also time a device copy moving each kernel's algorithmic bytes")
    ap.add_argument("--kernels-only", action="store_true",
                    help="only time the velocity and stress kernels on the live model (for counter runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MB",
                    help="also run one device-to-device copy of this many MB (counter calibration)")
    ap.add_argument("--serve", action="store_true", help="serving trace (Poisson arrivals) on the request server")
    ap.add_argument("--post", action="store_true",
                    help="time the gradient post-processing and trace-filter kernels on the workload's sizes")
    ap.add_argument("--fwi", action="store_true",
                    help="time whole FWI iterations of the workload (gradient, PCG, step search, update)")
    ap.add_argument("--dropin", action="store_true",
                    help="one shot gradient through the drop-in layer (the reference's call sequence)")
    ap.add_argument("--batched-vs-plain", action="store_true",
                    help="the same shots on the plain FP32 path and on the batched path (--batch, --snap)")
    ap.add_argument("--precision-errors", action="store_true",
                    help="BF16 and FP16 snapshot errors against the FP64 build at the workload's full size")
    ap.add_argument("--shots", type=int, default=0, help="precision errors: first N shots (0: all)")
    ap.add_argument("--segment", type=int, default=0,
                    help="precision errors: FP64 checkpoint segment in steps (0: DTINV*ceil(sqrt(NT/DTINV)))")
    ap.add_argument("--preset", default="", choices=["", "full"],
                    help="serving: 'full' draws from the whole serving workload (20/s, 300x150-3000x1200, 1-64 "
                         "shots, NT 1000-6000)")
    ap.add_argument("--requests", type=int, default=0, help="serving: number of requests (0: rate x duration)")
    ap.add_argument("--rate", type=float, default=4.0, help="serving: requests per second")
    ap.add_argument("--duration", type=float, default=15.0, help="serving: seconds of arrivals")
    ap.add_argument("--max-nx", type=int, default=1200)
    ap.add_argument("--min-nt", type=int, default=600)
    ap.add_argument("--max-nt", type=int, default=2000)
    ap.add_argument("--max-shots", type=int, default=8)
    ap.add_argument("--workers", type=int, default=4)
    ap.add_argument("--max-slots", type=int, default=4)
    ap.add_argument("--surveys", type=int, default=8,
                    help="serving: distinct survey configurations the requests draw from (0: one per request)")
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--seed", type=int, default=7)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only: --gpus {args.gpus} is not supported (every box here has one GPU; "
              f"the multi-GPU plan is shot sharding with one gradient all-reduce per iteration)", file=sys.stderr)
        return 2

    if args.serve:
        return serving(args)
    if args.dropin:
        return dropin(args)
    if args.fwi:
        return fwi_iteration(args)
    if args.post:
        return post_kernels(args)
    if args.precision_errors:
        return precision_errors(args)
    if args.batched_vs_plain:
        return batched_vs_plain(args)
    import ifos_amd
    w = build_workload(args.workload)
    cfg = w.cfg
    S = max(1, args.batch)
    nbatches = args.warmup + args.steps
    shots = [w.shots[k % len(w.shots)] for k in range(nbatches * S)]
    batches = [shots[b * S:(b + 1) * S] for b in range(nbatches)]
    log(f"[bench] workload {w.name}: {cfg.NX}x{cfg.NY} L={cfg.L} FDORDER={cfg.FDORDER} NT={cfg.NT} "
        f"DTINV={cfg.DTINV}, {len(shots)} shots in batches of {S}, {args.snap} snapshots")

    eng = ifos_amd.Engine(cfg, slots=S)
    if args.vel_rows:
        eng.set_velocity_rows(args.vel_rows)
    if args.vel_lds:
        eng.set_velocity_lds(True)
    if args.copy_probe:
        nb = args.copy_probe * (1 << 20)
        a, c = ifos_amd.DeviceBuffer(nb), ifos_amd.DeviceBuffer(nb)
        import ctypes as C
        ifos_amd._check(ifos_amd.lib().ifos_memcpy_d2d(C.c_void_p(c.ptr), C.c_void_p(a.ptr), nb))
        log(f"[bench] copy probe: {args.copy_probe} MB device-to-device")
        a.free()
        c.free()
    if args.kernels_only:
        eng.set_model(**w.start)
        eng.set_shots([(s.srcpos, s.recpos) for s in batches[0]])
        if not args.no_autotune:
            spb, grid, ms = eng.autotune(reps=5)
            log(f"[bench] autotune: {spb} shots per workgroup, grids {eng.get_launch_grids()}, "
                f"{'slot' if eng.get_work_order() else 'tile'}-major order, {ms:.4f} ms per step")
        if args.order != "auto":
            eng.set_work_order(args.order == "slot")
        if args.vel_rows:
            eng.set_velocity_rows(args.vel_rows)
        if args.vel_lds:
            eng.set_velocity_lds(True)
        log(f"[bench] launch {eng.get_launch()}, {'slot' if eng.get_work_order() else 'tile'}-major, "
            f"velocity rows per thread {eng.get_velocity_rows()}, LDS {eng.get_velocity_lds()}")
        for kern in (0, 1):
            ms = eng.time_kernel(kern, reps=20)
            nbytes = eng.kernel_bytes_per_cell(kern) * cfg.NX * cfg.NY * S
            gbs = nbytes / (ms * 1e-3) / 1e9
            log(f"[bench] kernel {kern}: {ms:.5f} ms/launch, {gbs:.1f} GB/s algorithmic")
            if args.copy_floor:
                mc = eng.time_copy(nbytes)
                log(f"[bench] copy of the same {nbytes / 1e6:.1f} MB (half read, half written), back to back: "
                    f"{mc:.5f} ms/launch, {nbytes / (mc * 1e-3) / 1e9:.1f} GB/s; kernel / copy {ms / mc:.3f}")
        if args.copy_floor:
            log(f"[bench] copy of 4 KB, back to back: {eng.time_copy(4096):.5f} ms/launch")
        eng.close()
        return 0
    # observed data: forward runs on the true model (untimed), kept in HBM
    eng.set_model(**w.true)
    obs = []
    for k, batch in enumerate(batches):
        eng.set_shots([(s.srcpos, s.recpos) for s in batch])
        eng.forward(snapshots=False)
        vx, vy = eng.seismograms()
        vx[:, 0] = 0.0
        vy[:, 0] = 0.0
        obs.append((ifos_amd.DeviceBuffer.from_array(vx), ifos_amd.DeviceBuffer.from_array(vy)))
        log(f"[bench] observed data for batch {k + 1}/{len(batches)}")
    eng.set_model(**w.start)
    eng.reset_gradient()
    eng.set_snapshot_precision(args.snap)
    shape = None
    if not args.no_autotune:
        eng.set_shots([(s.srcpos, s.recpos) for s in batches[0]])
        shape = eng.autotune(reps=5) + (eng.get_work_order(),)
        if args.order != "auto":
            eng.set_work_order(args.order == "slot")
            shape = shape[:3] + (eng.get_work_order(),)
        gv, gs = eng.get_launch_grids()
        log(f"[bench] autotune: {shape[0]} shots per workgroup, grids {gv or 'one per item'} (velocity) / "
            f"{gs or 'one per item'} (stress), {'slot' if shape[3] else 'tile'}-major order, "
            f"velocity rows per thread {eng.get_velocity_rows()} (LDS {eng.get_velocity_lds()}), "
            f"{shape[2]:.4f} ms per step")
    if args.vel_rows:
        eng.set_velocity_rows(args.vel_rows)
    if args.vel_lds:
        eng.set_velocity_lds(True)
    vel_rows = eng.get_velocity_rows()
    vel_lds = eng.get_velocity_lds()
    graphs = eng.get_graphs()

    def step(k):
        batch = batches[k]
        eng.set_shots([(s.srcpos, s.recpos) for s in batch])
        kill = np.concatenate([s.kill if s.kill is not None else np.zeros(len(s.recpos), np.int32) for s in batch])
        return eng.shot_gradient(obs[k][0], obs[k][1], kill=kill if cfg.TRKILL else None)

    for k in range(args.warmup):
        step(k)
        log(f"[bench] warmup step {k + 1} done")
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        step(k)
        log(f"[bench] step {k - args.warmup + 1}/{args.steps} done")
    elapsed = time.perf_counter() - t0
    ms_per_step = 1000.0 * elapsed / args.steps
    rate = args.steps * S / elapsed
    cell_updates = 2.0 * cfg.NX * cfg.NY * cfg.NT  # forward + adjoint per shot-gradient
    gcells = cell_updates * rate / 1e9

    # dominant kernel: the stress update; HIP-event timing on the engine stream
    kern = 1
    ms_k = eng.time_kernel(kern, reps=50)
    bpc = eng.kernel_bytes_per_cell(kern)
    bytes_launch = bpc * cfg.NX * cfg.NY * S
    achieved = bytes_launch / (ms_k * 1e-3) / 1e9
    ms_v = eng.time_kernel(0, reps=50)
    kname = "k_str_step"
    roof = dict(bound="hbm", achieved=round(achieved, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                frac=round(achieved / HBM_PEAK_GBS, 4), traffic=measured_traffic(kname, w.name) if S == 1 else None,
                kernel=kname,
                bytes_per_cell=bpc, cells=cfg.NX * cfg.NY, avg_ms=round(ms_k, 5),
                velocity_kernel_avg_ms=round(ms_v, 5),
                velocity_kernel_GBs=round(eng.kernel_bytes_per_cell(0) * cfg.NX * cfg.NY * S / (ms_v * 1e-3) / 1e9, 1))
    eng.close()

    # the same job with several shots per launch (bit-identical gradients),
    # measured in a child process of its own so that its engine starts clean
    several = None
    B = max(2, min(8, ifos_amd.mall_slots(cfg)))
    prec = {"fp32": 0, "bf16": 1, "fp16": 2}[args.snap]
    ntr, nsrc = len(w.shots[0].recpos), len(np.asarray(w.shots[0].srcpos).reshape(-1, 8))
    fits = ifos_amd.lib().ifos_engine_bytes(ctypes.byref(cfg), B, prec, 1, ntr * B, nsrc * B) < 250e9
    if S == 1 and not args.no_several and fits:
        cmd = [sys.executable, os.path.abspath(__file__), "--workload", args.workload, "--batch", str(B),
               "--steps", str(args.steps), "--warmup", str(args.warmup), "--snap", args.snap,
               "--no-cpu-baseline", "--no-several"]
        log(f"[bench] {B} shots per launch, child process")
        r = subprocess.run(cmd, stdout=subprocess.PIPE, text=True, timeout=900)
        lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
        if r.returncode == 0 and lines:
            d = json.loads(lines[-1])
            several = dict(shots_per_launch=B, value=d["value"], unit=d["unit"], ms_per_step=d["ms_per_step"],
                           note="the same shots, gradients bit-identical to one shot per launch",
                           roofline=dict(frac=d["roofline"]["frac"], achieved=d["roofline"]["achieved"],
                                         avg_ms=d["roofline"]["avg_ms"], bytes_per_launch=d["roofline"]["bytes_per_cell"]
                                         * cfg.NX * cfg.NY * B))
        else:
            several = dict(shots_per_launch=B, value=None, error=f"child exited with {r.returncode}")

    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline(w, steps=args.cpu_steps)
        except Exception as e:  # reported, not fatal
            cpu = dict(value=None, error=str(e))

    out = {
        "metric": "shot-gradients/s & Gcell-updates/s (% HBM roofline) at <=1e-4 rel. gradient err",
        "value": round(rate, 5),
        "unit": "shot-gradients/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded layered model with faults; observed traces from the true model)",
        "config": {"workload": w.name, "grid": [cfg.NX, cfg.NY], "dh": cfg.DH, "FDORDER": cfg.FDORDER, "L": cfg.L,
                   "NT": cfg.NT, "DTINV": cfg.DTINV, "snapshots": args.snap, "shots_per_step": S, "receivers": int(len(w.shots[0].recpos)),
                   "velocity_rows_per_thread": vel_rows, "velocity_lds": vel_lds,
                   "time_loop": "captured HIP graphs" if graphs else "eager launches",
                   "parallelism": "one GPU, one shot at a time" if S == 1 else
                   f"one GPU, {S} shots per launch" + (f" ({shape[0]} per workgroup, {'slot' if shape[3] else 'tile'}-major)"
                                                       if shape else "")},
        "gcell_updates_per_s": round(gcells, 3),
        "gradient_parity": ("bit-identical to the reference FP32 kernels (tests/test_gpu_parity.py)" if args.snap == "fp32"
                            else f"{args.snap} snapshots: bit-identical to the oracle with the same quantisation, "
                                 f"<=1e-3 rel. L2 vs the FP64 build (tests/test_gpu_precision.py)"),
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if several is not None:
        out["several_shots_per_launch"] = several
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
