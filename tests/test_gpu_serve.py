"""Request server (include/ifos2d_serve.h): mixed forward and gradient
requests of different grids and shot counts, submitted while others run,
give exactly the results of the same requests run directly on a one-slot
engine, shot after shot; shots of different requests of the same survey
share launches and still give those results; admission keeps the engines
inside the HBM budget; resident engines are reused."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
from ifos_amd.serve import FORWARD, GRADIENT, Server

pytestmark = pytest.mark.gpu


def request_set():
    a = workloads.small_case(seed=51, NX=110, NY=60, NT=240, FDORDER=4, L=0, VELOCITY=1, DTINV=2)
    b = workloads.small_case(seed=52, NX=90, NY=70, NT=270, FDORDER=8, L=2, VELOCITY=1, DTINV=3)
    reqs = []
    rng = np.random.default_rng(0)
    for k, (w, kind, nshots) in enumerate([(a, GRADIENT, 3), (b, FORWARD, 2), (b, GRADIENT, 5), (a, FORWARD, 1),
                                           (a, GRADIENT, 2), (b, GRADIENT, 1), (a, FORWARD, 4)]):
        shots = []
        for s in range(nshots):
            sp = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8).copy()
            sp[:, 0] = rng.integers(w.cfg.FW + 3, w.cfg.NX - w.cfg.FW - 3)
            rec = np.array(w.shots[0].recpos, np.int32)
            keep = rng.random(len(rec)) < 0.7
            keep[0] = True
            shots.append((sp, rec[keep]))
        reqs.append((w, kind, shots, "fp16" if k == 2 else "fp32"))
    return reqs


def observed(w, shots):
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.true)
    vx, vy = [], []
    for sp, rp in shots:
        eng.set_shot(sp, rp)
        eng.forward()
        a, b = eng.seismograms()
        a[:, 0] = 0.0
        b[:, 0] = 0.0
        vx.append(a)
        vy.append(b)
    eng.close()
    return np.concatenate(vx), np.concatenate(vy)


def run_all(server, reqs, obs):
    ids = []
    for (w, kind, shots, snap), o in zip(reqs, obs):
        ids.append(server.submit(kind, w.cfg, w.start, shots, obs=o if kind == GRADIENT else None, snap=snap,
                                 model_id=w.cfg.NX))
    for rid in ids:
        assert server.wait(rid, timeout=120.0)
    return [server.result(rid) for rid in ids]


def direct(w, kind, shots, snap, obs):
    """The request run on a one-slot engine, one shot after the other."""
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.start)
    out = {}
    if kind == FORWARD:
        vx, vy = [], []
        for sp, rp in shots:
            eng.set_shot(sp, rp)
            eng.forward()
            a, b = eng.seismograms()
            vx.append(a)
            vy.append(b)
        out["vx"], out["vy"] = np.concatenate(vx), np.concatenate(vy)
    else:
        eng.set_snapshot_precision(snap)
        eng.reset_gradient()
        mis, t0 = [], 0
        for sp, rp in shots:
            n = len(rp)
            eng.set_shot(sp, rp)
            eng.shot_gradient(obs[0][t0:t0 + n], obs[1][t0:t0 + n])
            mis.append(eng.shot_misfit()[0])
            t0 += n
        out["grad"] = eng.gradient(total=True)
        out["misfit"] = np.array(mis)
    eng.close()
    return out


def same(a, b):
    for k in ("vx", "vy", "grad", "misfit"):
        if k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_mixed_requests_equal_alone(gpu_lib):
    reqs = request_set()
    obs = [observed(w, shots) if kind == GRADIENT else None for w, kind, shots, _ in reqs]
    # alone: each request directly on a one-slot engine
    alone = [direct(w, kind, shots, snap, o) for (w, kind, shots, snap), o in zip(reqs, obs)]
    # in the mix: three workers, up to four shots per launch, all submitted at once
    s3 = Server(hbm_budget=8 << 30, workers=3, max_slots=4)
    mixed = run_all(s3, reqs, obs)
    st = s3.stats()
    print("[serve] stats", st, flush=True)
    s3.close()
    for k, (a, b) in enumerate(zip(alone, mixed)):
        print(f"[serve] request {k}: slots {b['slots']} times {b['times']}", flush=True)
        same(a, b)
    assert st["finished"] == len(reqs) and st["failed"] == 0
    assert max(r["slots"] for r in mixed) > 1
    # overlap: some request started before an earlier one finished
    spans = sorted((r["times"][1], r["times"][2]) for r in mixed)
    assert any(spans[i + 1][0] < spans[i][1] for i in range(len(spans) - 1))


def test_budget_admission_and_reuse(gpu_lib):
    reqs = request_set()
    obs = [observed(w, shots) if kind == GRADIENT else None for w, kind, shots, _ in reqs]
    # a budget that holds one small engine at a time forces serial admission
    s = Server(hbm_budget=600 << 20, workers=3, max_slots=1)
    out = run_all(s, reqs + reqs[:2], obs + obs[:2])
    st = s.stats()
    s.close()
    print("[serve] stats", st, flush=True)
    assert st["peak_bytes"] <= st["budget"] and st["failed"] == 0
    assert st["engines_reused"] > 0
    # a request larger than the budget is refused at submission
    w, kind, shots, _ = reqs[0]
    tiny = Server(hbm_budget=1 << 20, workers=1, max_slots=1)
    with pytest.raises(ifos_amd.IfosError):
        tiny.submit(kind, w.cfg, w.start, shots, obs=obs[0])
    tiny.close()


def test_requests_share_launches(gpu_lib):
    """A request that arrives while another runs joins the next launch of its
    survey: with one worker and four slots, the two small gradient requests
    of survey a submitted during the first launch of survey b's five-shot
    request run together in one launch, and each still equals its direct run."""
    reqs = request_set()
    big = reqs[2]                                   # survey b, gradient, 5 shots (fp16)
    a1 = (reqs[0][0], GRADIENT, reqs[0][2][:1], "fp32")
    a2 = (reqs[4][0], GRADIENT, reqs[4][2], "fp32")
    fw = (reqs[6][0], FORWARD, reqs[6][2][:3], "fp32")
    fw2 = (reqs[3][0], FORWARD, reqs[3][2], "fp32")
    order = [big, a1, a2, fw, fw2]
    obs = [observed(w, shots) if kind == GRADIENT else None for w, kind, shots, _ in order]
    want = [direct(w, kind, shots, snap, o) for (w, kind, shots, snap), o in zip(order, obs)]
    s = Server(hbm_budget=8 << 30, workers=1, max_slots=4)
    got = run_all(s, order, obs)
    st = s.stats()
    s.close()
    print("[serve] stats", st, [(r["slots"], r["shared"]) for r in got], flush=True)
    for k, (a, b) in enumerate(zip(want, got)):
        same(a, b)
    assert st["shared_launches"] >= 1
    assert got[1]["shared"] >= 1 and got[2]["shared"] >= 1      # a1 and a2 ran in one launch
    assert got[0]["shared"] == 0                                # survey b's request shares with no other


def test_large_request_in_mix(gpu_lib):
    """A 3000x1200, FDORDER 8, L=2 gradient request (2 shots, NT 600) among
    smaller forward and gradient requests submitted at once on two workers:
    every result equals its direct run on a one-slot engine."""
    big = workloads.layered_visc(seed=61, NX=3000, NY=1200, DH=4.0, NT=600, DTINV=4, nshots=2, nrec=500, fc=8.0)
    small = request_set()
    rng = np.random.default_rng(5)
    shots = []
    for s in big.shots:
        rec = big.shots[0].recpos[rng.random(len(big.shots[0].recpos)) < 0.8]
        shots.append((s.srcpos, rec))
    reqs = [small[0], (big, GRADIENT, shots, "fp32"), small[1], small[3]]
    print(f"[serve] large request {big.cfg.NX}x{big.cfg.NY} NT {big.cfg.NT} among {len(reqs) - 1} small ones",
          flush=True)
    obs = [observed(w, sh) if kind == GRADIENT else None for w, kind, sh, _ in reqs]
    want = [direct(w, kind, sh, snap, o) for (w, kind, sh, snap), o in zip(reqs, obs)]
    s = Server(hbm_budget=64 << 30, workers=2, max_slots=4)
    got = run_all(s, reqs, obs)
    st = s.stats()
    s.close()
    print("[serve] stats", st, flush=True)
    for a, b in zip(want, got):
        same(a, b)
    assert st["failed"] == 0 and np.abs(got[1]["grad"]).max() > 0
