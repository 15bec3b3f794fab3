"""Request server (include/ifos2d_serve.h): mixed forward and gradient
requests of different grids and shot counts, submitted while others run,
give exactly the results of the same requests run alone; admission keeps
the engines inside the HBM budget; resident engines are reused."""
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


def same(a, b):
    for k in ("vx", "vy", "grad", "misfit"):
        if k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_mixed_requests_equal_alone(gpu_lib):
    reqs = request_set()
    obs = [observed(w, shots) if kind == GRADIENT else None for w, kind, shots, _ in reqs]
    # alone: one worker, one shot per launch, each request waited for before the next
    alone = []
    s1 = Server(hbm_budget=8 << 30, workers=1, max_slots=1)
    for r, o in zip(reqs, obs):
        alone += run_all(s1, [r], [o])
    s1.close()
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
