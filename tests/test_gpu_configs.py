"""The benchmark workloads at their full sizes.

- 400x200 elastic case (FDORDER 4, NT 2000, DTINV 2, VELOCITY 0): small
  enough for the oracle, so seismograms and gradients are compared with it
  bit for bit.
- 2000x800 viscoelastic case (FDORDER 8, L=2, NT 4500, DTINV 3): too large
  for the oracle; checked through properties that do not depend on size:
  two shots advanced together give each shot's gradient bit for bit as run
  alone, and the batch's running sum equals the slot-order sum.
"""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O
from test_precision import params_of

pytestmark = pytest.mark.gpu


def test_toy_elastic_full_size_matches_oracle(gpu_lib):
    w = workloads.toy_elastic()
    cfg = w.cfg
    s = w.shots[0]
    print(f"[full size] toy_elastic {cfg.NX}x{cfg.NY} NT={cfg.NT}", flush=True)
    p = params_of(w)
    ovx, ovy, _ = O.forward(p, w.true, s.srcpos, s.recpos)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.true)
    eng.set_shot(s.srcpos, s.recpos)
    eng.forward()
    vx, vy = eng.seismograms()
    np.testing.assert_array_equal(vx, ovx)
    np.testing.assert_array_equal(vy, ovy)
    ovx[:, 0] = 0
    ovy[:, 0] = 0
    ref = O.gradient(p, w.start, s.srcpos, s.recpos, ovx, ovy)
    eng.set_model(**w.start)
    mis = eng.shot_gradient(ovx, ovy)
    g = eng.gradient()
    for k, n in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(g[k], ref["grad_" + n], err_msg=n)
    assert abs(mis[0] - ref["misfit"][3]) <= 1e-12 * ref["misfit"][3]
    eng.close()


def test_layered_visc_full_size_batch_equals_alone(gpu_lib):
    w = workloads.layered_visc()
    cfg = w.cfg
    shots = [(w.shots[k].srcpos, w.shots[k].recpos) for k in (5, 40)]
    print(f"[full size] layered_visc {cfg.NX}x{cfg.NY} NT={cfg.NT}, 2 shots", flush=True)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.true)
    obs = []
    for sp, rp in shots:
        eng.set_shot(sp, rp)
        eng.forward()
        vx, vy = eng.seismograms()
        vx[:, 0] = 0
        vy[:, 0] = 0
        obs.append((vx, vy))
    eng.set_model(**w.start)
    eng.reset_gradient()
    alone = []
    for (sp, rp), (ovx, ovy) in zip(shots, obs):
        eng.set_shot(sp, rp)
        eng.shot_gradient(ovx, ovy)
        alone.append(eng.gradient())
    total_alone = eng.gradient(total=True)
    eng.close()
    print("[full size] layered_visc, the two shots in one launch", flush=True)
    bat = ifos_amd.Engine(cfg, slots=2)
    bat.set_model(**w.start)
    bat.reset_gradient()
    bat.set_shots(shots)
    bat.shot_gradient(np.concatenate([o[0] for o in obs]), np.concatenate([o[1] for o in obs]))
    total = bat.gradient(total=True)
    bat.close()
    np.testing.assert_array_equal(total, total_alone)
    expect = (np.float32(0) + alone[0]) + alone[1]
    np.testing.assert_array_equal(total, expect.astype(np.float32))
    assert np.isfinite(total).all() and np.abs(total).max() > 0
