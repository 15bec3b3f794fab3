"""Batched shots: an engine with S slots advances S shots in the same kernel
launches.  Each shot's seismograms, misfit and gradient must equal the same
shot run alone on a one-slot engine bit for bit, for every launch shape
(shots per workgroup 1/2/4, one workgroup per item or a persistent grid),
and the batch gradient must equal the sequential sum over the shots in slot
order (the reference's shot loop, src/IFOS2D.c:2703-2713); each slot's own
gradient (ifos_engine_slot_gradient) equals that shot's gradient alone."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O

pytestmark = pytest.mark.gpu


def make_shots(w, n, seed=0):
    """n shots with different source positions and uneven receiver spreads."""
    rng = np.random.default_rng(seed)
    cfg = w.cfg
    base = w.shots[0]
    shots = []
    for k in range(n):
        sp = np.array(base.srcpos, np.float32).reshape(-1, 8).copy()
        sp[:, 0] = rng.integers(cfg.FW + 3, cfg.NX - cfg.FW - 3)
        sp[:, 7] = [1, 2, 3, 4][k % 4]
        rec = np.array(base.recpos, np.int32).reshape(-1, 2)
        keep = rng.random(len(rec)) < rng.uniform(0.3, 1.0)
        keep[0] = True
        shots.append((sp, rec[keep]))
    return shots


def observed(w, shots):
    out = []
    for sp, rp in shots:
        vx, vy, _ = O.forward(O_params(w), w.true, sp, rp)
        vx[:, 0] = 0.0
        vy[:, 0] = 0.0
        out.append((vx.astype(np.float32), vy.astype(np.float32)))
    return out


def O_params(w):
    cfg = w.cfg
    return O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                         **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})


CASES = [dict(seed=41, NX=150, NY=70, NT=240, FDORDER=8, L=2, VELOCITY=1, DTINV=3),
         dict(seed=42, NX=130, NY=60, NT=200, FDORDER=4, L=0, VELOCITY=0, DTINV=2)]


@pytest.mark.parametrize("case", CASES, ids=["visc8", "el4"])
def test_batch_equals_shots_alone(gpu_lib, case):
    w = workloads.small_case(**case)
    shots = make_shots(w, 5)
    obs = observed(w, shots)
    # each shot alone, in order, on a one-slot engine
    one = ifos_amd.Engine(w.cfg)
    one.set_model(**w.start)
    one.reset_gradient()
    alone = []
    for (sp, rp), (ovx, ovy) in zip(shots, obs):
        one.set_shot(sp, rp)
        mis = one.shot_gradient(ovx, ovy)
        alone.append((one.seismograms(), mis, one.gradient()))
    total_alone = one.gradient(total=True)
    one.close()
    obs_vx = np.concatenate([o[0] for o in obs])
    obs_vy = np.concatenate([o[1] for o in obs])
    eng = ifos_amd.Engine(w.cfg, slots=6)
    eng.set_model(**w.start)
    for spb, grid, order in [(1, 0, 0), (2, 0, 0), (4, 0, 0), (1, 512, 0), (2, 256, 0), (4, 1024, 0),
                             (1, 0, 1), (2, 0, 1), (4, 512, 1)]:
        print(f"[batch] case {case['seed']} spb {spb} grid {grid} order {order}", flush=True)
        eng.set_launch(spb, grid)
        eng.set_work_order(order)
        eng.reset_gradient()
        eng.set_shots(shots)
        eng.shot_gradient(obs_vx, obs_vy)
        vx, vy = eng.seismograms()
        per = eng.shot_misfit()
        t0 = 0
        for k, ((avx, avy), mis, _) in enumerate(alone):
            n = avx.shape[0]
            np.testing.assert_array_equal(vx[t0:t0 + n], avx, err_msg=f"shot {k} vx")
            np.testing.assert_array_equal(vy[t0:t0 + n], avy, err_msg=f"shot {k} vy")
            assert per[k, 0] == mis[0] and per[k, 1] == mis[1], (k, per[k], mis)
            t0 += n
        # each slot keeps its own shot's gradient, as that shot alone gives it
        for k, (_, _, g) in enumerate(alone):
            np.testing.assert_array_equal(eng.slot_gradient(k), g, err_msg=f"slot {k} gradient")
        # the batch gradient is the slot-order sum of the per-shot gradients
        want = alone[0][2].copy()
        for _, _, g in alone[1:]:
            want = want + g
        np.testing.assert_array_equal(eng.gradient(), want)
        np.testing.assert_array_equal(eng.gradient(total=True), total_alone)
    eng.close()


def test_batch_16bit_snapshots_and_autotune(gpu_lib):
    w = workloads.small_case(**CASES[0])
    shots = make_shots(w, 3, seed=7)
    obs = observed(w, shots)
    one = ifos_amd.Engine(w.cfg)
    one.set_model(**w.start)
    one.set_snapshot_precision("fp16")
    grads = []
    for (sp, rp), (ovx, ovy) in zip(shots, obs):
        one.set_shot(sp, rp)
        one.shot_gradient(ovx, ovy)
        grads.append(one.gradient())
    one.close()
    eng = ifos_amd.Engine(w.cfg, slots=3)
    eng.set_model(**w.start)
    eng.set_snapshot_precision("fp16")
    eng.set_shots(shots)
    spb, grid, ms = eng.autotune(reps=3)
    print(f"[batch] autotune picked spb {spb} grid {grid} ({ms:.4f} ms per step)", flush=True)
    assert spb in (1, 2) and grid % 8 == 0 and ms > 0
    eng.shot_gradient(np.concatenate([o[0] for o in obs]), np.concatenate([o[1] for o in obs]))
    np.testing.assert_array_equal(eng.gradient(), grads[0] + grads[1] + grads[2])
    eng.close()


def test_batch_rejects_too_many_shots(gpu_lib):
    w = workloads.small_case(seed=1, NX=70, NY=40, NT=60)
    eng = ifos_amd.Engine(w.cfg, slots=2)
    shots = [(w.shots[0].srcpos, w.shots[0].recpos)] * 3
    with pytest.raises(ifos_amd.IfosError):
        eng.set_shots(shots)
    with pytest.raises(ifos_amd.IfosError):
        eng.set_launch(3, 0)
    with pytest.raises(ifos_amd.IfosError):
        eng.set_launch(1, 12)
    eng.close()


def test_kernel_and_copy_timing(gpu_lib):
    """ifos_engine_time_kernel and ifos_engine_time_copy (the figures of
    bench.py --kernels-only --copy-floor): positive times, a copy of 256 MB
    takes longer than one of 4 KB, and the engine still runs afterwards."""
    w = workloads.small_case(seed=2, NX=200, NY=120, NT=60)
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.true)
    eng.set_shot(w.shots[0].srcpos, w.shots[0].recpos)
    for k in (0, 1):
        assert eng.time_kernel(k, reps=3) > 0
    small, large = eng.time_copy(4096, reps=5), eng.time_copy(256 << 20, reps=5)
    assert 0 < small < large
    eng.forward()  # the engine still runs after the timings
    assert np.isfinite(eng.seismograms()[1]).all()
    eng.close()
