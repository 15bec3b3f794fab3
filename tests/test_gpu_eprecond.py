"""Energy preconditioning on the GPU (EPRECOND 1 and 3, per shot or summed
over shots) against the oracle, itself bit-identical to the reference's
eprecond / eprecond1 (tests/test_oracle.py); with the energy sums in FP32 as
the reference forms them, and in FP64 (set_eprecond_fp64, the oracle's
EPRECOND_FP64)."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O
from test_precision import params_of

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fp64", [0, 1], ids=["fp32_sums", "fp64_sums"])
@pytest.mark.parametrize("mode,per_shot,L", [(1, 1, 0), (1, 0, 2), (3, 1, 2), (3, 0, 0)])
def test_eprecond_matches_oracle(gpu_lib, mode, per_shot, L, fp64):
    w = workloads.small_case(seed=23 + mode + L, NX=90, NY=50, NT=200, FDORDER=4, L=L, VELOCITY=1, DTINV=2)
    shots = [w.shots[0]]
    sp2 = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8).copy()
    sp2[:, 0] += 20
    shots.append(type(w.shots[0])(srcpos=sp2, recpos=w.shots[0].recpos))
    p = params_of(w, EPRECOND=mode, EPRECOND_PER_SHOT=per_shot, EPSILON_WE=0.005, EPRECOND_FP64=fp64)
    eng = ifos_amd.Engine(w.cfg, slots=2)
    eng.set_model(**w.start)
    eng.set_eprecond(mode, 0.005, per_shot)
    eng.set_eprecond_fp64(fp64)
    eng.reset_gradient()
    obs, want, wes = [], [], []
    for s in shots:
        ovx, ovy, _ = O.forward(p, w.true, s.srcpos, s.recpos)
        ovx[:, 0] = 0
        ovy[:, 0] = 0
        obs.append((ovx, ovy))
        g, we = O.gradient_we(p, w.start, s.srcpos, s.recpos, ovx, ovy)
        want.append(g)
        wes.append(we)
    # one shot alone, then both shots in one launch
    eng.set_shots([(shots[0].srcpos, shots[0].recpos)])
    eng.shot_gradient(obs[0][0], obs[0][1])
    np.testing.assert_array_equal(eng.energy_weight(0), wes[0])
    g = eng.gradient()
    for k, n in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(g[k], want[0]["grad_" + n], err_msg=n)
    eng.reset_gradient()
    if not per_shot:
        eng.apply_eprecond()  # clears the 1/We sum of the first run
        eng.reset_gradient()
    eng.set_shots([(s.srcpos, s.recpos) for s in shots])
    eng.shot_gradient(np.concatenate([o[0] for o in obs]), np.concatenate([o[1] for o in obs]))
    total = eng.gradient(total=True)
    sums = [want[0]["grad_" + n] + want[1]["grad_" + n] for n in ("vp", "vs", "rho")]
    if per_shot:
        for k in range(3):
            np.testing.assert_array_equal(total[k], sums[k])
    else:
        wsum = np.zeros_like(wes[0])
        for we in wes:
            wsum = (wsum.astype(np.float64) + 1.0 / we.astype(np.float64)).astype(np.float32)
        np.testing.assert_array_equal(eng.energy_weight(1), wsum)
        eng.apply_eprecond()
        scale = wsum / np.float32(np.abs(wsum).max())
        total = eng.gradient(total=True)
        for k in range(3):
            np.testing.assert_array_equal(total[k], sums[k] * scale)
    eng.close()
