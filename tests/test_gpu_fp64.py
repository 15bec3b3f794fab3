"""The FP64 build of the scheme on the device (ifos_fp64_*, the yardstick
for 16-bit snapshot errors) against the oracle's FP64 build
(oracle/liboracle64.so) on the same inputs: seismograms and gradients
bit-identical, misfit to 1e-12 (the device sums per trace first).  The
checkpointed recomputation of the forward run must not change a bit: every
segment length, from DTINV steps to the whole run, gives the same
gradients."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O
from test_precision import params_of

pytestmark = pytest.mark.gpu

CASES = [
    dict(seed=41, NX=72, NY=52, NT=150, FDORDER=8, L=2, VELOCITY=1, DTINV=3),
    dict(seed=42, NX=64, NY=48, NT=120, FDORDER=4, L=0, VELOCITY=0, DTINV=2),
    dict(seed=43, NX=60, NY=50, NT=96, FDORDER=6, L=3, VELOCITY=1, DTINV=2, FREE_SURF=0),
    dict(seed=44, NX=58, NY=46, NT=100, FDORDER=2, L=1, VELOCITY=0, DTINV=4, stype=4),
    dict(seed=45, NX=56, NY=44, NT=90, FDORDER=12, L=0, VELOCITY=1, DTINV=3, stype=2),
]


def _obs(w, p64):
    s = w.shots[0]
    ovx, ovy, _ = O.forward(p64, w.true, s.srcpos, s.recpos, fp64=True)
    ovx, ovy = ovx.astype(np.float32), ovy.astype(np.float32)
    ovx[:, 0] = 0
    ovy[:, 0] = 0
    return ovx, ovy


@pytest.mark.parametrize("case", CASES, ids=["visc8", "el4_v0", "visc6_L3_nofs", "visc2_force", "el12_force"])
def test_fp64_engine_equals_fp64_oracle(gpu_lib, case):
    print(f"[fp64] case {case}", flush=True)
    w = workloads.small_case(**case)
    s = w.shots[0]
    p64 = params_of(w)
    ovx, ovy = _obs(w, p64)
    kill = np.zeros(len(s.recpos), np.int32)
    kill[1::5] = 1
    p64.TRKILL = 1
    want = O.gradient(p64, w.start, s.srcpos, s.recpos, ovx, ovy, kill=kill, fp64=True)
    w.cfg.TRKILL = 1
    eng = ifos_amd.Fp64Engine(w.cfg, segment_steps=2 * case["DTINV"])
    eng.set_model(**w.start)
    eng.set_shot(s.srcpos, s.recpos)
    mis = eng.shot_gradient(ovx, ovy, kill)
    vx, vy = eng.seismograms()
    g = eng.gradient()
    eng.close()
    np.testing.assert_array_equal(vx, want["seis_vx"])
    np.testing.assert_array_equal(vy, want["seis_vy"])
    for k, name in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(g[k], want["grad_" + name], err_msg=f"grad {name}")
    assert abs(mis[0] - want["misfit"][3]) <= 1e-12 * want["misfit"][3]
    assert abs(mis[1] - want["misfit"][4]) <= 1e-12 * want["misfit"][4]


def test_fp64_segments_do_not_change_the_gradient(gpu_lib):
    case = dict(seed=46, NX=80, NY=56, NT=180, FDORDER=8, L=2, VELOCITY=1, DTINV=3)
    w = workloads.small_case(**case)
    s = w.shots[0]
    ovx, ovy = _obs(w, params_of(w))
    got = []
    for seg in (3, 21, 60, 0, 180):
        print(f"[fp64] segment {seg}", flush=True)
        eng = ifos_amd.Fp64Engine(w.cfg, segment_steps=seg)
        if seg:
            assert eng.segment_steps == seg
        else:
            assert eng.segment_steps == 3 * 8  # DTINV * ceil(sqrt(60))
        eng.set_model(**w.start)
        eng.set_shot(s.srcpos, s.recpos)
        eng.shot_gradient(ovx, ovy)
        eng.shot_gradient(ovx, ovy)        # a second shot adds to the sum
        got.append((eng.gradient(), eng.gradient(total=True)))
        eng.close()
    for g, t in got[1:]:
        np.testing.assert_array_equal(g, got[0][0])
        np.testing.assert_array_equal(t, got[0][1])
    np.testing.assert_array_equal(got[0][1], got[0][0] + got[0][0])


def test_fp64_rejects_bad_segments(gpu_lib):
    w = workloads.small_case(seed=47, NX=40, NY=30, NT=60, DTINV=3)
    with pytest.raises(ifos_amd.IfosError):
        ifos_amd.Fp64Engine(w.cfg, segment_steps=4)    # not a multiple of DTINV
    assert ifos_amd.fp64_bytes(w.cfg, 4) == 0
    assert ifos_amd.fp64_bytes(w.cfg, 6) > 0
