"""A full FWI iteration on the GPU (ifos_amd.fwi): gradient over the shots,
PCG or L-BFGS direction, the reference's step-length search with the
parabola, bounded update.  The data misfit must fall from iteration to
iteration on a small elastic anomaly model, and the chosen step must come
from three bracketing test misfits."""
import numpy as np
import pytest

from ifos_amd import workloads
from ifos_amd.fwi import FWI, FWISettings

pytestmark = pytest.mark.gpu


def anomaly_case():
    # VELOCITY 1: with VELOCITY 0 the reference's gradient is anti-parallel to the
    # gradient of its own displacement misfit (tests/test_oracle.py pins that),
    # so its step search finds no descent step
    w = workloads.small_case(seed=11, NX=120, NY=70, NT=500, FDORDER=4, L=0, VELOCITY=1, DTINV=2, FW=12)
    y, x = np.mgrid[0:70, 0:120]
    bump = np.exp(-((x - 60) ** 2 + (y - 35) ** 2) / (2 * 8.0 ** 2)).astype(np.float32)
    base = dict(vp=np.full((70, 120), 3000.0, np.float32), vs=np.full((70, 120), 1750.0, np.float32),
                rho=np.full((70, 120), 2100.0, np.float32))
    w.true = dict(vp=base["vp"] * (1 + 0.08 * bump), vs=base["vs"] * (1 + 0.08 * bump), rho=base["rho"].copy())
    w.start = base
    # three shots along the surface
    src = w.shots[0]
    shots = []
    for sx in (25, 60, 95):
        sp = np.array(src.srcpos, np.float32).reshape(-1, 8).copy()
        sp[:, 0] = sx
        shots.append(type(src)(srcpos=sp, recpos=src.recpos))
    w.shots = shots
    return w


@pytest.mark.parametrize("method,slots", [(1, 1), (1, 3), (2, 2), (1, 0)])
def test_fwi_iterations_reduce_misfit(gpu_lib, method, slots):
    w = anomaly_case()
    s = FWISettings(GRAD_METHOD=method, TESTSHOT_START=1, TESTSHOT_END=3, TESTSHOT_INCR=1, EPS_SCALE=0.01)
    f = FWI(w, s, slots=slots)
    misfits = []
    for it in range(1, 4):
        rec = f.iteration(it)
        print(f"[fwi] method {method} slots {slots} iter {it}: misfit {rec['misfit']:.5e} "
              f"step {rec['opteps']:.4g} L2t {rec['L2t']} epst {rec['epst']}", flush=True)
        misfits.append(rec["misfit"])
        assert rec["opteps"] > 0
        assert rec["epst"][1] > 0 and rec["L2t"][1] < rec["L2t"][0]
    final = f.test_misfit(f.model)
    f.close()
    assert misfits[1] < misfits[0] and misfits[2] < misfits[1] and final < misfits[2]


def test_fwi_batched_equals_one_slot(gpu_lib):
    """The iteration is the same computation whatever the number of slots:
    gradients and models agree bit for bit."""
    w = anomaly_case()
    s = FWISettings(TESTSHOT_START=1, TESTSHOT_END=3)
    a, b = FWI(w, s, slots=1), FWI(w, s, slots=3)
    ra, rb = a.iteration(1), b.iteration(1)
    assert ra["misfit"] == rb["misfit"] and ra["opteps"] == rb["opteps"]
    for k in ("vp", "vs", "rho"):
        np.testing.assert_array_equal(a.model[k], b.model[k])
    a.close()
    b.close()
