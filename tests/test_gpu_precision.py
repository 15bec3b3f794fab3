"""16-bit snapshots on the GPU: gradients equal the oracle run with the same
snapshot quantisation bit for bit, the forward seismograms are untouched,
and the gradient error against the FP64 build of the scheme stays within
1e-3 (relative L2 per parameter)."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O
from test_precision import PREC_CASES, params_of, rel

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", PREC_CASES, ids=["visc8_v1", "el4_v0"])
def test_snapshot16_gradient_matches_oracle(gpu_lib, case):
    w = workloads.small_case(**case)
    s = w.shots[0]
    p64 = params_of(w)
    ovx, ovy, _ = O.forward(p64, w.true, s.srcpos, s.recpos, fp64=True)
    ovx, ovy = ovx.astype(np.float32), ovy.astype(np.float32)
    g64 = O.gradient(p64, w.start, s.srcpos, s.recpos, ovx, ovy, fp64=True)
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.start)
    eng.set_shot(s.srcpos, s.recpos)
    eng.shot_gradient(ovx, ovy)
    seis32 = eng.seismograms()
    for prec in (1, 2):
        print(f"[precision] case {case} prec {prec}", flush=True)
        want = O.gradient(params_of(w, SNAP_PREC=prec), w.start, s.srcpos, s.recpos, ovx, ovy)
        eng.set_snapshot_precision(prec)
        eng.shot_gradient(ovx, ovy)
        vx, vy = eng.seismograms()
        np.testing.assert_array_equal(vx, seis32[0])
        np.testing.assert_array_equal(vy, seis32[1])
        g = eng.gradient()
        for k, name in enumerate(("vp", "vs", "rho")):
            np.testing.assert_array_equal(g[k], want["grad_" + name], err_msg=f"prec {prec} grad {name}")
            assert rel(g[k], g64["grad_" + name]) <= 1e-3
    eng.close()


def test_snapshot_precision_switch_and_sizes(gpu_lib):
    w = workloads.small_case(seed=5, NX=70, NY=50, NT=120, L=0, DTINV=2)
    b32 = ifos_amd.snapshot_bytes(w.cfg, "fp32")
    b16 = ifos_amd.snapshot_bytes(w.cfg, "bf16")
    plane, nsnap = 70 * 50, 60
    ntile = 2 * 7
    assert b32 == 4 * 5 * nsnap * plane
    assert b16 == 2 * 5 * nsnap * plane + 4 * 5 * nsnap * ntile
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.start)
    eng.set_shot(w.shots[0].srcpos, w.shots[0].recpos)
    eng.forward(snapshots=True)
    eng.set_snapshot_precision("fp16")   # frees the FP32 snapshots
    with pytest.raises(ifos_amd.IfosError):
        eng.adjoint()
    with pytest.raises(ifos_amd.IfosError):
        eng.set_snapshot_precision(3)
    eng.close()
