"""Source wavelets given by the caller (ifos_engine_set_wavelets, the
SOURCE_SHAPE 7 path the reference takes after the source-time-function
inversion, src/stf.c:265-272 and src/IFOS2D.c:1556) and the inversion loop
of src/stf.c:24 around the engine.

  * the generated wavelet handed back in gives the same seismograms bit
    for bit (the given signal replaces the generated one and nothing else);
  * twice the wavelet gives the same seismograms as the generated wavelet
    of a source with twice the amplitude, bit for bit; they are twice the
    first seismograms only to about 1e-5 of their peak, because values
    ahead of the wavefront pass through the subnormal range, where rounding
    is absolute (the oracle shows the same differences);
  * observed traces made with the wavelet convolved with a known short
    filter: the inverted correction filter, convolved with the wavelet,
    gives that convolved wavelet (times 1/dt, the library's scaling of the
    impulse response) to the accuracy the window's end allows, and the
    forward run with it reproduces the observed traces;
  * set_shots drops the given wavelets."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import stf as S
from ifos_amd import workloads
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _params(w):
    cfg = w.cfg
    return O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                         **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})


@pytest.mark.parametrize("case", [dict(seed=41, NX=150, NY=70, NT=600, FDORDER=8, L=0, VELOCITY=0, DTINV=3),
                                  dict(seed=42, NX=130, NY=60, NT=500, FDORDER=4, L=2, VELOCITY=1, DTINV=2)],
                         ids=["el8", "visc4"])
def test_given_wavelets_and_stf_loop(gpu_lib, case):
    w = workloads.small_case(**case)
    cfg = w.cfg
    sh = w.shots[0]
    sp = np.array(sh.srcpos, np.float32).reshape(-1, 8)
    rp = np.array(sh.recpos, np.int32).reshape(-1, 2)
    sig = O.wavelet(_params(w), sp)
    eng = ifos_amd.Engine(cfg)
    try:
        eng.set_model(**w.true)
        eng.set_shot(sp, rp)
        eng.forward()
        vx0, vy0 = eng.seismograms()
        assert np.abs(vy0).max() > 0
        print(f"[stf] {case['seed']} generated wavelet handed back", flush=True)
        eng.set_wavelets(sig)
        eng.forward()
        vx1, vy1 = eng.seismograms()
        np.testing.assert_array_equal(vx1, vx0)
        np.testing.assert_array_equal(vy1, vy0)
        print(f"[stf] {case['seed']} doubled wavelet", flush=True)
        eng.set_wavelets(2 * sig)
        eng.forward()
        vx2, vy2 = eng.seismograms()
        sp2 = sp.copy()
        sp2[:, 5] *= 2  # amplitude column
        eng.set_shot(sp2, rp)
        eng.forward()
        vxa, vya = eng.seismograms()
        np.testing.assert_array_equal(vx2, vxa)
        np.testing.assert_array_equal(vy2, vya)
        assert np.abs(vx2 - 2 * vx0).max() <= 1e-5 * np.abs(vx0).max()
        assert np.abs(vy2 - 2 * vy0).max() <= 1e-5 * np.abs(vy0).max()
        eng.set_shot(sp, rp)
        print(f"[stf] {case['seed']} inversion loop", flush=True)
        h = np.zeros(cfg.NT)
        h[0], h[4], h[9] = 0.8, 0.3, -0.2
        sig_h = np.stack([np.convolve(s.astype(np.float64), h)[:cfg.NT] for s in sig]).astype(np.float32)
        eng.set_wavelets(sig_h)
        eng.forward()
        _, obs_vy = eng.seismograms()
        new, filt, conv = S.invert("fdlsq:waterlevel=1e-6", obs_vy, vy0, sig[0], cfg.DT)
        scale = np.abs(sig_h[0]).max()
        assert np.abs(new * cfg.DT - sig_h[0]).max() <= 5e-2 * scale
        assert np.abs(conv - obs_vy).max() <= 2e-2 * np.abs(obs_vy).max()
        eng.set_wavelets((new * cfg.DT)[None, :])
        eng.forward()
        _, vy3 = eng.seismograms()
        assert np.abs(vy3 - obs_vy).max() <= 5e-2 * np.abs(obs_vy).max()
        # a new shot set: generated wavelets again
        eng.set_shot(sp, rp)
        eng.forward()
        _, vy4 = eng.seismograms()
        np.testing.assert_array_equal(vy4, vy0)
    finally:
        eng.close()
