"""Time-domain Butterworth filter on the GPU (TIME_FILT): the device filter
against the oracle (bit for bit; the oracle equals the reference's
timedomain_filt + libcseife, tests/test_filter.py), and its use in the
engine: observed traces filtered before the residual, source wavelets
filtered before the forward run."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order,method,fhp", [(4, 1, 0.0), (3, 1, 0.0), (5, 2, 0.0), (6, 1, 4.0), (1, 1, 0.0)])
def test_device_filter_matches_oracle(gpu_lib, order, method, fhp):
    rng = np.random.default_rng(order * 10 + method)
    d = (rng.standard_normal((37, 1500)) * 1e-5).astype(np.float32)
    want = O.time_filter(d, 4e-4, 30.0, order, method, fhp)
    buf = ifos_amd.DeviceBuffer.from_array(d)
    ifos_amd.time_filter(buf, 37, 1500, 4e-4, 30.0, order, method, fhp)
    got = buf.to_array(d.shape)
    buf.free()
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_engine_filters_observed_traces_and_wavelets(gpu_lib):
    w = workloads.small_case(seed=61, NX=90, NY=50, NT=300, FDORDER=4, L=0, VELOCITY=1, DTINV=2)
    sp, rp = w.shots[0].srcpos, w.shots[0].recpos
    fc, order = 60.0, 4
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.true)
    eng.set_shot(sp, rp)
    eng.forward()
    raw = eng.seismograms()
    # wavelets: the wave equation is linear and time-invariant, so the
    # seismograms of the filtered wavelet are the filtered seismograms
    eng.set_time_filter(fc, order)
    eng.forward()
    filt = eng.seismograms()
    for k in range(2):
        want = O.time_filter(raw[k], w.cfg.DT, fc, order, 1)
        err = np.linalg.norm(filt[k] - want) / np.linalg.norm(want)
        assert err < 1e-4, err
    # observed traces are filtered in the engine's copy before the residual:
    # the misfit is that of the engine's (filtered-wavelet) synthetics
    # against the oracle-filtered observed traces
    obs = [O.time_filter(r, w.cfg.DT, 25.0, 3, 1) for r in raw]  # some other traces as "observed"
    fobs = [O.time_filter(o, w.cfg.DT, fc, order, 1) for o in obs]
    m_host = eng.residual(obs[0], obs[1])
    l2 = sum(float(np.sum((filt[k] - fobs[k]).astype(np.float64) ** 2)) for k in range(2))
    en = sum(float(np.sum(fobs[k].astype(np.float64) ** 2)) for k in range(2))
    np.testing.assert_allclose(m_host, [l2, en], rtol=1e-10)
    # device observed traces: the same misfit, the caller's buffers untouched
    dev = [ifos_amd.DeviceBuffer.from_array(o) for o in obs]
    m_dev = eng.residual(dev[0], dev[1])
    np.testing.assert_array_equal(m_dev, m_host)
    for k in range(2):
        np.testing.assert_array_equal(dev[k].to_array(obs[k].shape), obs[k])
        dev[k].free()
    # turning the filter off restores the unfiltered wavelets
    eng.set_time_filter(0.0)
    eng.forward()
    np.testing.assert_array_equal(eng.seismograms()[0], raw[0])
    eng.close()
