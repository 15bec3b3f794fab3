"""Time-domain Butterworth filter (multi-scale inversion, TIME_FILT): the
oracle (oracle/filt_oracle.c) against the reference's timedomain_filt and
its vendored libcseife compiled from source (oracle/_ref), bit for bit, and
against closed forms (DC gain of the lowpass, a highpass removing a
constant)."""
import numpy as np
import pytest

import oracle_lib as ol

needs_ref = pytest.mark.skipif(not ol.ref_available(), reason="reference tree not built into oracle/_ref")


def _traces(ntr, ns, seed):
    rng = np.random.default_rng(seed)
    t = np.arange(ns) * 1e-3
    d = np.sin(2 * np.pi * 12.0 * t)[None, :] + 0.5 * rng.standard_normal((ntr, ns))
    return (d * 1e-6).astype(np.float32)


@needs_ref
@pytest.mark.parametrize("order", [1, 2, 3, 4, 5, 6])
@pytest.mark.parametrize("method,fhp", [(1, 0.0), (2, 0.0), (1, 3.0)])
def test_filter_matches_reference(order, method, fhp):
    d = _traces(7, 900, order)
    ref = ol.time_filter(d, 5e-4, 25.0, order, method, fhp, which="ref")
    orc = ol.time_filter(d, 5e-4, 25.0, order, method, fhp)
    np.testing.assert_array_equal(orc.view(np.uint32), ref.view(np.uint32))
    assert not np.array_equal(orc, d)


def test_lowpass_dc_gain_and_highpass_constant():
    ns = 6000
    c = np.full((2, ns), 3.0, np.float32)
    low = ol.time_filter(c, 1e-3, 10.0, 4, 1)
    np.testing.assert_allclose(low[:, -1], 3.0, rtol=1e-5)
    # a step: the highpass output decays to zero
    high = ol.time_filter(c, 1e-3, 10.0, 4, 2)
    assert abs(high[0, -1]) < 1e-4 * 3.0
    # a sinusoid far below the corner keeps its amplitude (the IIR sections
    # only delay it); one far above is removed
    t = np.arange(ns) * 1e-3
    s = np.sin(2 * np.pi * 0.5 * t).astype(np.float32)[None, :]
    out = ol.time_filter(s, 1e-3, 20.0, 4, 1)
    assert abs(np.max(np.abs(out[0, 3000:])) - 1.0) < 1e-3
    s = np.sin(2 * np.pi * 200.0 * t).astype(np.float32)[None, :]
    out = ol.time_filter(s, 1e-3, 20.0, 4, 1)
    assert np.max(np.abs(out[0, 3000:])) < 1e-3
