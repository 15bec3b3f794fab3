"""Source-time-function inversion (include/ifos2d_stf.h) on the host.

The reference hands the traces to the stfinv library (contrib/stfinv,
FFTW-based; FFTW is absent here, so the library is not built) and
convolves with FFTW in conv_FD.  What pins this restatement:
  * the transform against the brute-force DFT sum, every length 1-70 and
    some large ones, both signs (power-of-two and Bluestein paths);
  * conv_fd against numpy's direct linear convolution (the convolution
    theorem; the padding of src/conv_FD.c:41 makes the cyclic convolution
    the linear one);
  * fdlsq: observed traces made as the synthetics convolved with a known
    short filter; the inverted filter is that filter (small waterlevel),
    the convolved synthetics are the observed traces; the waterlevel
    shrinks it by the closed-form factor on a single frequency; a time
    shift of a whole number of samples delays the impulse response by that
    many samples (periodic in the padded length); offset weights with
    exp=0 equal no weights; the padded length follows fpad/fpow2/fdiv;
  * taper: the reference's own taper() (src/taper.c) compiled into
    oracle/_ref, bit for bit.
Parity against the library's own output is unpinned (it cannot be built)."""
import numpy as np
import pytest

import oracle_lib as O
from ifos_amd import stf as S

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")


def brute_dft(x, sign):
    n = len(x)
    t = np.arange(n)
    return np.array([np.sum(x * np.exp(sign * 2j * np.pi * t * k / n)) for k in range(n)])


@pytest.mark.parametrize("sign", [-1, 1])
def test_dft_matches_brute_force(sign):
    rng = np.random.default_rng(1)
    for n in list(range(1, 71)) + [97, 128, 243, 1000, 1024]:
        x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        want = brute_dft(x, sign)
        got = S.dft(x, sign)
        assert np.abs(got - want).max() <= 1e-11 * max(1.0, np.abs(want).max()), n


def test_dft_round_trip_large_prime():
    rng = np.random.default_rng(2)
    n = 6751  # a prime, like 1.5 x 4500 padded lengths that are not powers of two
    x = rng.standard_normal(n)
    back = S.dft(S.dft(x, -1), 1) / n
    assert np.abs(back - x).max() <= 1e-10


def test_conv_fd_is_linear_convolution():
    rng = np.random.default_rng(3)
    for ns in (1, 2, 7, 100, 501, 2048):
        a = rng.standard_normal(ns).astype(np.float32)
        b = rng.standard_normal(ns).astype(np.float32)
        want = np.convolve(a.astype(np.float64), b.astype(np.float64))[:ns]
        got = S.conv_fd(a, b)
        assert np.abs(got - want).max() <= 1e-5 * max(1.0, np.abs(want).max()), ns


def _ricker(ns, dt, fc, t0):
    t = np.arange(ns) * dt - t0
    a = (np.pi * fc * t) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def test_fdlsq_recovers_a_known_filter():
    rng = np.random.default_rng(4)
    ns, ntr, dt = 600, 12, 1e-3
    # broadband synthetics that end 16 samples early: the whole linear
    # convolution with the short filter lies inside the window, and the padded
    # length 1.5*ns keeps the cyclic convolution equal to it
    syn = rng.standard_normal((ntr, ns)).astype(np.float32)
    syn[:, ns - 16:] = 0.0
    h = np.zeros(ns, np.float64)
    h[0], h[3], h[7] = 1.0, -0.5, 0.25          # a short causal filter
    obs = np.stack([np.convolve(s.astype(np.float64), h)[:ns] for s in syn]).astype(np.float32)
    stf, conv = S.fdlsq("fdlsq:waterlevel=1e-12", obs, syn, dt)
    # the library returns the impulse response scaled by 1/dt (scale_series 1/(n dt) of an unscaled c2r)
    assert np.abs(stf * dt - h).max() <= 1e-5
    assert np.abs(conv - obs).max() <= 1e-5 * np.abs(obs).max()


def test_fdlsq_waterlevel_closed_form():
    # one trace, obs = 2 * syn: H = 2 |S|^2 / (wl + |S|^2) per frequency; for a
    # single sinusoid filling whole periods of the padded length only one
    # frequency pair carries energy, with |S|^2 = (n A / 2)^2 and the energy sum n A^2 / 2
    n, k0, A, eps = 64, 5, 0.7, 0.25
    t = np.arange(n)
    syn = (A * np.cos(2 * np.pi * k0 * t / n))[None, :].astype(np.float64)
    obs = 2 * syn
    stf, conv = S.fdlsq(f"fdlsq:waterlevel={eps}:fpad=1", obs.astype(np.float32), syn.astype(np.float32), 1.0)
    S2 = (n * A / 2) ** 2
    wl = eps * n * A * A / 2
    gain = 2 * S2 / (wl + S2)
    # impulse response: gain * 2 cos(2 pi k0 t / n) / n (two bins), scaled by 1/(n dt) of the unscaled c2r
    want = gain * 2 * np.cos(2 * np.pi * k0 * t / n) / n
    np.testing.assert_allclose(stf, want, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(conv[0], gain * syn[0], rtol=1e-5, atol=1e-6)


def test_fdlsq_tshift_delays_the_filter():
    rng = np.random.default_rng(5)
    ns, ntr, dt = 300, 6, 2e-3
    syn = np.stack([_ricker(ns, dt, rng.uniform(10, 30), rng.uniform(0.1, 0.3)) for _ in range(ntr)])
    obs = (syn * rng.uniform(0.5, 2.0)).astype(np.float32)
    p0 = "fdlsq:fpad=1:fpow2=true:waterlevel=1e-6"
    h0, c0 = S.fdlsq(p0, obs, syn, dt)
    h5, c5 = S.fdlsq(p0 + f":tshift={5 * dt}", obs, syn, dt)
    n = 512
    full0 = np.zeros(n)
    full0[:ns] = h0  # the filter has its energy near zero lag and at the end of the period
    # the delayed filter is the periodic shift by 5 samples; compare on the window
    np.testing.assert_allclose(h5[5:], h0[:-5], rtol=0, atol=1e-4 * np.abs(h0).max())
    np.testing.assert_array_equal(c5, c0)   # the convolved synthetics are formed before the shift


def test_fdlsq_weights_and_padding():
    rng = np.random.default_rng(6)
    ns, ntr = 200, 5
    syn = rng.standard_normal((ntr, ns)).astype(np.float32)
    obs = rng.standard_normal((ntr, ns)).astype(np.float32)
    off = rng.uniform(10, 500, ntr)
    a, _ = S.fdlsq("fdlsq", obs, syn, 1e-3)
    b, _ = S.fdlsq("fdlsq:exp=0", obs, syn, 1e-3, off)
    np.testing.assert_array_equal(a, b)
    c, _ = S.fdlsq("fdlsq:exp=1.0", obs, syn, 1e-3, off)
    assert not np.array_equal(a, c)
    # weights scaled by a constant leave the result unchanged (the waterlevel scales with them)
    d, _ = S.fdlsq("fdlsq:exp=1.0", obs, syn, 1e-3, off * 1000.0)
    np.testing.assert_allclose(c, d, rtol=1e-5, atol=1e-6 * np.abs(c).max())
    p = S.parse("fdlsq:fpad=1.5:fdiv=64:tshift=0.0:irtap=0.,0.01,0.05,0.08")
    assert p.use_tshift == 1 and p.fdiv == 64 and list(p.irtap) == [0.0, 0.01, 0.05, 0.08]
    with pytest.raises(Exception):
        S.parse("fixed:tshift=0.0")
    with pytest.raises(Exception):
        S.parse("fdlsq:waterlevel=0")


@needs_ref
@pytest.mark.parametrize("fc,dt,ns", [(10.0, 1e-3, 2000), (25.0, 4e-4, 1500), (7.5, 2e-3, 700)])
def test_taper_matches_reference(fc, dt, ns):
    rng = np.random.default_rng(7)
    x = rng.standard_normal(ns).astype(np.float32)
    want = x.copy()
    O.ref_stf_taper(want, fc, dt)
    np.testing.assert_array_equal(S.taper(x, fc, dt), want)
