"""Gradient post-processing on the GPU against the oracle, itself
bit-identical to the reference's taper_grad_shot / taper_grad / smooth
(tests/test_post.py): the per-shot source taper fused into the gradient
kernel, the whole-gradient tapers and the smoothing of the optimizer
library.  Tapers bit-exact; smoothing value-exact (a selection)."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import opt, workloads
import oracle_lib as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,radius,filtsize", [(1, 12.0, 1), (2, 30.0, 6)])
def test_shot_taper_matches_oracle(gpu_lib, shape, radius, filtsize):
    w = workloads.small_case(seed=41 + shape, NX=90, NY=50, NT=160, FDORDER=4, L=0, VELOCITY=1, DTINV=2)
    base = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8)
    shots = []
    # the last source sits 3 cells from the left edge: a 6-cell box wraps
    for x in (30, 61, 3):
        sp = base.copy()
        sp[:, 0] = x
        shots.append((sp, w.shots[0].recpos))
    eng1 = ifos_amd.Engine(w.cfg, slots=1)
    eng1.set_model(**w.true)
    obs = []
    for sp, rp in shots:
        eng1.set_shot(sp, rp)
        eng1.forward()
        obs.append(eng1.seismograms())
    eng1.set_model(**w.start)
    want = None
    for (sp, rp), (ovx, ovy) in zip(shots, obs):
        eng1.set_shot(sp, rp)
        eng1.shot_gradient(ovx, ovy)
        g = eng1.gradient()
        is_, js = int(sp[0, 0]), int(sp[0, 1])
        t = np.stack([O.taper_sources(g[k], w.cfg.DH, radius, shape, filtsize, [is_], [js]) for k in range(3)])
        want = t if want is None else want + t
    eng1.close()
    eng = ifos_amd.Engine(w.cfg, slots=3)
    eng.set_model(**w.start)
    eng.set_shot_taper(True, radius, shape, filtsize)
    eng.set_shots(shots)
    eng.shot_gradient(np.concatenate([o[0] for o in obs]), np.concatenate([o[1] for o in obs]))
    got = eng.gradient()
    for k, n in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(got[k].view(np.uint32), want[k].view(np.uint32), err_msg=n)
    eng.set_shot_taper(False)
    eng.close()


def test_whole_gradient_tapers_match_oracle(gpu_lib):
    NX, NY = 70, 45
    rng = np.random.default_rng(5)
    g = rng.standard_normal((NY, NX)).astype(np.float32)
    o = opt.Optimizer()
    buf = ifos_amd.DeviceBuffer(g.nbytes)
    src = np.array([[12, 4], [40, 6], [2, 20]], np.int32)
    rec = np.array([[x, 3] for x in range(5, 66, 6)], np.int32)
    cases = [
        (opt.TAPER_VERTICAL, dict(gradt=(6, 14, 50, 62), srcpos=src, recpos=rec),
         lambda a: O.taper_vertical(a, (6, 14, 50, 62), src[:, 0], src[:, 1], rec)),
        (opt.TAPER_HORIZONTAL, dict(gradt=(4, 11, 0, 0)), lambda a: O.taper_horizontal(a, (4, 11))),
        (opt.TAPER_SOURCES, dict(DH=2.0, SRTRADIUS=14.0, SRTSHAPE=1, FILTSIZE=4, srcpos=src),
         lambda a: O.taper_sources(a, 2.0, 14.0, 1, 4, src[:, 0], src[:, 1], ishot=0)),
        (opt.TAPER_SOURCES, dict(DH=2.0, SRTRADIUS=9.0, SRTSHAPE=2, FILTSIZE=1, srcpos=src),
         lambda a: O.taper_sources(a, 2.0, 9.0, 2, 1, src[:, 0], src[:, 1], ishot=0)),
    ]
    for kind, kw, ref in cases:
        coef, zero = opt.taper_coeff(kind, NX, NY, **kw)
        buf.upload(g)
        o.taper(buf, coef, zero)
        np.testing.assert_array_equal(buf.to_array((NY, NX)).view(np.uint32), ref(g).view(np.uint32),
                                      err_msg=str(kind))
    buf.free()
    o.close()


@pytest.mark.parametrize("fs", [3, 4, 7, 15, -5])
def test_smooth_matches_oracle(gpu_lib, fs):
    NX, NY = 70, 45
    rng = np.random.default_rng(9 + abs(fs))
    g = rng.standard_normal((NY, NX)).astype(np.float32)
    g[rng.random((NY, NX)) < 0.1] = 0.0
    g[10:14, 20:30] = 0.75
    o = opt.Optimizer()
    buf = ifos_amd.DeviceBuffer.from_array(g)
    side = o.smooth(NX, NY, fs, buf)
    want = O.median_smooth(g, fs)
    assert side == abs(fs) + (1 - abs(fs) % 2)
    np.testing.assert_array_equal(buf.to_array((NY, NX)), want)
    buf.free()
    o.close()


@pytest.mark.parametrize("size,first", [(3, 1), (6, 5), (1, 2)])
def test_spat_filt_matches_oracle(gpu_lib, size, first):
    NX, NY = 70, 45
    g = np.random.default_rng(size + first).standard_normal((NY, NX)).astype(np.float32)
    o = opt.Optimizer()
    buf = ifos_amd.DeviceBuffer.from_array(g)
    o.spat_filt(NX, NY, size, first, buf)
    want = O.spat_filt(g, size, first)
    np.testing.assert_array_equal(buf.to_array((NY, NX)).view(np.uint32), want.view(np.uint32))
    buf.free()
    o.close()


def test_shot_taper_files(gpu_lib):
    """Per-shot taper grids multiply each shot's gradients after the
    circular taper and before the sum (src/IFOS2D.c:2681-2698)."""
    w = workloads.small_case(seed=47, NX=80, NY=44, NT=140, FDORDER=4, L=0, VELOCITY=1, DTINV=2)
    base = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8)
    shots = []
    for x in (25, 55):
        sp = base.copy()
        sp[:, 0] = x
        shots.append((sp, w.shots[0].recpos))
    rng = np.random.default_rng(3)
    tapers = [[rng.random((44, 80)).astype(np.float32) for _ in range(3)] for _ in shots]
    eng1 = ifos_amd.Engine(w.cfg, slots=1)
    eng1.set_model(**w.true)
    obs = []
    for sp, rp in shots:
        eng1.set_shot(sp, rp)
        eng1.forward()
        obs.append(eng1.seismograms())
    eng1.set_model(**w.start)
    eng1.set_shot_taper(True, 10.0, 1, 1)
    want = None
    for (sp, rp), (ovx, ovy), t in zip(shots, obs, tapers):
        eng1.set_shot(sp, rp)
        eng1.shot_gradient(ovx, ovy)
        g = eng1.gradient() * np.stack(t)
        want = g if want is None else want + g
    eng1.close()
    eng = ifos_amd.Engine(w.cfg, slots=2)
    eng.set_model(**w.start)
    eng.set_shot_taper(True, 10.0, 1, 1)
    for k, t in enumerate(tapers):
        eng.set_shot_taper_file(k, *t)
    eng.set_shots(shots)
    eng.shot_gradient(np.concatenate([o[0] for o in obs]), np.concatenate([o[1] for o in obs]))
    got = eng.gradient()
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    eng.set_shot_taper_file(0)
    eng.close()
