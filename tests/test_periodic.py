"""Periodic sides (BOUNDARY 1) in the oracle.

The reference makes the left and right edges periodic through its halo
exchanges, which on one MPI rank copy the opposite edge into the halo
columns (src/exchange_v.c:143-228 after the velocity update,
src/exchange_s.c:145-229 after the stress update, sources and free
surface), and switches the CPML off on those edges (src/update_v_PML.c:84-95,
src/update_s_visc_PML.c:73-83, src/surface_PML.c:109-116).

  * the oracle against the reference's own kernels and exchanges
    (oracle/_ref compiles src/exchange_v.c and src/exchange_s.c and runs them
    on one MPI rank): forward seismograms and final wavefields, and the shot
    gradient, bit for bit;
  * what the mathematics fixes: on a model that does not vary along x the
    periodic scheme commutes with circular shifts in x, so moving the source
    by k columns moves the seismograms of a full receiver row and the final
    wavefields by k columns, bit for bit.  A halo column left out or copied
    from the wrong column breaks this at the edges.  With BOUNDARY 0 (CPML
    sides) the same shift changes the result, so the check can fail.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as O
from ifos_amd import workloads

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")


def params_of(w, boundary=1):
    cfg = w.cfg
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                      **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    p.BOUNDARY = boundary
    return p


def uniform_in_x(model, NX):
    """The model's first column repeated along x."""
    return {k: np.ascontiguousarray(np.repeat(v[:, :1], NX, axis=1)) for k, v in model.items()}


@needs_ref
@pytest.mark.parametrize("fd,L,fs", list(itertools.product([2, 4, 6, 8], [0, 2], [0, 1])))
def test_periodic_forward_matches_reference(fd, L, fs):
    w = workloads.small_case(seed=300 + fd + L + fs, NX=44, NY=32, NT=90, FDORDER=fd, L=L, VELOCITY=1,
                             FREE_SURF=fs, FW=8, DTINV=3)
    p = params_of(w)
    s = w.shots[0]
    sp = np.array(s.srcpos, np.float32).reshape(-1, 8).copy()
    sp[:, 0] = 3  # next to the left edge, so the wave crosses it within NT
    a = O.forward(p, w.true, sp, s.recpos)
    b = O.forward(p, w.true, sp, s.recpos, which="ref")
    for x, y, n in zip(a, b, ("vx", "vy", "fields")):
        np.testing.assert_array_equal(x, y, err_msg=n)
    # and the periodic sides did change the result
    c = O.forward(params_of(w, 0), w.true, sp, s.recpos)
    assert not np.array_equal(a[2], c[2])


@needs_ref
@pytest.mark.parametrize("fd,L", list(itertools.product([4, 8], [0, 2])))
def test_periodic_gradient_matches_reference(fd, L):
    w = workloads.small_case(seed=320 + fd + L, NX=48, NY=36, NT=120, FDORDER=fd, L=L, VELOCITY=1, DTINV=3, FW=8)
    p = params_of(w)
    s = w.shots[0]
    sp = np.array(s.srcpos, np.float32).reshape(-1, 8).copy()
    sp[:, 0] = 4
    ovx, ovy, _ = O.forward(p, w.true, sp, s.recpos)
    ovx[:, 0] = 0
    ovy[:, 0] = 0
    g = O.gradient(p, w.start, sp, s.recpos, ovx, ovy)
    dvx, _ = O.calc_res(p, g["seis_vx"], ovx)
    dvy, _ = O.calc_res(p, g["seis_vy"], ovy)
    r = O.ref_gradient(p, w.start, sp, s.recpos, dvx, dvy)
    for k in ("seis_vx", "seis_vy", "grad_vp", "grad_vs", "grad_rho"):
        np.testing.assert_array_equal(g[k], r[k], err_msg=k)


def shifted_runs(fd, L, fs, boundary, k):
    NX, NY = 40, 30
    w = workloads.small_case(seed=340 + fd + L, NX=NX, NY=NY, NT=80, FDORDER=fd, L=L, VELOCITY=1, FREE_SURF=fs,
                             FW=6, DTINV=2)
    p = params_of(w, boundary)
    model = uniform_in_x(w.true, NX)
    sp = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8).copy()
    sp[:, 0], sp[:, 1] = 5, 9
    rec = np.stack([np.arange(1, NX + 1), np.full(NX, 12)], 1).astype(np.int32)
    a = O.forward(p, model, sp, rec)
    sp2 = sp.copy()
    sp2[:, 0] = (sp[:, 0] - 1 + k) % NX + 1
    b = O.forward(p, model, sp2, rec)
    return a, b


@pytest.mark.parametrize("fd,L,fs", list(itertools.product([2, 4, 6, 8, 10, 12], [0, 2], [0, 1])))
def test_periodic_scheme_commutes_with_shifts(fd, L, fs):
    k = 7
    (vx, vy, f), (vx2, vy2, f2) = shifted_runs(fd, L, fs, 1, k)
    # receiver t of the shifted run sees what receiver t - k saw
    np.testing.assert_array_equal(vx2, np.roll(vx, k, axis=0), err_msg="vx")
    np.testing.assert_array_equal(vy2, np.roll(vy, k, axis=0), err_msg="vy")
    np.testing.assert_array_equal(f2, np.roll(f, k, axis=2), err_msg="fields")
    assert np.abs(f).max() > 0


def test_cpml_sides_do_not_commute_with_shifts():
    (vx, vy, f), (vx2, vy2, f2) = shifted_runs(4, 0, 1, 0, 7)
    assert not np.array_equal(f2, np.roll(f, 7, axis=2))
