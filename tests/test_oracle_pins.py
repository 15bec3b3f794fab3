"""Pins of the oracle's stencils at every FDORDER (2-12), its free surface,
and its PCG, against the reference's own compiled functions (oracle/_ref)
and against what the mathematics fixes.

Stencils: the oracle's eight staggered difference sums (the ones its
velocity and stress updates use) are read through orc_probe_derivs and
compared with
  * the reference's update_v_PML / update_s_elastic_PML, which return the
    same sums unscaled when run once with DT = DH = 1 and unit materials
    (ref_probe_derivs), bit for bit, including orders 10 and 12 whose
    reference branches (src/update_v_PML.c:466-547,
    src/update_s_elastic_PML.c:442-535) differ from orders 2-8 only in the
    scaling and the missing CPML, not in the sums;
  * exactness on polynomials of degree <= FDORDER for the Taylor rows
    (MAXRELERROR 0), in the FP64 build, and inexactness at degree FDORDER+1;
  * Holberg's design bound: for MAXRELERROR 1-4 the phase-velocity error of
    the operator applied to sampled plane waves stays within E (0.1, 0.5, 1,
    3 %) up to hc[0] points per wavelength (the table's own rounding allows
    5 % of E).
Free surface: surface_elastic_PML and surface_PML of the reference against
the oracle on random fields, every order 2-12 and L 0-3, bit for bit.
PCG: PCG() of src/PCG.c, driven through its own state files, against
orc_pcg chained over iterations, bit for bit.
"""
import itertools
import os
import tempfile

import numpy as np
import pytest

import oracle_lib as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")

# derivative k of orc_probe_derivs: (field index in vx,vy,sxx,syy,sxy; axis 'x'/'y'; stagger offset)
DERIVS = [(2, "x", +0.5), (4, "x", -0.5), (4, "y", -0.5), (3, "y", +0.5),
          (0, "x", -0.5), (1, "x", +0.5), (0, "y", +0.5), (1, "y", -0.5)]


def probe_params(fd, e, NX=40, NY=34, FW=6):
    return O.make_params(NX=NX, NY=NY, FDORDER=fd, MAXRELERROR=e, FW=FW, NT=2, DH=1.0, DT=1.0, FREE_SURF=1)


def padded_coords(p):
    nd, (h, w) = O.probe_shape(p)
    i = np.arange(-nd + 1, p.NX + nd + 1, dtype=np.float64)
    j = np.arange(-nd + 1, p.NY + nd + 1, dtype=np.float64)
    assert len(i) == w and len(j) == h
    return np.meshgrid(i, j)  # ii[j][i], jj[j][i]


# ----------------------------------------------------- against the reference --
@needs_ref
@pytest.mark.parametrize("fd,e", list(itertools.product([2, 4, 6, 8, 10, 12], [0, 1, 4])))
def test_difference_sums_match_reference(fd, e):
    p = probe_params(fd, e)
    nd, shp = O.probe_shape(p)
    rng = np.random.default_rng(fd * 7 + e)
    f = rng.standard_normal((5, *shp)).astype(np.float32)
    a = O.probe_derivs(p, f)
    b = O.probe_derivs(p, f, which="ref")
    rows = slice(0, p.NY - p.FW)  # the reference's bottom CPML strip modifies j > NY-FW at orders 2-8
    for k in range(8):
        np.testing.assert_array_equal(a[k][rows], b[k][rows], err_msg=f"difference {k}")
    assert np.abs(a[:, rows]).min() >= 0 and np.abs(a[:, rows]).max() > 0.1


@needs_ref
@pytest.mark.parametrize("fd,L", list(itertools.product([2, 4, 6, 8, 10, 12], [0, 1, 2, 3])))
def test_free_surface_matches_reference(fd, L):
    p = O.make_params(NX=36, NY=24, FDORDER=fd, MAXRELERROR=1, FW=6, NT=10, DH=2.0, DT=2.5e-4, FREE_SURF=1, L=L,
                      FL=[10.0, 40.0, 70.0][:max(L, 1)], F_REF=25.0, VPPML=3000.0, FPML=25.0)
    nd, shp = O.probe_shape(p)
    rng = np.random.default_rng(100 + fd + L)
    model = dict(vp=(3000 * (1 + 0.1 * rng.random((p.NY, p.NX)))).astype(np.float32),
                 vs=(1700 * (1 + 0.1 * rng.random((p.NY, p.NX)))).astype(np.float32),
                 rho=(2000 * (1 + 0.1 * rng.random((p.NY, p.NX)))).astype(np.float32))
    if L:
        model["taus"] = (0.1 * (1 + 0.2 * rng.random((p.NY, p.NX)))).astype(np.float32)
        model["taup"] = (0.08 * (1 + 0.2 * rng.random((p.NY, p.NX)))).astype(np.float32)
    f = rng.standard_normal((8, *shp)).astype(np.float32)
    prow = rng.standard_normal((max(L, 1), p.NX)).astype(np.float32)
    fa, pa = O.probe_surface(p, model, f, prow)
    fb, pb = O.probe_surface(p, model, f, prow, which="ref")
    np.testing.assert_array_equal(fa, fb)
    if L:
        np.testing.assert_array_equal(pa, pb)
    # the mirror rows: syy[1-m] = -syy[1+m], sxy[1-m] = -sxy[m], m = 1..FDORDER/2
    cols = slice(nd, nd + p.NX)  # interior columns 1..NX; row j = 1 sits at index nd
    for m in range(1, fd // 2 + 1):
        np.testing.assert_array_equal(fa[3][nd - m, cols], -fa[3][nd + m, cols])
        np.testing.assert_array_equal(fa[4][nd - m, cols], -fa[4][nd + m - 1, cols])
    assert not np.any(fa[3][nd, cols])


# ------------------------------------------------------ against mathematics --
@pytest.mark.parametrize("fd", [2, 4, 6, 8, 10, 12])
def test_taylor_operators_exact_on_polynomials(fd):
    """FDORDER-th order staggered differences (Taylor coefficients, FP64
    build) are exact for polynomials of degree <= FDORDER (the centred form
    also cancels the even degree FDORDER); the coefficients are FP32 values
    of the table, which bounds "exact" at 1e-6 relative.  At degree
    FDORDER+1 they are not: x^(FDORDER+1) about a staggered node has derivative
    0 there, and the operator returns its first non-vanishing moment
    2 sum_m c_m (m-1/2)^(FDORDER+1) instead."""
    p = probe_params(fd, 0)
    ii, jj = padded_coords(p)
    hc = O.holberg(fd, 0).astype(np.float64)
    ci, cj = np.meshgrid(np.arange(1, p.NX + 1, dtype=np.float64), np.arange(1, p.NY + 1, dtype=np.float64))
    s = 8.0
    rng = np.random.default_rng(fd)
    for deg in (fd - 1, fd):
        P = np.polynomial.Polynomial(rng.uniform(0.5, 1.5, deg + 1))
        dP = P.deriv()
        for axis in ("x", "y"):
            u = ii if axis == "x" else jj
            out = O.probe_derivs(p, np.repeat(P((u - 20.0) / s)[None], 5, axis=0), fp64=True)
            for k, (_, ax, off) in enumerate(DERIVS):
                if ax != axis:
                    assert np.all(out[k] == 0), (k, axis)
                    continue
                want = dP(((ci if ax == "x" else cj) + off - 20.0) / s) / s
                err = np.abs(out[k] - want).max() / np.abs(want).max()
                assert err < 1e-6, (k, axis, deg, err)
    moment = 2 * sum(hc[m] * (m - 0.5) ** (fd + 1) for m in range(1, fd // 2 + 1))
    assert abs(moment) > 1e-3
    for axis in ("x", "y"):
        for k, (_, ax, off) in enumerate(DERIVS):
            if ax != axis:
                continue
            node = 17.0 + off  # staggered node of cell 17
            u = ii if axis == "x" else jj
            out = O.probe_derivs(p, np.repeat(((u - node) ** (fd + 1))[None], 5, axis=0), fp64=True)
            got = out[k][16, 16]
            assert abs(got - moment) < 1e-6 * abs(moment) * 10 ** (fd / 2), (k, axis, got, moment)


@pytest.mark.parametrize("fd,e", list(itertools.product([2, 4, 6, 8, 10, 12], [1, 2, 3, 4])))
def test_holberg_operators_meet_design_error(fd, e):
    """Plane waves sin/cos(k x) through the oracle's operators: the numerical
    wavenumber D(k) = |response| satisfies |D/k - 1| <= E for k up to
    2 pi / hc[0] (Holberg 1987, the table's design band)."""
    E = {1: 1e-3, 2: 5e-3, 3: 1e-2, 4: 3e-2}[e]
    p = probe_params(fd, e)
    ppw = float(O.holberg(fd, e)[0])
    ii, jj = padded_coords(p)
    worst = 0.0
    for k in np.linspace(0.02, 2 * np.pi / ppw, 40):
        for axis in ("x", "y"):
            u = ii if axis == "x" else jj
            fs = np.repeat(np.sin(k * u)[None], 5, axis=0)
            fc = np.repeat(np.cos(k * u)[None], 5, axis=0)
            a = O.probe_derivs(p, fs, fp64=True)
            b = O.probe_derivs(p, fc, fp64=True)
            for kk, (_, ax, _) in enumerate(DERIVS):
                if ax != axis:
                    continue
                D = np.sqrt(a[kk] ** 2 + b[kk] ** 2)
                assert np.ptp(D) < 1e-9 * k
                worst = max(worst, abs(D.mean() / k - 1))
    assert worst <= 1.05 * E, worst
    assert worst >= 0.5 * E  # the band edge reaches the design error


# --------------------------------------------------------------------- PCG --
@needs_ref
@pytest.mark.parametrize("start", [0, 2])
def test_pcg_matches_reference(start):
    """PCG() over five iterations (src/PCG.c:27-560; Polak-Ribiere beta,
    the iter == 2 / iter > 2 rules, restart at PCG_iter_start, the beta < 0
    reset) against orc_pcg chained the same way, bit for bit."""
    rng = np.random.default_rng(11 + start)
    ny, nx, n = 23, 37, 5
    g = np.zeros((n, 3, ny, nx), np.float32)
    g[0] = rng.standard_normal((3, ny, nx))
    for it in range(1, n):
        g[it] = 0.7 * g[it - 1] + 0.5 * rng.standard_normal((3, ny, nx))
    g[3, 1] = 0.5 * g[2, 1]  # sum gp (gp - p_last) < 0: a negative beta for Vs at iteration 4
    C3 = [0.5, 2.0, 1.0]
    # the reference's file names live in 74-byte buffers (STRING_SIZE,
    # src/fd.h:22), so the state files go to a short directory
    with tempfile.TemporaryDirectory(prefix="pcg", dir="/tmp") as d:
        dirs, gps = O.ref_pcg(g, C3, start, os.path.join(d, "j"))
    for k in range(3):
        p_last = np.zeros((ny, nx), np.float32)
        c_last = np.zeros((ny, nx), np.float32)
        betas = []
        for it in range(1, n + 1):
            beta, gp, d = O.pcg(g[it - 1, k], C3[k], p_last, c_last, it, start)
            np.testing.assert_array_equal(gp, gps[it - 1, k], err_msg=f"gradp param {k} iter {it}")
            np.testing.assert_array_equal(d, dirs[it - 1, k], err_msg=f"direction param {k} iter {it}")
            betas.append(beta)
            p_last = gp
            if it > 1 and (it - start) >= 1:
                c_last = d
        if k == 1 and start == 0:
            assert betas[3] == 0.0 and betas[1] > 0
