"""Optimizer oracle (oracle/opt_oracle.c) against the reference's own
compiled calc_opt_step / solvelin, lbfgs_core and calc_mat_change_test
(oracle/_ref), bit for bit, and against FP64 closed forms: the parabola
vertex, the Polak-Ribiere beta, a dense two-loop L-BFGS recursion."""
import numpy as np
import pytest

import oracle_lib as O

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference tree absent)")


def random_steps(rng, n):
    out = []
    for _ in range(n):
        e = np.sort(rng.uniform(0.0, 0.05, 3)).astype(np.float32)
        e[0] = 0.0
        a, b, c = rng.uniform(0.5, 2.0), rng.uniform(-40, 10), rng.uniform(100, 900) * rng.choice([1, -1])
        L2 = (a + b * e.astype(np.float64) + c * e.astype(np.float64) ** 2).astype(np.float32)
        out.append((L2, e))
    return out


@needs_ref
def test_opt_step_matches_reference():
    rng = np.random.default_rng(1)
    for L2, e in random_steps(rng, 300):
        assert O.opt_step(L2, e) == O.opt_step(L2, e, which="ref"), (L2, e)


def test_opt_step_is_parabola_vertex():
    # a convex parabola through three points: the vertex, clipped to [epst1, epst3]
    for a, b, c in [(1.0, -20.0, 400.0), (2.0, -1.0, 30.0), (0.8, -5.0, 150.0)]:
        e = np.array([0.0, 0.02, 0.04], np.float32)
        L2 = (a + b * e.astype(np.float64) + c * e.astype(np.float64) ** 2).astype(np.float32)
        vertex = -b / (2 * c)
        want = min(max(vertex, 0.0), 0.04)
        assert abs(O.opt_step(L2, e) - want) < 2e-5 * max(1.0, abs(want)) + 1e-6
    # concave with L2[2] > L2[3]: take epst3
    e = np.array([0.0, 0.02, 0.04], np.float32)
    L2 = np.array([1.0, 0.9, 0.7], np.float32)
    assert O.opt_step(L2, e) == e[2]


def test_pcg_beta_and_direction():
    rng = np.random.default_rng(2)
    ny, nx = 30, 41
    g = rng.standard_normal((ny, nx)).astype(np.float32)
    pl = rng.standard_normal((ny, nx)).astype(np.float32)
    cl = rng.standard_normal((ny, nx)).astype(np.float32)
    C_ = 0.5
    gp64 = (np.float32(C_) * g).astype(np.float64)
    beta64 = max(0.0, float(np.sum(gp64 * (gp64 - pl)) / np.sum(pl.astype(np.float64) ** 2)))
    # first iteration: no conjugation
    beta, gp, d = O.pcg(g, C_, pl, cl, 1, 0)
    assert beta == 0.0 and np.array_equal(d, np.float32(C_) * g) and np.array_equal(gp, d)
    # iter == 2 uses the previous gradient, iter > 2 the previous direction
    beta, gp, d = O.pcg(g + 0.8 * pl, C_, pl, cl, 2, 0)
    gp64 = (np.float32(C_) * (g + 0.8 * pl).astype(np.float32)).astype(np.float64)
    beta64 = max(0.0, float(np.sum(gp64 * (gp64 - pl)) / np.sum(pl.astype(np.float64) ** 2)))
    assert abs(beta - beta64) < 1e-4 * abs(beta64) and beta > 0
    assert np.array_equal(d, gp + pl * np.float32(beta))
    beta, gp, d = O.pcg(g + 0.8 * pl, C_, pl, cl, 3, 0)
    assert np.array_equal(d, gp + cl * np.float32(beta))
    # iter > 2 right after a restart: beta computed, direction not conjugated (src/PCG.c:206-218)
    beta, gp, d = O.pcg(g + 0.8 * pl, C_, pl, cl, 5, 4)
    assert beta > 0 and np.array_equal(d, gp)
    # a negative beta (gp = pl/2: sum gp (gp - pl) < 0) resets the direction
    beta, gp, d = O.pcg(0.5 * pl, 1.0, pl, cl, 2, 0)
    assert beta == 0.0 and np.array_equal(d, 0.5 * pl)


def models(rng, ny=24, nx=33):
    vp = (3000 + 200 * rng.standard_normal((ny, nx))).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = (2200 + 50 * rng.standard_normal((ny, nx))).astype(np.float32)
    g = [rng.standard_normal((ny, nx)).astype(np.float32) * s for s in (1e-3, 2e-3, 5e-4)]
    return vp, vs, rho, g


MAT_CASES = [
    dict(),
    dict(eps_scale=0.2, VPUPPERLIM=3200.0, VPLOWERLIM=2900.0, VSUPPERLIM=1750.0, VSLOWERLIM=1600.0),
    dict(eps_scale=0.1, VP_VS_RATIO=1.9, RHOUPPERLIM=2250.0, RHOLOWERLIM=2150.0, iter=3, INV_RHO_ITER=1),
    dict(eps_scale=0.3, S=1, S_VP=2.0, S_VS=3.0, S_RHO=1.0),
    dict(iter=2, INV_VP_ITER=3, INV_VS_ITER=1, INV_RHO_ITER=5),
    dict(GRAD_METHOD=2, iter=4, LBFGS_iter_start=2, eps_scale=0.5),
]


@needs_ref
@pytest.mark.parametrize("kw", MAT_CASES)
def test_mat_change_matches_reference(kw):
    rng = np.random.default_rng(3)
    vp, vs, rho, g = models(rng)
    vp0, vs0, rho0 = vp * 1.01, vs * 0.99, rho * 1.005
    u = O.update_params(**kw)
    a = O.mat_change(u, vp, vs, rho, *g, vp0, vs0, rho0)
    b = O.mat_change(u, vp, vs, rho, *g, vp0, vs0, rho0, which="ref")
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)


def test_mat_change_scaling_closed_form():
    rng = np.random.default_rng(4)
    vp, vs, rho, g = models(rng)
    u = O.update_params(eps_scale=0.01)
    vp1, vs1, rho1, eps = O.mat_change(u, vp, vs, rho, *g, vp, vs, rho)
    for m, gg, e, m1 in ((vp, g[0], eps[0], vp1), (vs, g[1], eps[1], vs1), (rho, g[2], eps[2], rho1)):
        assert e == np.float32(np.float32(0.01) * (np.float32(m.max()) / np.float32(np.abs(gg).max())))
        np.testing.assert_array_equal(m1, m - e * gg)
        # the largest model change is 1 % of the largest model value
        assert abs(np.abs(m1 - m).max() / m.max() - 0.01) < 1e-5


def two_loop64(s, y, q, it, N):
    m = max(1, it - N)
    idx = lambda v: (v % N or N) - 1
    w = idx(it - 1)
    s, y, q = s.astype(np.float64), y.astype(np.float64), q.astype(np.float64).copy()
    rho = np.zeros(N)
    for v in range(m, it):
        rho[idx(v)] = 1.0 / np.dot(y[idx(v)], s[idx(v)])
    h0 = np.dot(y[w], s[w]) / np.dot(y[w], y[w])
    alpha = np.zeros(N)
    for v in range(it - 1, m - 1, -1):
        k = idx(v)
        alpha[k] = rho[k] * np.dot(s[k], q)
        q -= alpha[k] * y[k]
    r = h0 * q
    for v in range(m, it):
        k = idx(v)
        beta = rho[k] * np.dot(y[k], r)
        r += s[k] * (alpha[k] - beta)
    return r


@pytest.mark.parametrize("it", [2, 4, 6, 9])
def test_lbfgs_core(it):
    rng = np.random.default_rng(5 + it)
    N, n = 5, 400
    # a quadratic with SPD Hessian: y = H s
    Q = rng.standard_normal((n, n)) / np.sqrt(n)
    H = Q @ Q.T + np.eye(n)
    s = (rng.standard_normal((N, n)) * 0.1).astype(np.float32)
    y = (s.astype(np.float64) @ H).astype(np.float32)
    q = rng.standard_normal(n).astype(np.float32)
    rho0 = np.zeros(N, np.float32)
    # rho of the older pairs is set by earlier iterations: fill them like the reference would have
    for v in range(max(1, it - N), it - 1):
        k = (v % N or N) - 1
        rho0[k] = np.float32(1.0) / np.float32(np.dot(y[k].astype(np.float32), s[k].astype(np.float32)))
    r, rho, alpha, _ = O.lbfgs_core(it, s, y, rho0, q, np.zeros(N, np.float32))
    want = two_loop64(s, y, q, it, N)
    assert np.linalg.norm(r - want) / np.linalg.norm(want) < 1e-4
    if O.ref_available():
        r2, rho2, alpha2, q2 = O.lbfgs_core(it, s, y, rho0, q, np.zeros(N, np.float32), which="ref")
        np.testing.assert_array_equal(r, r2)
        np.testing.assert_array_equal(rho, rho2)
        np.testing.assert_array_equal(alpha, alpha2)
