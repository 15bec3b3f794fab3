"""Device optimizer (include/ifos2d_opt.h) against the oracle's FP32
restatement (oracle/opt_oracle.c, itself bit-identical to the reference's
calc_opt_step / calc_mat_change_test / lbfgs_core) and FP64 numpy:
- dot products equal the FP64 sum to 1e-12 relative, maxima exactly;
- the PCG beta agrees with the oracle's FP32 beta to 1e-5 and the direction
  is exactly gp + last*beta in FP32;
- the model update equals the oracle bit for bit (maxima are exact);
- the step length agrees with the oracle to 1e-4 relative (FP64 vs FP32
  Gauss elimination);
- L-BFGS directions agree with the oracle to 1e-4 relative L2."""
import numpy as np
import pytest

import ifos_amd
import oracle_lib as O
from ifos_amd.opt import Optimizer, update_params

pytestmark = pytest.mark.gpu

DB = ifos_amd.DeviceBuffer


def dev(a):
    return DB.from_array(np.ascontiguousarray(a, np.float32))


def test_reductions(gpu_lib):
    rng = np.random.default_rng(1)
    opt = Optimizer()
    for n in (1, 255, 1000, 123457, 1 << 20):
        a = rng.standard_normal(n).astype(np.float32)
        b = rng.standard_normal(n).astype(np.float32)
        da, db = dev(a), dev(b)
        want = float(np.dot(a.astype(np.float64), b.astype(np.float64)))
        got = opt.dot(da, db, n)
        assert abs(got - want) <= 1e-12 * max(1.0, np.abs(a).dot(np.abs(b)))
        assert opt.dot(da, db, n) == got  # deterministic
        assert opt.max(da, n, absolute=True) == np.abs(a).max()
        assert opt.max(da, n) == a.max()
    opt.close()


def test_pcg_matches_oracle(gpu_lib):
    rng = np.random.default_rng(2)
    ny, nx = 37, 53
    n = ny * nx
    opt = Optimizer()
    pl = rng.standard_normal((ny, nx)).astype(np.float32)
    cl = rng.standard_normal((ny, nx)).astype(np.float32)
    for it, start in [(1, 0), (2, 0), (3, 0), (5, 4), (6, 4)]:
        g = (rng.standard_normal((ny, nx)) + 0.7 * pl).astype(np.float32)
        beta_o, gp_o, d_o = O.pcg(g, 0.25, pl, cl, it, start)
        dpl, dcl, dd = dev(pl), dev(cl), DB(4 * n)
        beta = opt.pcg(n, dev(g), 0.25, dpl, dcl, it, start, dd)
        d = dd.to_array((ny, nx))
        if beta_o == 0.0:
            assert beta == 0.0
        else:
            assert abs(beta - beta_o) <= 1e-5 * abs(beta_o)
        gp = np.float32(0.25) * g
        use1 = it - start >= 1
        use2 = it - start >= 2
        conj = it > 1 and use1
        form = 0 if not conj else (1 if it == 2 else (2 if it > 2 and use2 else 0))
        want = gp if form == 0 else gp + (pl if form == 1 else cl) * np.float32(beta)
        np.testing.assert_array_equal(d, want)
        # state updates: p_last <- gp; c_last <- direction when beta was formed
        np.testing.assert_array_equal(dpl.to_array((ny, nx)), gp)
        np.testing.assert_array_equal(dcl.to_array((ny, nx)), want if conj else cl)
    opt.close()


def test_step_length_matches_oracle(gpu_lib):
    rng = np.random.default_rng(3)
    opt = Optimizer()
    for _ in range(200):
        e = np.sort(rng.uniform(0.0, 0.05, 3)).astype(np.float32)
        e[0] = 0.0
        a, b, c = rng.uniform(0.5, 2.0), rng.uniform(-40, 10), rng.uniform(100, 900) * rng.choice([1, -1])
        L2 = (a + b * e.astype(np.float64) + c * e.astype(np.float64) ** 2).astype(np.float32)
        got = opt.step_length(L2, e)
        want = O.opt_step(L2, e)
        assert abs(got - want) <= 1e-4 * max(abs(want), 1e-3), (L2, e, got, want)
    opt.close()


@pytest.mark.parametrize("kw", [dict(), dict(eps_scale=0.2, VPUPPERLIM=3200.0, VPLOWERLIM=2900.0,
                                            VSUPPERLIM=1750.0, VSLOWERLIM=1600.0),
                                dict(eps_scale=0.1, VP_VS_RATIO=1.9, RHOUPPERLIM=2250.0, RHOLOWERLIM=2150.0,
                                     iter=3, INV_RHO_ITER=1),
                                dict(eps_scale=0.3, S=1, S_VP=2.0, S_VS=3.0, S_RHO=1.0),
                                dict(GRAD_METHOD=2, iter=4, LBFGS_iter_start=2, eps_scale=0.5)])
def test_model_update_matches_oracle(gpu_lib, kw):
    rng = np.random.default_rng(4)
    ny, nx = 40, 61
    vp = (3000 + 200 * rng.standard_normal((ny, nx))).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = (2200 + 50 * rng.standard_normal((ny, nx))).astype(np.float32)
    g = [rng.standard_normal((ny, nx)).astype(np.float32) * s for s in (1e-3, 2e-3, 5e-4)]
    m0 = [vp * np.float32(1.01), vs * np.float32(0.99), rho * np.float32(1.005)]
    want = O.mat_change(O.update_params(**kw), vp, vs, rho, *g, *m0)
    opt = Optimizer()
    outs = [DB(4 * ny * nx) for _ in range(3)]
    eps = opt.model_update(ny * nx, update_params(**kw), *[dev(a) for a in (vp, vs, rho, *g, *m0)], *outs)
    np.testing.assert_array_equal(eps, want[3])
    for k in range(3):
        np.testing.assert_array_equal(outs[k].to_array((ny, nx)), want[k])
    opt.close()


def test_lbfgs_matches_oracle(gpu_lib):
    """Three L-BFGS iterations on a quadratic; the device two-loop (FP64
    dots) against the oracle's lbfgs_core (FP32, the reference's)."""
    rng = np.random.default_rng(5)
    ny, nx, npar, N = 12, 17, 3, 5
    ncell = ny * nx
    n = npar * ncell
    Q = rng.standard_normal((n, n)) / np.sqrt(n)
    H = Q @ Q.T + np.eye(n)
    avg = np.array([1800.0, 2100.0, 3200.0], np.float32)
    opt = Optimizer()
    s_hist, y_hist, work = DB(4 * N * n), DB(4 * N * n), DB(4 * 2 * n)
    s_hist.upload(np.zeros(N * n, np.float32))
    y_hist.upload(np.zeros(N * n, np.float32))
    rho_hist = np.zeros(N, np.float64)
    m = rng.standard_normal(n).astype(np.float32)
    s_o = np.zeros((N, n), np.float32)
    y_o = np.zeros((N, n), np.float32)
    rho_o = np.zeros(N, np.float32)
    start = 1
    for it in range(1, 5):
        grad = (H @ m.astype(np.float64)).astype(np.float32)   # normalised model space
        gpar = [grad[k * ncell:(k + 1) * ncell] / avg[k] for k in range(npar)]
        dg = [dev(x) for x in gpar]
        opt.lbfgs(ncell, dg, avg, s_hist, y_hist, rho_hist, work, N, it, start)
        d = np.concatenate([dg[k].to_array(ncell) for k in range(npar)])
        # oracle: same bookkeeping as src/LBFGS.c around lbfgs_core
        q = np.concatenate([gpar[k] * avg[k] for k in range(npar)]).astype(np.float32)
        if it == start:
            w = (it % N or N) - 1
            y_o[w] = -q
            want = np.concatenate(gpar)
        else:
            w = ((it - 1) % N or N) - 1
            y_o[w] = y_o[w] + q
            r, rho_o, _, _ = O.lbfgs_core(it, s_o, y_o, rho_o, q, np.zeros(N, np.float32))
            w2 = (it % N or N) - 1
            y_o[w2] = -q
            want = np.concatenate([r[k * ncell:(k + 1) * ncell] * avg[k] for k in range(npar)])
        assert np.linalg.norm(d - want) / np.linalg.norm(want) < 1e-4, it
        # model step and the s store of src/calc_mat_change_test.c:297-316
        step = -0.05 * np.concatenate([d[k * ncell:(k + 1) * ncell] for k in range(npar)])
        m_new = (m + step).astype(np.float32)
        m0 = [dev(m[k * ncell:(k + 1) * ncell] * avg[k]) for k in range(npar)]
        m1 = [dev(m_new[k * ncell:(k + 1) * ncell] * avg[k]) for k in range(npar)]
        opt.lbfgs_store_s(ncell, m1, m0, avg, s_hist, N, it)
        w = (it % N or N) - 1
        s_o[w] = np.concatenate([(m_new[k * ncell:(k + 1) * ncell] * avg[k] - m[k * ncell:(k + 1) * ncell] * avg[k])
                                 / avg[k] for k in range(npar)]).astype(np.float32)
        np.testing.assert_array_equal(s_hist.to_array(N * n).reshape(N, n)[w], s_o[w])
        m = m_new
    opt.close()
