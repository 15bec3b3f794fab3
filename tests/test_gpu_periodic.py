"""Periodic sides (BOUNDARY 1) on the GPU.

The step kernels copy the edge columns into the opposite halo as the
reference's exchanges do on one rank (src/exchange_v.c:143-228,
src/exchange_s.c:145-229), and the CPML is off on the left and right edges.

  * forward seismograms, final wavefields and shot gradients equal the
    oracle's bit for bit, for every velocity-step shape, a persistent grid,
    grids whose last tile column is partly outside, and several shots per
    launch (the oracle's exchanges are pinned to the reference's own
    src/exchange_v.c / src/exchange_s.c in tests/test_periodic.py);
  * at a size the oracle would take minutes for (1100x500, launched
    eagerly), the property the mathematics fixes: on a model that does not
    vary along x, moving the source by k columns moves a full receiver row's
    seismograms and the final wavefields by k columns, bit for bit;
  * the one-kernel step refuses periodic sides (its tiles form their
    neighbours' velocities from stresses beyond the halo copies), and the
    autotuner does not pick it.
"""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O

pytestmark = pytest.mark.gpu


def params_of(cfg):
    return O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                         **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})


def periodic_case(seed, NX, NY, NT, FDORDER, L, FS, DTINV=3):
    w = workloads.small_case(seed=seed, NX=NX, NY=NY, NT=NT, FDORDER=FDORDER, L=L, VELOCITY=1, DTINV=DTINV,
                             FREE_SURF=FS, FW=10)
    w.cfg.BOUNDARY = 1
    sp = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8).copy()
    sp[:, 0] = 4  # next to the left edge: the wave crosses it within NT
    return w, sp


@pytest.mark.parametrize("FDORDER,L,FS,rows,grid,NX", [(4, 0, 1, 1, 0, 72), (8, 2, 1, 2, 0, 72),
                                                        (8, 2, 0, 4, 64, 128), (12, 2, 1, 2, 0, 90),
                                                        (6, 3, 1, 1, 32, 100), (2, 0, 0, 2, 0, 64)])
def test_periodic_forward_matches_oracle(gpu_lib, FDORDER, L, FS, rows, grid, NX):
    w, sp = periodic_case(400 + FDORDER + L, NX, 52, 150, FDORDER, L, FS)
    print(f"[periodic forward] FDORDER={FDORDER} L={L} FS={FS} rows={rows} grid={grid} NX={NX}", flush=True)
    p = params_of(w.cfg)
    rp = w.shots[0].recpos
    ovx, ovy, ofin = O.forward(p, w.true, sp, rp)
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.true)
    eng.set_shot(sp, rp)
    eng.set_velocity_rows(rows)
    eng.set_launch(1, grid)
    eng.forward()
    vx, vy = eng.seismograms()
    np.testing.assert_array_equal(vx, ovx, err_msg="vx")
    np.testing.assert_array_equal(vy, ovy, err_msg="vy")
    np.testing.assert_array_equal(eng.fields(), ofin, err_msg="fields")
    eng.close()


@pytest.mark.parametrize("FDORDER,L", [(8, 2), (4, 0)])
def test_periodic_gradient_matches_oracle(gpu_lib, FDORDER, L):
    w, sp = periodic_case(420 + FDORDER + L, 80, 50, 180, FDORDER, L, 1)
    print(f"[periodic gradient] FDORDER={FDORDER} L={L}", flush=True)
    p = params_of(w.cfg)
    rp = w.shots[0].recpos
    obs_vx, obs_vy, _ = O.forward(p, w.true, sp, rp)
    obs_vx[:, 0] = 0
    obs_vy[:, 0] = 0
    ref = O.gradient(p, w.start, sp, rp, obs_vx, obs_vy)
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.start)
    eng.set_shot(sp, rp)
    mis = eng.shot_gradient(obs_vx, obs_vy)
    g = eng.gradient()
    for k, n in enumerate(("grad_vp", "grad_vs", "grad_rho")):
        np.testing.assert_array_equal(g[k], ref[n], err_msg=n)
    assert abs(mis[0] - ref["misfit"][3]) <= 1e-12 * ref["misfit"][3]
    eng.close()


def test_periodic_several_shots_per_launch(gpu_lib):
    """Three shots, one near each edge and one inside, in one launch: each
    slot's seismograms equal that shot run alone by the oracle."""
    w, sp = periodic_case(431, 96, 48, 150, 8, 2, 1)
    print("[periodic batch] 3 slots", flush=True)
    p = params_of(w.cfg)
    rp = w.shots[0].recpos
    shots = []
    for x in (3, 50, 95):
        s = sp.copy()
        s[:, 0] = x
        shots.append((s, rp))
    eng = ifos_amd.Engine(w.cfg, slots=3)
    eng.set_model(**w.true)
    eng.set_shots(shots)
    eng.forward()
    vx, vy = eng.seismograms()  # the slots' traces in slot order
    n = len(rp)
    for k, (s, r) in enumerate(shots):
        ovx, ovy, _ = O.forward(p, w.true, s, r)
        np.testing.assert_array_equal(vx[k * n:(k + 1) * n], ovx, err_msg=f"vx slot {k}")
        np.testing.assert_array_equal(vy[k * n:(k + 1) * n], ovy, err_msg=f"vy slot {k}")
    eng.close()


def test_periodic_full_size_commutes_with_shifts(gpu_lib):
    NX, NY, k = 1100, 500, 377
    w = workloads.small_case(seed=440, NX=NX, NY=NY, NT=700, FDORDER=8, L=2, VELOCITY=1, DTINV=2, FREE_SURF=1, FW=20)
    w.cfg.BOUNDARY = 1
    model = {key: np.ascontiguousarray(np.repeat(v[:, :1], NX, axis=1)) for key, v in w.true.items()}
    sp = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8).copy()
    sp[:, 0], sp[:, 1] = 40, 30
    rec = np.stack([np.arange(1, NX + 1), np.full(NX, 25)], 1).astype(np.int32)
    print(f"[periodic shift] {NX}x{NY} NT=700 FDORDER 8 L=2, shift {k}", flush=True)
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**model)
    out = []
    for x in (sp[0, 0], (sp[0, 0] - 1 + k) % NX + 1):
        s = sp.copy()
        s[:, 0] = x
        eng.set_shot(s, rec)
        assert not eng.get_graphs()  # above the graph threshold: eager launches
        eng.forward()
        out.append((*eng.seismograms(), eng.fields()))
    eng.close()
    (vx, vy, f), (vx2, vy2, f2) = out
    assert np.abs(f).max() > 0
    np.testing.assert_array_equal(vx2, np.roll(vx, k, axis=0), err_msg="vx")
    np.testing.assert_array_equal(vy2, np.roll(vy, k, axis=0), err_msg="vy")
    np.testing.assert_array_equal(f2, np.roll(f, k, axis=2), err_msg="fields")


def test_periodic_one_kernel_step_refused(gpu_lib):
    w, sp = periodic_case(450, 72, 52, 60, 4, 0, 1)
    eng = ifos_amd.Engine(w.cfg)
    eng.set_model(**w.true)
    eng.set_shot(sp, w.shots[0].recpos)
    with pytest.raises(ifos_amd.IfosError):
        eng.set_fused(True)
    eng.autotune(reps=2)
    assert not eng.get_fused()
    eng.close()


@pytest.mark.parametrize("FDORDER,L", [(8, 2), (4, 0)])
def test_periodic_dropin_matches_oracle(gpu_lib, FDORDER, L):
    """The drop-in layer driven as src/IFOS2D.c drives its kernels, with the
    exchange_v / exchange_s replacements after each half step: forward
    seismograms and final wavefields, and the shot gradient, equal the
    oracle's bit for bit."""
    from ifos_amd.dropin import DropinShot
    lib = ifos_amd.lib()
    w, sp = periodic_case(460 + FDORDER + L, 80, 50, 150, FDORDER, L, 1)
    print(f"[periodic drop-in] FDORDER={FDORDER} L={L}", flush=True)
    p = params_of(w.cfg)
    rp = np.asarray(w.shots[0].recpos, np.int32)
    ovx, ovy, ofin = O.forward(p, w.true, sp, rp)
    run = DropinShot(lib, w.cfg, w.true)
    vx, vy = run.forward(sp, rp, snapshots=False)
    np.testing.assert_array_equal(vx, ovx, err_msg="vx")
    np.testing.assert_array_equal(vy, ovy, err_msg="vy")
    fin = np.stack([run.download(k) for k in ("vx", "vy", "sxx", "syy", "sxy")])
    np.testing.assert_array_equal(fin, ofin, err_msg="fields")
    run.close()
    ovx[:, 0] = 0
    ovy[:, 0] = 0
    ref = O.gradient(p, w.start, sp, rp, ovx, ovy)
    run = DropinShot(lib, w.cfg, w.start)
    vx, vy = run.forward(sp, rp, snapshots=True)
    dvx, _ = O.calc_res(p, vx, ovx)
    dvy, _ = O.calc_res(p, vy, ovy)
    g = run.adjoint_gradient(rp, dvx, dvy)
    run.close()
    for k, n in enumerate(("grad_vp", "grad_vs", "grad_rho")):
        np.testing.assert_array_equal(g[k], ref[n], err_msg=n)
