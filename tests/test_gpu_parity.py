"""GPU parity: the gfx950 engine against the reference's own kernels (golden
vectors written from oracle/_ref) and against the oracle restatement.

The bar is bit-identical FP32 results: the kernels keep the reference's
operand order and promotions and are compiled without FMA contraction.
Misfit sums are FP64 on the device (the reference sums in FP32); they are
compared with an FP64 numpy sum to 1e-12 relative.
"""
import json
import os

import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["el4_fs_v0", "visc8_fs_v1", "el2_nofs_rot", "visc6_fs_fx_v1"]

pytestmark = pytest.mark.gpu


def load_case(name):
    z = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    cfgd = dict(meta["config"])
    fl = cfgd.pop("FL")
    cfg = ifos_amd.make_config(FL=fl[:7], **cfgd)
    keys = ("vp", "vs", "rho", "taus", "taup")
    model = lambda pre: {k: z[pre + k] for k in keys if pre + k in z}
    return cfg, z, model("true_"), model("start_")


def assert_bitexact(a, b, what):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, what
    bad = np.flatnonzero(~((a == b) | (np.isnan(a) & np.isnan(b))))
    if bad.size:
        k = bad[0]
        rel = np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-30)
        raise AssertionError(f"{what}: {bad.size} of {a.size} values differ; first at {k}: "
                             f"{a.ravel()[k]!r} vs {b.ravel()[k]!r}; rel L2 {rel:.3e}")


@pytest.mark.parametrize("name", CASES)
def test_forward_matches_reference(gpu_lib, name):
    cfg, z, true, _ = load_case(name)
    print(f"[forward] case {name}", flush=True)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**true)
    eng.set_shot(z["srcpos"], z["recpos"])
    eng.forward(snapshots=False)
    vx, vy = eng.seismograms()
    assert_bitexact(vx, z["seis_vx"], "seismogram vx")
    assert_bitexact(vy, z["seis_vy"], "seismogram vy")
    assert_bitexact(eng.fields(), z["final5"], "final wavefields")
    eng.close()


@pytest.mark.parametrize("name", CASES)
def test_gradient_matches_reference(gpu_lib, name):
    cfg, z, _, start = load_case(name)
    print(f"[gradient] case {name}", flush=True)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**start)
    eng.set_shot(z["srcpos"], z["recpos"])
    eng.reset_gradient()
    mis = eng.shot_gradient(z["obs_vx"], z["obs_vy"])
    vx, vy = eng.seismograms()
    assert_bitexact(vx, z["start_seis_vx"], "start-model seismogram vx")
    g = eng.gradient()
    assert_bitexact(g[0], z["grad_vp"], "Vp gradient")
    assert_bitexact(g[1], z["grad_vs"], "Vs gradient")
    assert_bitexact(g[2], z["grad_rho"], "rho gradient")
    # FP64 misfit of the integrated residual (src/calc_misfit.c:193-194)
    p = O.make_params(**{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    dvx, _ = O.calc_res(p, z["start_seis_vx"], z["obs_vx"])
    dvy, _ = O.calc_res(p, z["start_seis_vy"], z["obs_vy"])
    l2 = float(np.sum(dvx.astype(np.float64) ** 2) + np.sum(dvy.astype(np.float64) ** 2))
    assert abs(mis[0] - l2) <= 1e-12 * l2
    tot = eng.gradient(total=True)
    assert_bitexact(tot, g, "running gradient sum after one shot")
    eng.close()


def test_graph_replay_equals_eager(gpu_lib, monkeypatch):
    cfg, z, true, start = load_case("visc8_fs_v1")
    print("[graph vs eager] visc8_fs_v1", flush=True)
    out = []
    for mode in ("0", "1"):
        monkeypatch.setenv("IFOS_NO_GRAPH", mode)
        eng = ifos_amd.Engine(cfg)
        eng.set_model(**start)
        eng.set_shot(z["srcpos"], z["recpos"])
        eng.shot_gradient(z["obs_vx"], z["obs_vy"])
        out.append((eng.seismograms(), eng.gradient()))
        eng.close()
    assert_bitexact(out[0][0][0], out[1][0][0], "graph vs eager seismograms")
    assert_bitexact(out[0][1], out[1][1], "graph vs eager gradient")


def test_graph_mode_by_size(gpu_lib, monkeypatch):
    """Captured graphs below 2^19 cells x slots, eager launches above; the
    setter overrides the choice, and the environment sets it at creation."""
    monkeypatch.delenv("IFOS_NO_GRAPH", raising=False)
    small = ifos_amd.make_config(NX=400, NY=200, NT=20, FDORDER=4, DH=10.0, DT=1e-3, FW=10)
    large = ifos_amd.make_config(NX=1024, NY=600, NT=20, FDORDER=4, DH=10.0, DT=1e-3, FW=10)
    sp = np.array([[50, 20, 0, 0, 10, 1e10, 0, 1]], np.float32)
    rp = np.array([[60, 30]], np.int32)
    want = {}
    for name, cfg in (("small", small), ("large", large)):
        eng = ifos_amd.Engine(cfg)
        m = np.full((cfg.NY, cfg.NX), 3000.0, np.float32)
        eng.set_model(m, m / 1.7, m * 0.7)
        eng.set_shot(sp, rp)
        want[name] = eng.get_graphs()
        eng.set_graphs(1)
        assert eng.get_graphs()
        eng.set_graphs(0)
        assert not eng.get_graphs()
        eng.close()
    assert want == {"small": True, "large": False}
    monkeypatch.setenv("IFOS_NO_GRAPH", "0")
    eng = ifos_amd.Engine(large)
    m = np.full((large.NY, large.NX), 3000.0, np.float32)
    eng.set_model(m, m / 1.7, m * 0.7)
    eng.set_shot(sp, rp)
    assert eng.get_graphs()
    eng.close()


@pytest.mark.parametrize("FDORDER,L,VEL,FS,stype", [(10, 0, 0, 1, 1), (12, 2, 1, 1, 1), (8, 3, 1, 0, 3),
                                                    (4, 2, 0, 1, 2)])
def test_forward_matches_oracle(gpu_lib, FDORDER, L, VEL, FS, stype):
    w = workloads.small_case(seed=11 + FDORDER + L, NX=72, NY=52, NT=150, FDORDER=FDORDER, L=L, VELOCITY=VEL,
                             DTINV=3, FREE_SURF=FS, stype=stype, FW=12)
    print(f"[oracle forward] FDORDER={FDORDER} L={L} VEL={VEL} FS={FS} type={stype}", flush=True)
    cfg = w.cfg
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                      **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    shot = w.shots[0]
    ovx, ovy, ofin = O.forward(p, w.true, shot.srcpos, shot.recpos)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.true)
    eng.set_shot(shot.srcpos, shot.recpos)
    eng.forward()
    vx, vy = eng.seismograms()
    assert_bitexact(vx, ovx, "vx vs oracle")
    assert_bitexact(vy, ovy, "vy vs oracle")
    assert_bitexact(eng.fields(), ofin, "fields vs oracle")
    eng.close()


@pytest.mark.parametrize("kmax,npower", [(4.0, 2.0), (2.5, 3.0)])
def test_stretched_pml_matches_oracle(gpu_lib, kmax, npower):
    """CPML with a coordinate stretch K > 1 (k_max_PML, src/PML_pro.c:
    73-119): the memory-variable updates divide by K, a division the
    kernels skip when every K is one."""
    w = workloads.small_case(seed=17, NX=70, NY=50, NT=160, FDORDER=6, L=1, VELOCITY=1, DTINV=2, FW=12)
    print(f"[oracle stretched pml] kmax={kmax} npower={npower}", flush=True)
    cfg = w.cfg
    cfg.k_max_PML = kmax
    cfg.npower = npower
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                      **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    shot = w.shots[0]
    ovx, ovy, ofin = O.forward(p, w.true, shot.srcpos, shot.recpos)
    ref = O.gradient(p, w.start, shot.srcpos, shot.recpos, ovx, ovy)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.true)
    eng.set_shot(shot.srcpos, shot.recpos)
    eng.forward()
    vx, vy = eng.seismograms()
    assert_bitexact(vx, ovx, "vx vs oracle (stretched PML)")
    assert_bitexact(vy, ovy, "vy vs oracle (stretched PML)")
    assert_bitexact(eng.fields(), ofin, "fields vs oracle (stretched PML)")
    eng.set_model(**w.start)
    eng.shot_gradient(ovx, ovy)
    g = eng.gradient()
    assert_bitexact(g[0], ref["grad_vp"], "Vp gradient (stretched PML)")
    assert_bitexact(g[1], ref["grad_vs"], "Vs gradient (stretched PML)")
    assert_bitexact(g[2], ref["grad_rho"], "rho gradient (stretched PML)")
    eng.close()


@pytest.mark.parametrize("adj", [2, 3])
def test_adjoint_type_matches_oracle(gpu_lib, adj):
    """ADJOINT_TYPE 2 (vy residuals only) and 3 (vx only)."""
    w = workloads.small_case(seed=19 + adj, NX=70, NY=48, NT=150, FDORDER=6, L=2, VELOCITY=1, DTINV=3, FW=10)
    print(f"[oracle gradient] ADJOINT_TYPE {adj}", flush=True)
    cfg = w.cfg
    cfg.ADJOINT_TYPE = adj
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                      **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    shot = w.shots[0]
    ovx, ovy, _ = O.forward(p, w.true, shot.srcpos, shot.recpos)
    ovx[:, 0] = 0
    ovy[:, 0] = 0
    ref = O.gradient(p, w.start, shot.srcpos, shot.recpos, ovx, ovy)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.start)
    eng.set_shot(shot.srcpos, shot.recpos)
    mis = eng.shot_gradient(ovx, ovy)
    g = eng.gradient()
    assert_bitexact(g[0], ref["grad_vp"], "Vp gradient")
    assert_bitexact(g[1], ref["grad_vs"], "Vs gradient")
    assert_bitexact(g[2], ref["grad_rho"], "rho gradient")
    assert abs(mis[0] - ref["misfit"][3]) <= 1e-12 * ref["misfit"][3]
    eng.close()


def test_gradient_matches_oracle_with_kill(gpu_lib):
    w = workloads.small_case(seed=5, NX=80, NY=50, NT=180, FDORDER=8, L=2, VELOCITY=1, DTINV=3)
    print("[oracle gradient] kill mask", flush=True)
    cfg = w.cfg
    cfg.TRKILL = 1
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                      **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    shot = w.shots[0]
    obs_vx, obs_vy, _ = O.forward(p, w.true, shot.srcpos, shot.recpos)
    obs_vx[:, 0] = 0
    obs_vy[:, 0] = 0
    kill = np.zeros(len(shot.recpos), np.int32)
    kill[::3] = 1
    ref = O.gradient(p, w.start, shot.srcpos, shot.recpos, obs_vx, obs_vy, kill=kill)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**w.start)
    eng.set_shot(shot.srcpos, shot.recpos)
    mis = eng.shot_gradient(obs_vx, obs_vy, kill=kill)
    g = eng.gradient()
    assert_bitexact(g[0], ref["grad_vp"], "Vp gradient (kill)")
    assert_bitexact(g[1], ref["grad_vs"], "Vs gradient (kill)")
    assert_bitexact(g[2], ref["grad_rho"], "rho gradient (kill)")
    assert abs(mis[0] - ref["misfit"][3]) <= 1e-12 * ref["misfit"][3]
    assert abs(mis[1] - ref["misfit"][4]) <= 1e-12 * ref["misfit"][4]
    eng.close()


def test_two_shots_sum(gpu_lib):
    cfg, z, true, start = load_case("el4_fs_v0")
    print("[two shots]", flush=True)
    eng = ifos_amd.Engine(cfg)
    eng.set_model(**start)
    eng.reset_gradient()
    src2 = z["srcpos"].copy()
    src2[0, 0] += 7
    gs = []
    for src in (z["srcpos"], src2):
        eng.set_shot(src, z["recpos"])
        eng.shot_gradient(z["obs_vx"], z["obs_vy"])
        gs.append(eng.gradient())
    tot = eng.gradient(total=True)
    expect = (np.float32(0) + gs[0]) + gs[1]
    assert_bitexact(tot, expect.astype(np.float32), "sum over shots")
    eng.close()
