"""The reference-signature entry points (include/ifos2d_dropin.h) driven the
way src/IFOS2D.c drives its CPU kernels -- forward loop (:1651-1871) with
the snapshot store, adjoint loop (:2224-2419) with the zero-lag
correlation, chain rule (:2471-2611) -- on device-resident
Numerical-Recipes arrays, compared bit for bit with the golden vectors
written from the reference's own kernels, elastic and viscoelastic."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
from ifos_amd.dropin import DropinShot
import oracle_lib as O
from test_precision import params_of

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = ["el4_fs_v0", "visc8_fs_v1", "el2_nofs_rot", "visc6_fs_fx_v1"]
pytestmark = pytest.mark.gpu


def load_case(name):
    z = dict(np.load(os.path.join(GOLDEN, name + ".npz")))
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    cd = dict(meta["config"])
    fl = cd.pop("FL")
    cfg = ifos_amd.make_config(FL=fl[:7], **cd)
    keys = ("vp", "vs", "rho", "taus", "taup")
    model = lambda pre: {k: z[pre + k] for k in keys if pre + k in z}
    return cfg, z, model("true_"), model("start_")


@pytest.mark.parametrize("name", CASES)
def test_dropin_forward_matches_reference(gpu_lib, name):
    lib = ifos_amd.lib()
    cfg, z, true, _ = load_case(name)
    print(f"[drop-in forward] {name}", flush=True)
    run = DropinShot(lib, cfg, true)
    vx, vy = run.forward(z["srcpos"], z["recpos"], snapshots=False)
    np.testing.assert_array_equal(vx, z["seis_vx"])
    np.testing.assert_array_equal(vy, z["seis_vy"])
    fin = np.stack([run.download(k) for k in ("vx", "vy", "sxx", "syy", "sxy")])
    np.testing.assert_array_equal(fin, z["final5"])
    run.close()


@pytest.mark.parametrize("name", CASES)
def test_dropin_gradient_matches_reference(gpu_lib, name):
    """Forward with ifos_snapshot_store, residual (oracle calc_res, the
    reference's host code), adjoint with ifos_correlate, then
    ifos_gradient_assemble: the shot's Vp, Vs and rho gradients."""
    lib = ifos_amd.lib()
    cfg, z, _, start = load_case(name)
    print(f"[drop-in gradient] {name}", flush=True)
    run = DropinShot(lib, cfg, start)
    vx, vy = run.forward(z["srcpos"], z["recpos"], snapshots=True)
    np.testing.assert_array_equal(vx, z["start_seis_vx"])
    p = O.make_params(**{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    dvx, _ = O.calc_res(p, vx, z["obs_vx"])
    dvy, _ = O.calc_res(p, vy, z["obs_vy"])
    g = run.adjoint_gradient(z["recpos"], dvx, dvy)
    for k, n in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(g[k], z["grad_" + n], err_msg=n)
    run.close()


def test_c_client_forward_matches_oracle(gpu_lib, tmp_path):
    """tests/c/dropin_forward.c: the reference's forward loop written in C
    against include/ifos2d_dropin.h (built by __graft_entry__.build() as a
    shared library and called in-process); seismograms and final wavefields
    bit for bit against the oracle."""
    so = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "libdropin_forward.so")
    if not os.path.exists(so):
        pytest.fail("tests/c/libdropin_forward.so is missing: run __graft_entry__.build()")
    client = C.CDLL(so)
    client.dropin_forward_run.argtypes = [C.c_char_p, C.c_char_p]
    w = workloads.small_case(seed=71, NX=80, NY=56, NT=160, FDORDER=4, L=0, VELOCITY=1, DTINV=1, FW=10, stype=1)
    cfg, s = w.cfg, w.shots[0]
    src = np.ascontiguousarray(np.array(s.srcpos, np.float32).reshape(-1, 8))
    rec = np.ascontiguousarray(np.array(s.recpos, np.int32).reshape(-1, 2))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([cfg.NX, cfg.NY, cfg.NT, cfg.FDORDER, cfg.FW, cfg.FREE_SURF, cfg.VELOCITY, len(src), len(rec)],
                         np.int32).tobytes())
        f.write(np.array([cfg.DT, cfg.DH, cfg.VPPML, cfg.FPML, cfg.npower, cfg.k_max_PML], np.float32).tobytes())
        f.write(src.tobytes())
        f.write(rec.tobytes())
        for k in ("vp", "vs", "rho"):
            f.write(np.ascontiguousarray(w.true[k], np.float32).tobytes())
    print("[c client] forward 80x56 FDORDER 4", flush=True)
    assert client.dropin_forward_run(str(inp).encode(), str(out).encode()) == 0
    got = np.fromfile(out, np.float32)
    ntr, NT, n = len(rec), cfg.NT, cfg.NX * cfg.NY
    vx = got[:ntr * NT].reshape(ntr, NT)
    vy = got[ntr * NT:2 * ntr * NT].reshape(ntr, NT)
    fin = got[2 * ntr * NT:].reshape(5, cfg.NY, cfg.NX)
    ovx, ovy, ofin = O.forward(params_of(w), w.true, s.srcpos, s.recpos)
    np.testing.assert_array_equal(vx, ovx)
    np.testing.assert_array_equal(vy, ovy)
    np.testing.assert_array_equal(fin, ofin)


@pytest.mark.parametrize("name", ["el4_fs_v0", "el2_nofs_rot", "visc6_fs_fx_v1"])
def test_source_series_changed_between_steps(gpu_lib, name):
    """update_v_PML and psource keep a device copy of the host signals and
    check at every step the samples that step reads (psource reads columns
    nt-1..nt+1, update_v_PML column nt).  Signals overwritten in place after
    step k with a series equal to the first up to column k+2 must give the
    same bits as running with the second series from the start: the first
    step that reads a changed sample (k+2) has to see it."""
    from ifos_amd.dropin import HostNR
    cfg, z, true, _ = load_case(name)
    lib = ifos_amd.lib()
    src = np.ascontiguousarray(z["srcpos"], np.float32).reshape(-1, 8)
    nsrc, NT = src.shape[0], cfg.NT
    k = NT // 2
    s1 = np.zeros((nsrc, NT), np.float32)
    lib.ifos_wavelet(C.byref(cfg), nsrc, src.ctypes.data_as(C.POINTER(C.c_float)),
                     s1.ctypes.data_as(C.POINTER(C.c_float)))
    s2 = s1.copy()
    s2[:, k + 2:] *= np.float32(-0.7)  # 1-based columns k+3.. differ
    srcpos_loc = HostNR(1, 8, 1, nsrc)
    for l in range(nsrc):
        for c in range(8):
            srcpos_loc[c + 1, l + 1] = src[l, c]
    out = []
    for in_place in (True, False):
        print(f"[drop-in series] {name} in place {in_place}", flush=True)
        run = DropinShot(lib, cfg, true)
        signals = HostNR(1, nsrc, 1, NT)
        signals.a[:] = s1 if in_place else s2
        run.zero()
        for nt in range(1, NT + 1):
            if in_place and nt == k + 1:
                signals.a[:] = s2
            run.step(nt, srcpos_loc.ptr, signals.ptr, signals.ptr, nsrc, 0)
        assert lib.ifos_dev_sync() == 0
        out.append([run.download(f) for f in ("vx", "vy", "sxx", "syy", "sxy")])
        run.close()
    for a, b, f in zip(out[0], out[1], ("vx", "vy", "sxx", "syy", "sxy")):
        np.testing.assert_array_equal(a, b, err_msg=f)
    assert np.abs(out[1][1]).max() > 0
