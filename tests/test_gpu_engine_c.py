"""tests/c/engine_gradient.c: a C client of the native engine and of the
source-time-function calls (built by __graft_entry__.build() as a shared
library, called in-process).

  * engine_gradient_run makes the calls that replace the reference's shot
    loop (src/IFOS2D.c:1076-2724) for two shots: the summed gradient equals
    the oracle's gradients of the two shots added in shot order in FP32, bit
    for bit, and the summed misfit and energy equal the oracle's to 1e-12;
  * engine_stf_run makes the STF step of src/stf.c:24 (forward run, fdlsq
    filter from the vertical traces, conv_FD of the wavelet, the wavelet
    given to the engine, forward run): its filter, wavelet and seismograms
    equal the same sequence through the Python mirror bit for bit."""
import ctypes as C
import os

import numpy as np
import pytest

import ifos_amd
from ifos_amd import stf as S
from ifos_amd import workloads
import oracle_lib as O

pytestmark = pytest.mark.gpu

SO = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "libengine_gradient.so")


def client():
    if not os.path.exists(SO):
        pytest.fail("tests/c/libengine_gradient.so is missing: run __graft_entry__.build()")
    lib = C.CDLL(SO)
    lib.engine_gradient_run.argtypes = [C.c_char_p, C.c_char_p]
    lib.engine_stf_run.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    return lib


def params_of(w):
    cfg = w.cfg
    return O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                         **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})


def write_input(path, w, shots, obs, model):
    cfg = w.cfg
    with open(path, "wb") as f:
        f.write(bytes(cfg))
        f.write(np.int32(len(shots)).tobytes())
        for (sp, rp), (ovx, ovy) in zip(shots, obs):
            sp = np.ascontiguousarray(np.asarray(sp, np.float32).reshape(-1, 8))
            rp = np.ascontiguousarray(np.asarray(rp, np.int32).reshape(-1, 2))
            f.write(np.array([len(sp), len(rp)], np.int32).tobytes())
            f.write(sp.tobytes())
            f.write(rp.tobytes())
            f.write(np.ascontiguousarray(ovx, np.float32).tobytes())
            f.write(np.ascontiguousarray(ovy, np.float32).tobytes())
        for k in ("vp", "vs", "rho") + (("taus", "taup") if cfg.L else ()):
            f.write(np.ascontiguousarray(model[k], np.float32).tobytes())


def test_c_client_shot_loop_matches_oracle(gpu_lib, tmp_path):
    w = workloads.small_case(seed=81, NX=96, NY=64, NT=180, FDORDER=8, L=2, VELOCITY=1, DTINV=3)
    cfg = w.cfg
    p = params_of(w)
    base = w.shots[0]
    sp2 = np.array(base.srcpos, np.float32).reshape(-1, 8).copy()
    sp2[:, 0] = cfg.NX - sp2[:, 0]
    shots = [(base.srcpos, base.recpos), (sp2, base.recpos)]
    obs, refs = [], []
    for sp, rp in shots:
        vx, vy, _ = O.forward(p, w.true, sp, rp)
        vx[:, 0] = 0.0
        vy[:, 0] = 0.0
        obs.append((vx, vy))
        refs.append(O.gradient(p, w.start, sp, rp, vx, vy))
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, w, shots, obs, w.start)
    print("[c engine client] two shots 96x64 FDORDER 8 L=2", flush=True)
    assert client().engine_gradient_run(str(inp).encode(), str(out).encode()) == 0
    raw = open(out, "rb").read()
    mis = np.frombuffer(raw[:16], np.float64)
    g = np.frombuffer(raw[16:], np.float32).reshape(3, cfg.NY, cfg.NX)
    for k, n in enumerate(("grad_vp", "grad_vs", "grad_rho")):
        want = (np.float32(0) + refs[0][n].astype(np.float32)) + refs[1][n].astype(np.float32)
        np.testing.assert_array_equal(g[k], want, err_msg=n)
    l2 = refs[0]["misfit"][3] + refs[1]["misfit"][3]
    en = refs[0]["misfit"][4] + refs[1]["misfit"][4]
    assert abs(mis[0] - l2) <= 1e-12 * l2
    assert abs(mis[1] - en) <= 1e-12 * en


def test_c_client_stf_step_matches_python(gpu_lib, tmp_path):
    w = workloads.small_case(seed=82, NX=120, NY=60, NT=400, FDORDER=4, L=0, VELOCITY=1, DTINV=2)
    cfg = w.cfg
    s = w.shots[0]
    sp = np.array(s.srcpos, np.float32).reshape(-1, 8)
    rp = np.array(s.recpos, np.int32).reshape(-1, 2)
    sig = O.wavelet(params_of(w), sp)
    h = np.zeros(cfg.NT)
    h[0], h[3], h[7] = 0.9, 0.25, -0.15
    sig_h = np.stack([np.convolve(x.astype(np.float64), h)[:cfg.NT] for x in sig]).astype(np.float32)
    eng = ifos_amd.Engine(cfg)
    try:
        eng.set_model(**w.true)
        eng.set_shot(sp, rp)
        eng.set_wavelets(sig_h)
        eng.forward()
        obs_vx, obs_vy = eng.seismograms()
        # the same sequence through the Python mirror
        para = "fdlsq:waterlevel=1e-6"
        eng.set_shot(sp, rp)
        eng.forward()
        _, syn_vy = eng.seismograms()
        stf, _ = S.fdlsq(para, obs_vy, syn_vy, cfg.DT)
        new = S.conv_fd(sig[0], stf)
        eng.set_wavelets(new[None, :])
        eng.forward()
        _, vy_py = eng.seismograms()
    finally:
        eng.close()
    inp, out = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(inp, w, [(sp, rp)], [(obs_vx, obs_vy)], w.true)
    print("[c engine client] STF step 120x60", flush=True)
    assert client().engine_stf_run(str(inp).encode(), para.encode(), str(out).encode()) == 0
    got = np.fromfile(out, np.float32)
    NT, ntr = cfg.NT, len(rp)
    np.testing.assert_array_equal(got[:NT], stf)
    np.testing.assert_array_equal(got[NT:2 * NT], new)
    np.testing.assert_array_equal(got[2 * NT:].reshape(ntr, NT), vy_py)
    # and the loop recovers the filtered wavelet (times 1/dt, the library's scaling)
    assert np.abs(new * cfg.DT - sig_h[0]).max() <= 5e-2 * np.abs(sig_h[0]).max()
