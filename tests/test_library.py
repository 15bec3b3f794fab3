"""CPU-side checks of the product library: it loads, exports every function
declared in include/*.h, and its host setup routines (run once per model /
shot on the host) are bit-identical to the reference build."""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\(", text, flags=re.M):
            name = m.group(1)
            if name not in ("if", "while", "return", "sizeof", "defined"):
                names.add(name)
    return sorted(names)


def test_headers_declare_entry_points():
    names = declared_functions()
    for must in ("update_v_PML", "update_s_elastic_PML", "update_s_visc_PML", "surface_PML",
                 "surface_elastic_PML", "psource", "seismo_ssg", "ifos_engine_create", "ifos_engine_shot_gradient"):
        assert must in names


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = ifos_amd.lib()
    assert hasattr(lib, name), f"{name} declared in include/ but not exported"


def cfg_and_params(**kw):
    w = workloads.small_case(**kw)
    cfg = w.cfg
    p = O.make_params(FL=[cfg.FL[i] for i in range(1, 8)],
                      **{k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"})
    return w, cfg, p


@pytest.mark.parametrize("fd,e", [(2, 0), (4, 1), (8, 1), (12, 4), (6, 3)])
def test_holberg_bitexact(fd, e):
    hc = (C.c_float * 7)()
    ifos_amd.lib().ifos_holberg(fd, e, hc)
    np.testing.assert_array_equal(np.array(hc[:], np.float32), O.holberg(fd, e))


@pytest.mark.parametrize("FW,DH", [(10, 2.0), (20, 1.0), (30, 8.0)])
def test_pml_profiles_bitexact(FW, DH):
    w, cfg, p = cfg_and_params(NX=3 * FW, NY=2 * FW + 4, FW=FW)
    cfg.DH = DH
    p.DH = DH
    out = np.zeros((12, 2 * FW), np.float32)
    ifos_amd.lib().ifos_pml_profiles(C.byref(cfg), out.ctypes.data_as(C.POINTER(C.c_float)))
    np.testing.assert_array_equal(out, O.pml_pro(p))
    if O.ref_available():
        np.testing.assert_array_equal(out, O.pml_pro(p, "ref"))


@pytest.mark.parametrize("shape", [1, 2, 4, 5, 6, 8])
def test_wavelet_bitexact(shape):
    w, cfg, p = cfg_and_params()
    cfg.SOURCE_SHAPE = shape
    p.SOURCE_SHAPE = shape
    src = np.array([[5, 5, 0, 0.003, 25.0, 1.5, 0, 1], [6, 6, 0, 0.0, 40.0, -2.0, 0, 1]], np.float32)
    out = np.zeros((2, cfg.NT), np.float32)
    ifos_amd.lib().ifos_wavelet(C.byref(cfg), 2, src.ctypes.data_as(C.POINTER(C.c_float)),
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    np.testing.assert_array_equal(out, O.wavelet(p, src))


def test_engine_rejects_bad_config():
    cfg = ifos_amd.make_config(NX=50, NY=40, DH=1.0, DT=1e-4, NT=101, FDORDER=4, DTINV=2)
    with pytest.raises(ifos_amd.IfosError, match="multiple of DTINV"):
        ifos_amd.Engine(cfg)
    cfg = ifos_amd.make_config(NX=50, NY=40, DH=1.0, DT=1e-4, NT=100, FDORDER=5, DTINV=2)
    with pytest.raises(ifos_amd.IfosError, match="FDORDER"):
        ifos_amd.Engine(cfg)
    cfg = ifos_amd.make_config(NX=50, NY=40, DH=1.0, DT=1e-4, NT=100, FDORDER=4, DTINV=2, BOUNDARY=2)
    with pytest.raises(ifos_amd.IfosError, match="BOUNDARY"):
        ifos_amd.Engine(cfg)
    # periodic sides copy N = FDORDER/2 edge columns into each halo
    cfg = ifos_amd.make_config(NX=6, NY=40, DH=1.0, DT=1e-4, NT=100, FDORDER=8, FW=2, DTINV=2, BOUNDARY=1)
    with pytest.raises(ifos_amd.IfosError, match="NX >= FDORDER"):
        ifos_amd.Engine(cfg)


def test_mall_slots_rule():
    """ifos_engine_mall_slots: the largest slot count whose per-step working
    set (per slot 4 B x padded field x (5 + 3L), shared 4 B x padded field x
    (8 if L else 5)) fits 160 MB; host arithmetic only."""
    import ifos_amd
    from ifos_amd import workloads

    def rule(cfg):
        N = cfg.FDORDER // 2
        H = N + 1
        pitch = ((16 + cfg.NX + H) + 63) // 64 * 64
        field = 4.0 * pitch * (cfg.NY + 2 * H + 1)
        n = (160e6 - field * (8 if cfg.L else 5)) / (field * (5 + 3 * cfg.L))
        return max(1, min(4096, int(n)))

    for name, want in (("batched_aperture", 5), ("layered_visc", 1), ("precision_large", 1)):
        cfg = getattr(workloads, name)().cfg
        assert ifos_amd.mall_slots(cfg) == want == rule(cfg), name
    cfg = workloads.toy_elastic().cfg
    assert ifos_amd.mall_slots(cfg) == rule(cfg) > 16
    bad = ifos_amd.make_config(NX=0, NY=10, NT=10, FDORDER=4, DH=1.0, DT=1e-3)
    with pytest.raises(ifos_amd.IfosError):
        ifos_amd.mall_slots(bad)
