"""The reference-signature entry points (include/ifos2d_dropin.h) driven the
way src/IFOS2D.c:1651-1803 drives its CPU kernels, on device-resident
Numerical-Recipes arrays, compared bit for bit with the golden vectors
written from the reference's own kernels."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import ifos_amd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
pytestmark = pytest.mark.gpu

FP = C.POINTER(C.c_float)
VP = C.c_void_p


class HostNR:
    """Host Numerical-Recipes matrix (src/util.c:184): rows of one block,
    the pointer handed out is offset by -nrl rows and -ncl columns."""

    def __init__(self, nrl, nrh, ncl, nch, dtype=np.float32):
        self.nrl, self.ncl = nrl, ncl
        self.a = np.zeros((nrh - nrl + 1, nch - ncl + 1), dtype)
        isz = self.a.itemsize
        base = self.a.ctypes.data
        nr, nc = self.a.shape
        self.rows = (C.c_void_p * nr)(*[base + (r * nc - ncl) * isz for r in range(nr)])
        self.ptr = C.c_void_p(C.addressof(self.rows) - nrl * C.sizeof(C.c_void_p))

    def __getitem__(self, jk):
        j, k = jk
        return self.a[j - self.nrl, k - self.ncl]

    def __setitem__(self, jk, v):
        j, k = jk
        self.a[j - self.nrl, k - self.ncl] = v


def host_vec1(values):
    """1-based host vector (src/util.c:95): pointer to element 0 = data - 1."""
    arr = np.ascontiguousarray(np.concatenate([[0.0], values]), dtype=np.float32)
    return arr, C.c_void_p(arr.ctypes.data)


def set_globals(lib, cfg):
    for name in ("DT", "DH", "F_REF"):
        C.c_float.in_dll(lib, name).value = getattr(cfg, name)
    for name in ("NX", "NY", "NT", "FDORDER", "FW", "FREE_SURF", "BOUNDARY", "L", "PARAMETERIZATION",
                 "VELOCITY", "ADJOINT_TYPE", "SOURCE_TYPE"):
        C.c_int.in_dll(lib, name).value = getattr(cfg, name)
    C.c_int.in_dll(lib, "SEISMO").value = 1


def materials(m):
    """av_rho / av_mue (src/av_rho.c:27-31, src/av_mue.c:45-62) on a
    periodic one-cell halo (src/matcopy_elastic.c), float32 semantics."""
    def wrap(a):
        return np.pad(a, 1, mode="wrap")
    rho, vs = wrap(m["rho"]), wrap(m["vs"])
    f64 = np.float64
    rjp = (1.0 / (0.5 * (rho[1:-1, 1:-1] + rho[2:, 1:-1]).astype(f64))).astype(np.float32)
    rip = (1.0 / (0.5 * (rho[1:-1, 1:-1] + rho[1:-1, 2:]).astype(f64))).astype(np.float32)
    mu = rho * vs * vs
    u1, u2, u3, u4 = mu[1:-1, 1:-1], mu[1:-1, 2:], mu[2:, 1:-1], mu[2:, 2:]
    s = (((1.0 / u1.astype(f64)) + (1.0 / u2.astype(f64))) + (1.0 / u3.astype(f64))) + (1.0 / u4.astype(f64))
    uipjp = (4.0 / s).astype(np.float32)
    uipjp[(u1 == 0) | (u2 == 0) | (u3 == 0) | (u4 == 0)] = 0
    return rip, rjp, uipjp


def test_dropin_forward_elastic_matches_reference(gpu_lib):
    lib = ifos_amd.lib()
    name = "el4_fs_v0"
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    meta = json.load(open(os.path.join(GOLDEN, name + ".json")))
    cd = dict(meta["config"])
    fl = cd.pop("FL")
    cfg = ifos_amd.make_config(FL=fl[:7], **cd)
    print(f"[drop-in forward] {name}", flush=True)
    set_globals(lib, cfg)
    NX, NY, NT, nd = cfg.NX, cfg.NY, cfg.NT, cfg.FDORDER // 2 + 1
    lib.ifos_dev_matrix.restype = C.c_void_p
    lib.ifos_dev_copy_array.argtypes = [C.c_void_p, C.c_void_p, C.c_int]

    def dmat(r0, r1, c0, c1):
        return C.c_void_p(lib.ifos_dev_matrix(r0, r1, c0, c1))

    def grid():
        return dmat(-nd + 1, NY + nd, -nd + 1, NX + nd)

    names = ["vx", "vy", "sxx", "syy", "sxy", "ux", "uy", "uxy", "uyx", "vxp1", "vyp1", "vxm1", "vym1",
             "pi", "u", "rho", "rip", "rjp", "uipjp"]
    D = {k: grid() for k in names}
    fw2 = 2 * cfg.FW
    for k in ("psi_sxx_x", "psi_sxy_x", "psi_vxx", "psi_vyx", "psi_vxxs"):
        D[k] = dmat(1, NY, 1, fw2)
    for k in ("psi_syy_y", "psi_sxy_y", "psi_vyy", "psi_vxy"):
        D[k] = dmat(1, fw2, 1, NX)

    def upload(key, interior):
        h = np.zeros((NY + 2 * nd, NX + 2 * nd), np.float32)
        h[nd:nd + NY, nd:nd + NX] = interior
        assert lib.ifos_dev_copy_array(D[key], h.ctypes.data, 0) == 0
        return h

    def download(key):
        h = np.zeros((NY + 2 * nd, NX + 2 * nd), np.float32)
        assert lib.ifos_dev_copy_array(D[key], h.ctypes.data, 1) == 0
        return h[nd:nd + NY, nd:nd + NX]

    true = {k: z["true_" + k] for k in ("vp", "vs", "rho")}
    rip, rjp, uipjp = materials(true)
    keep = [upload("pi", true["vp"]), upload("u", true["vs"]), upload("rho", true["rho"]),
            upload("rip", rip), upload("rjp", rjp), upload("uipjp", uipjp)]

    prof = np.zeros((12, fw2), np.float32)
    lib.ifos_pml_profiles(C.byref(cfg), prof.ctypes.data_as(FP))
    pv = [host_vec1(prof[k]) for k in range(12)]
    hc = np.zeros(7, np.float32)
    lib.ifos_holberg(cfg.FDORDER, cfg.MAXRELERROR, hc.ctypes.data_as(FP))

    src = z["srcpos"]
    nsrc = src.shape[0]
    srcpos_loc = HostNR(1, 8, 1, nsrc)
    for l in range(nsrc):
        for k in range(8):
            srcpos_loc[k + 1, l + 1] = src[l, k]
    w = np.zeros((nsrc, NT), np.float32)
    lib.ifos_wavelet(C.byref(cfg), nsrc, np.ascontiguousarray(src).ctypes.data_as(FP), w.ctypes.data_as(FP))
    signals = HostNR(1, nsrc, 1, NT)
    signals.a[:] = w
    rec = z["recpos"]
    ntr = rec.shape[0]
    recpos = HostNR(1, 3, 1, ntr, np.int32)
    for t in range(ntr):
        recpos[1, t + 1], recpos[2, t + 1], recpos[3, t + 1] = rec[t, 0], rec[t, 1], t + 1
    svx, svy = HostNR(1, ntr, 1, NT), HostNR(1, ntr, 1, NT)

    K = [p[1] for p in pv]
    for nt in range(1, NT + 1):
        lib.update_v_PML(1, NX, 1, NY, nt, D["vx"], D["vxp1"], D["vxm1"], D["vy"], D["vyp1"], D["vym1"], None, None,
                         D["sxx"], D["syy"], D["sxy"], D["rip"], D["rjp"], srcpos_loc.ptr, signals.ptr, signals.ptr,
                         nsrc, None, hc.ctypes.data_as(FP), 0, 0, *K, D["psi_sxx_x"], D["psi_syy_y"],
                         D["psi_sxy_y"], D["psi_sxy_x"])
        lib.update_s_elastic_PML(1, NX, 1, NY, D["vx"], D["vy"], D["ux"], D["uy"], D["uxy"], D["uyx"], D["sxx"],
                                 D["syy"], D["sxy"], D["pi"], D["u"], D["uipjp"], None, D["rho"],
                                 hc.ctypes.data_as(FP), 0, *K, D["psi_vxx"], D["psi_vyy"], D["psi_vxy"], D["psi_vyx"])
        if C.c_int.in_dll(lib, "SOURCE_TYPE").value == 1:
            lib.psource(nt, D["sxx"], D["syy"], None, srcpos_loc.ptr, signals.ptr, nsrc, 0)
        lib.surface_elastic_PML(1, D["vx"], D["vy"], D["sxx"], D["syy"], D["sxy"], None, D["pi"], D["u"], D["rho"],
                                hc.ctypes.data_as(FP), K[0], K[1], K[2], D["psi_vxxs"], D["ux"], D["uy"], D["uxy"],
                                None, None, None)
        lib.seismo_ssg(nt, ntr, recpos.ptr, svx.ptr, svy.ptr, None, None, None, None, D["vx"], D["vy"], None,
                       D["sxx"], D["syy"], None, D["pi"], D["u"], hc.ctypes.data_as(FP))
    assert lib.ifos_dev_sync() == 0
    np.testing.assert_array_equal(svx.a, z["seis_vx"])
    np.testing.assert_array_equal(svy.a, z["seis_vy"])
    fin = np.stack([download(k) for k in ("vx", "vy", "sxx", "syy", "sxy")])
    np.testing.assert_array_equal(fin, z["final5"])
    for k, v in D.items():
        lib.ifos_dev_free_array(v)
