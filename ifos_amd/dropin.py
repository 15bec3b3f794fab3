"""Driver of the drop-in layer (include/ifos2d_dropin.h): the call
sequence of the reference's shot loop -- forward loop (src/IFOS2D.c:
1651-1871) with the snapshot store, adjoint loop (:2224-2419) with the
zero-lag correlation, chain rule (:2471-2611) -- made on device-resident
Numerical-Recipes arrays, as the reference driver makes it once its
allocations come from ifos_dev_matrix / ifos_dev_f3tensor.  The tests and
the bench use it; a C driver makes the same calls."""
import ctypes as C

import numpy as np

FP = C.POINTER(C.c_float)


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


MAXL = 8


class Materials(C.Structure):
    """struct ifos_materials of the library's host setup (av_rho, av_mue,
    av_tau, prepare_update_s on a periodic one-cell halo)."""
    _fields_ = [(n, FP) for n in ("rip", "rjp", "uipjp", "f", "g", "fipjp", "d", "e", "dip", "muss", "lamss",
                                  "mu_r", "pi_r", "muipjp_r", "tausipjp")] + \
               [(n, C.c_float * (MAXL + 1)) for n in ("eta", "bip", "bjm", "cip", "cjm")] + \
               [("sumu", C.c_float), ("sumpi", C.c_float)]


class DropinShot:
    """Device arrays and the reference's per-step call sequence for one shot."""

    def __init__(self, lib, cfg, model):
        self.lib, self.cfg = lib, cfg
        set_globals(lib, cfg)
        NX, NY, L = cfg.NX, cfg.NY, cfg.L
        self.nd = nd = cfg.FDORDER // 2 + 1
        lib.ifos_dev_matrix.restype = C.c_void_p
        lib.ifos_dev_f3tensor.restype = C.c_void_p
        lib.ifos_dev_copy_array.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
        self.fl = np.zeros(8, np.float32)
        for l in range(1, 8):
            self.fl[l] = cfg.FL[l]
        C.c_void_p.in_dll(lib, "FL").value = self.fl.ctypes.data
        self.D = D = {}
        names = ["vx", "vy", "sxx", "syy", "sxy", "ux", "uy", "uxy", "uyx", "vxp1", "vyp1", "vxm1", "vym1",
                 "pi", "u", "rho", "rip", "rjp", "uipjp", "wc", "wc_u", "wc_rho", "w_lam", "w_mu", "w_rho_s"]
        if L:
            names += ["fipjp", "f", "g", "taus", "taup"]
        for k in names:
            D[k] = self.dmat(-nd + 1, NY + nd, -nd + 1, NX + nd)
        fw2 = 2 * cfg.FW
        for k in ("psi_sxx_x", "psi_sxy_x", "psi_vxx", "psi_vyx", "psi_vxxs"):
            D[k] = self.dmat(1, NY, 1, fw2)
        for k in ("psi_syy_y", "psi_sxy_y", "psi_vyy", "psi_vxy"):
            D[k] = self.dmat(1, fw2, 1, NX)
        if L:
            for k in ("r", "p", "q", "d", "e", "dip"):
                D[k] = self.dten(1, L)
        nsnap = cfg.NT // cfg.DTINV
        for k in ("F_rx", "F_ry", "F_x", "F_y", "F_u"):
            D[k] = self.dten(1, nsnap)
        self.nsnap = nsnap
        self.keep = []
        m = Materials()
        f32 = {k: np.ascontiguousarray(model[k], dtype=np.float32) for k in model}
        ptr = lambda k: f32[k].ctypes.data_as(FP) if k in f32 else None
        rc = lib.ifos_setup_materials(C.byref(cfg), ptr("vp"), ptr("vs"), ptr("rho"), ptr("taus"), ptr("taup"),
                                        C.byref(m))
        if rc != 0:
            raise RuntimeError("ifos_setup_materials failed")
        plane = lambda a, k=0: np.ctypeslib.as_array(a, shape=((k + 1) * NY * NX,))[k * NY * NX:].reshape(NY, NX)
        self.upload("pi", f32["vp"])
        self.upload("u", f32["vs"])
        self.upload("rho", f32["rho"])
        for k in ("rip", "rjp", "uipjp"):
            self.upload(k, plane(getattr(m, k)))
        if L:
            self.upload("taus", f32["taus"])
            self.upload("taup", f32["taup"])
            for k in ("fipjp", "f", "g"):
                self.upload(k, plane(getattr(m, k)))
            for k in ("d", "e", "dip"):
                t = np.zeros((NY + 2 * nd, NX + 2 * nd, L), np.float32)
                for l in range(L):
                    t[nd:nd + NY, nd:nd + NX, l] = plane(getattr(m, k), l)
                self.keep.append(t)
                self._check(lib.ifos_dev_copy_array(D[k], t.ctypes.data, 0), "upload " + k)
            self.vec = {k: host_vec1(np.array(getattr(m, k)[1:L + 1], np.float32))
                        for k in ("bip", "bjm", "cip", "cjm", "eta")}
            self.etajm = host_vec1(np.array(m.eta[1:L + 1], np.float32))
        lib.ifos_free_materials(C.byref(m))
        prof = np.zeros((12, fw2), np.float32)
        lib.ifos_pml_profiles(C.byref(cfg), prof.ctypes.data_as(FP))
        self.pv = [host_vec1(prof[k]) for k in range(12)]
        self.K = [p[1] for p in self.pv]
        self.hc = np.zeros(7, np.float32)
        lib.ifos_holberg(cfg.FDORDER, cfg.MAXRELERROR, self.hc.ctypes.data_as(FP))

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"drop-in {what}: {self.lib.ifos_last_error().decode()}")

    def dmat(self, r0, r1, c0, c1):
        return C.c_void_p(self.lib.ifos_dev_matrix(r0, r1, c0, c1))

    def dten(self, l0, l1):
        nd, cfg = self.nd, self.cfg
        return C.c_void_p(self.lib.ifos_dev_f3tensor(-nd + 1, cfg.NY + nd, -nd + 1, cfg.NX + nd, l0, l1))

    def upload(self, key, interior):
        nd, NX, NY = self.nd, self.cfg.NX, self.cfg.NY
        h = np.zeros((NY + 2 * nd, NX + 2 * nd), np.float32)
        h[nd:nd + NY, nd:nd + NX] = interior
        self.keep.append(h)
        self._check(self.lib.ifos_dev_copy_array(self.D[key], h.ctypes.data, 0), "upload " + key)

    def download(self, key):
        nd, NX, NY = self.nd, self.cfg.NX, self.cfg.NY
        h = np.zeros((NY + 2 * nd, NX + 2 * nd), np.float32)
        self._check(self.lib.ifos_dev_copy_array(self.D[key], h.ctypes.data, 1), "download " + key)
        return h[nd:nd + NY, nd:nd + NX]

    def zero(self):
        D, cfg, lib = self.D, self.cfg, self.lib
        nd = self.nd
        args = [-nd + 1, cfg.NY + nd, -nd + 1, cfg.NX + nd, D["vx"], D["vy"], None, D["sxx"], D["syy"], D["sxy"],
                None, None, D["vxm1"], D["vym1"], D["uxy"], D["vxp1"], D["vyp1"], D["psi_sxx_x"], D["psi_sxy_x"],
                None, D["psi_vxx"], D["psi_vyx"], None, D["psi_syy_y"], D["psi_sxy_y"], None, D["psi_vyy"],
                D["psi_vxy"], None, D["psi_vxxs"]]
        if cfg.L:
            lib.zero_fdveps_visc(*args, D["r"], D["p"], D["q"], None, None)
        else:
            lib.zero_fdveps(*args)

    def step(self, nt, srcpos, signals, signals1, nsig, sw):
        """update_v, exchange_v, update_s, psource, surface, exchange_s:
        src/IFOS2D.c:1682-1785 (sw 0) and :2253-2366 (sw 1)"""
        lib, D, cfg, K = self.lib, self.D, self.cfg, self.K
        NX, NY = cfg.NX, cfg.NY
        hc = self.hc.ctypes.data_as(FP)
        lib.update_v_PML(1, NX, 1, NY, nt, D["vx"], D["vxp1"], D["vxm1"], D["vy"], D["vyp1"], D["vym1"], None, None,
                         D["sxx"], D["syy"], D["sxy"], D["rip"], D["rjp"], srcpos, signals, signals1, nsig, None, hc,
                         0, sw, *K, D["psi_sxx_x"], D["psi_syy_y"], D["psi_sxy_y"], D["psi_sxy_x"])
        lib.exchange_v(D["vx"], D["vy"], None, None, None, None, None, None, None, 1)
        if cfg.L:
            v = self.vec
            lib.update_s_visc_PML(1, NX, 1, NY, D["vx"], D["vy"], D["ux"], D["uy"], D["uxy"], D["uyx"], D["sxx"],
                                  D["syy"], D["sxy"], D["pi"], D["u"], D["uipjp"], D["rho"], hc, 0, D["r"], D["p"],
                                  D["q"], D["fipjp"], D["f"], D["g"], v["bip"][1], v["bjm"][1], v["cip"][1],
                                  v["cjm"][1], D["d"], D["e"], D["dip"], *K, D["psi_vxx"], D["psi_vyy"],
                                  D["psi_vxy"], D["psi_vyx"])
        else:
            lib.update_s_elastic_PML(1, NX, 1, NY, D["vx"], D["vy"], D["ux"], D["uy"], D["uxy"], D["uyx"], D["sxx"],
                                     D["syy"], D["sxy"], D["pi"], D["u"], D["uipjp"], None, D["rho"], hc, 0, *K,
                                     D["psi_vxx"], D["psi_vyy"], D["psi_vxy"], D["psi_vyx"])
        if sw == 0 and C.c_int.in_dll(lib, "SOURCE_TYPE").value == 1:
            lib.psource(nt, D["sxx"], D["syy"], None, srcpos, signals, nsig, 0)
        if cfg.FREE_SURF:
            if cfg.L:
                lib.surface_PML(1, D["vx"], D["vy"], D["sxx"], D["syy"], D["sxy"], None, D["p"], D["q"], D["pi"],
                                D["u"], D["rho"], D["taup"], D["taus"], self.etajm[1], self.vec["eta"][1], hc, K[0],
                                K[1], K[2], D["psi_vxxs"], D["ux"], D["uy"], D["uxy"], None, None, None)
            else:
                lib.surface_elastic_PML(1, D["vx"], D["vy"], D["sxx"], D["syy"], D["sxy"], None, D["pi"], D["u"],
                                        D["rho"], hc, K[0], K[1], K[2], D["psi_vxxs"], D["ux"], D["uy"], D["uxy"],
                                        None, None, None)
        lib.exchange_s(D["sxx"], D["syy"], D["sxy"], None, None, None, None, None, None, None, None, 1)

    def snaps(self):
        D = self.D
        return [D["F_rx"], D["F_ry"], D["F_x"], D["F_y"], D["F_u"]]

    def forward(self, src, rec, snapshots):
        lib, D, cfg = self.lib, self.D, self.cfg
        NT = cfg.NT
        nsrc = src.shape[0]
        srcpos_loc = HostNR(1, 8, 1, nsrc)
        for l in range(nsrc):
            for k in range(8):
                srcpos_loc[k + 1, l + 1] = src[l, k]
        w = np.zeros((nsrc, NT), np.float32)
        lib.ifos_wavelet(C.byref(cfg), nsrc, np.ascontiguousarray(src).ctypes.data_as(FP), w.ctypes.data_as(FP))
        signals = HostNR(1, nsrc, 1, NT)
        signals.a[:] = w
        ntr = rec.shape[0]
        recpos = HostNR(1, 3, 1, ntr, np.int32)
        for t in range(ntr):
            recpos[1, t + 1], recpos[2, t + 1], recpos[3, t + 1] = rec[t, 0], rec[t, 1], t + 1
        svx, svy = HostNR(1, ntr, 1, NT), HostNR(1, ntr, 1, NT)
        self.zero()
        hin, hin1 = 1, 1
        for nt in range(1, NT + 1):
            self.step(nt, srcpos_loc.ptr, signals.ptr, signals.ptr, nsrc, 0)
            lib.seismo_ssg(nt, ntr, recpos.ptr, svx.ptr, svy.ptr, None, None, None, None, D["vx"], D["vy"], None,
                           D["sxx"], D["syy"], None, D["pi"], D["u"], self.hc.ctypes.data_as(FP))
            if snapshots and nt == hin1:
                lib.ifos_snapshot_store(hin, D["vxp1"], D["vyp1"], D["sxx"], D["syy"], D["sxy"], D["ux"], D["uy"],
                                        D["uxy"], *self.snaps())
                hin += 1
                hin1 += cfg.DTINV
        self._check(lib.ifos_dev_sync(), "sync")
        return svx.a.copy(), svy.a.copy()

    def adjoint_gradient(self, rec, dvx, dvy):
        lib, D, cfg = self.lib, self.D, self.cfg
        NT, ntr = cfg.NT, rec.shape[0]
        back = HostNR(1, 6, 1, ntr)
        for t in range(ntr):
            back[1, t + 1], back[2, t + 1] = rec[t, 0], rec[t, 1]
        sx, sy = HostNR(1, ntr, 1, NT), HostNR(1, ntr, 1, NT)
        sx.a[:] = dvx
        sy.a[:] = dvy
        self.zero()
        hin = 1
        snap_step = set(range(1, NT + 1, cfg.DTINV))
        for nt in range(1, NT + 1):
            self.step(nt, back.ptr, sx.ptr, sy.ptr, ntr, 1)
            if NT - nt + 1 in snap_step:
                lib.ifos_correlate(self.nsnap - hin + 1, D["vxp1"], D["vyp1"], D["sxx"], D["syy"], D["sxy"],
                                   D["rho"], D["pi"], D["u"], *self.snaps(), D["wc"], D["wc_u"], D["wc_rho"])
                hin += 1
        lib.ifos_gradient_assemble(D["rho"], D["pi"], D["u"], D["wc"], D["wc_u"], D["wc_rho"], D["w_lam"],
                                   D["w_mu"], D["w_rho_s"])
        self._check(lib.ifos_dev_sync(), "sync")
        return [self.download(k) for k in ("wc", "wc_u", "wc_rho")]

    def close(self):
        for v in self.D.values():
            self.lib.ifos_dev_free_array(v)


