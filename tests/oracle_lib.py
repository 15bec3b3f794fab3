"""ctypes access to the checkers under oracle/ (test infrastructure).

liboracle.so / liboracle64.so are the CPU restatement (oracle/ifos_oracle.c);
_ref/libifosref.so drives the reference's own compiled hot-path files
(oracle/ref_harness.c).  Only tests, smoke() and bench.py's cpu_baseline leg
use this module.
"""
import ctypes as C
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_INT_FIELDS = ["NX", "NY", "FDORDER", "MAXRELERROR", "FW", "FREE_SURF", "BOUNDARY", "L",
               "VELOCITY", "ADJOINT_TYPE", "PARAMETERIZATION", "DTINV", "NT",
               "SOURCE_SHAPE", "SOURCE_TYPE", "TRKILL"]
_FLT_FIELDS = ["DH", "DT", "VPPML", "FPML", "npower", "k_max_PML", "F_REF", "TAU"]


class OrcParams(C.Structure):
    _fields_ = ([(n, C.c_int) for n in _INT_FIELDS] + [(n, C.c_float) for n in _FLT_FIELDS]
                + [("FL", C.c_float * 8), ("SNAP_PREC", C.c_int), ("EPRECOND", C.c_int),
                   ("EPRECOND_PER_SHOT", C.c_int), ("EPSILON_WE", C.c_float), ("EPRECOND_FP64", C.c_int)])


def make_params(**kw):
    p = OrcParams()
    defaults = dict(MAXRELERROR=1, FW=10, FREE_SURF=1, BOUNDARY=0, L=0, VELOCITY=0,
                    ADJOINT_TYPE=1, PARAMETERIZATION=1, DTINV=1, SOURCE_SHAPE=1, SOURCE_TYPE=1,
                    TRKILL=0, VPPML=3500.0, FPML=30.0, npower=2.0, k_max_PML=1.0, F_REF=0.0, TAU=0.0)
    defaults.update(kw)
    fl = defaults.pop("FL", None)
    for k, v in defaults.items():
        setattr(p, k, v)
    if fl is not None:
        for i, v in enumerate(fl):
            p.FL[i + 1] = v
    return p


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _ptr(a, ty=C.c_float):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ty))


class _Lib:
    def __init__(self, path):
        self.lib = C.CDLL(path)


_ORC = {}


def oracle(fp64=False):
    if fp64 not in _ORC:
        path = os.path.join(ORACLE_DIR, "liboracle64.so" if fp64 else "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle not built: run make -C oracle")
        lib = C.CDLL(path)
        lib.orc_calc_res.restype = C.c_double
        _ORC[fp64] = lib
    return _ORC[fp64]


def ref_available():
    return os.path.exists(os.path.join(ORACLE_DIR, "_ref", "libifosref.so"))


_REF = None


def ref():
    global _REF
    if _REF is None:
        _REF = C.CDLL(os.path.join(ORACLE_DIR, "_ref", "libifosref.so"))
    return _REF


def _ip_fp(p):
    base = C.addressof(p)
    ip = C.cast(base, C.POINTER(C.c_int))
    fp = C.cast(base + 4 * len(_INT_FIELDS), C.POINTER(C.c_float))
    fl = C.cast(base + 4 * (len(_INT_FIELDS) + len(_FLT_FIELDS)), C.POINTER(C.c_float))
    return ip, fp, fl


# ---------------------------------------------------------------------------
def holberg(FDORDER, MAXRELERROR, which="oracle"):
    out = np.zeros(7, np.float32)
    if which == "oracle":
        oracle().orc_holbergcoeff(C.c_int(FDORDER), C.c_int(MAXRELERROR), _ptr(out))
    else:
        p = make_params(NX=10, NY=10, FDORDER=FDORDER, MAXRELERROR=MAXRELERROR, NT=1, DH=1.0, DT=1e-4)
        ip, fp, _ = _ip_fp(p)
        ref().ref_holbergcoeff(ip, fp, _ptr(out))
    return out


def pml_pro(p, which="oracle"):
    out = np.zeros((12, 2 * p.FW), np.float32)
    if which == "oracle":
        oracle().orc_pml_pro(C.byref(p), _ptr(out))
    else:
        ip, fp, _ = _ip_fp(p)
        ref().ref_pml_pro(ip, fp, _ptr(out))
    return out


def wavelet(p, srcpos, which="oracle"):
    srcpos = _f32(srcpos).reshape(-1, 8)
    n = srcpos.shape[0]
    out = np.zeros((n, p.NT), np.float32)
    if which == "oracle":
        oracle().orc_wavelet(C.byref(p), C.c_int(n), _ptr(srcpos), _ptr(out))
    else:
        ip, fp, _ = _ip_fp(p)
        ref().ref_wavelet(ip, fp, C.c_int(n), _ptr(srcpos), _ptr(out))
    return out


def forward(p, model, srcpos, recpos, which="oracle", fp64=False):
    """Returns (seis_vx, seis_vy, final5) with final5 = vx,vy,sxx,syy,sxy interiors."""
    vp, vs, rho = (_f32(model[k]) for k in ("vp", "vs", "rho"))
    taus = _f32(model["taus"]) if p.L else None
    taup = _f32(model["taup"]) if p.L else None
    srcpos = _f32(srcpos).reshape(-1, 8)
    recpos = np.ascontiguousarray(recpos, dtype=np.int32).reshape(-1, 2)
    ntr = recpos.shape[0]
    dt = np.float64 if fp64 else np.float32
    ty = C.c_double if fp64 else C.c_float
    svx = np.zeros((ntr, p.NT), dt)
    svy = np.zeros((ntr, p.NT), dt)
    fin = np.zeros((5, p.NY, p.NX), dt)
    if which == "oracle":
        rc = oracle(fp64).orc_forward(C.byref(p), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                                      C.c_int(srcpos.shape[0]), _ptr(srcpos), C.c_int(ntr),
                                      _ptr(recpos, C.c_int), _ptr(svx, ty), _ptr(svy, ty), _ptr(fin, ty))
    else:
        ip, fp, fl = _ip_fp(p)
        rc = ref().ref_forward(ip, fp, fl, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                               C.c_int(srcpos.shape[0]), _ptr(srcpos), C.c_int(ntr),
                               _ptr(recpos, C.c_int), _ptr(svx), _ptr(svy), _ptr(fin))
    assert rc == 0, rc
    return svx, svy, fin


def calc_res(p, syn, obs, kill=None):
    syn = _f32(syn)
    obs = _f32(obs)
    ntr = syn.shape[0]
    diff = np.zeros_like(syn)
    k = None if kill is None else np.ascontiguousarray(kill, dtype=np.int32)
    l2 = oracle().orc_calc_res(C.byref(p), C.c_int(ntr), _ptr(syn), _ptr(obs), _ptr(k, C.c_int), _ptr(diff))
    return diff, l2


def gradient(p, model, srcpos, recpos, obs_vx, obs_vy, kill=None, fp64=False):
    """Oracle forward+adjoint+gradient.  Returns dict."""
    vp, vs, rho = (_f32(model[k]) for k in ("vp", "vs", "rho"))
    taus = _f32(model["taus"]) if p.L else None
    taup = _f32(model["taup"]) if p.L else None
    srcpos = _f32(srcpos).reshape(-1, 8)
    recpos = np.ascontiguousarray(recpos, dtype=np.int32).reshape(-1, 2)
    ntr = recpos.shape[0]
    dt = np.float64 if fp64 else np.float32
    ty = C.c_double if fp64 else C.c_float
    svx = np.zeros((ntr, p.NT), dt)
    svy = np.zeros((ntr, p.NT), dt)
    g = np.zeros((3, p.NY, p.NX), dt)
    mis = np.zeros(5, np.float64)
    k = None if kill is None else np.ascontiguousarray(kill, dtype=np.int32)
    rc = oracle(fp64).orc_gradient(C.byref(p), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                                   C.c_int(srcpos.shape[0]), _ptr(srcpos), C.c_int(ntr), _ptr(recpos, C.c_int),
                                   _ptr(_f32(obs_vx)), _ptr(_f32(obs_vy)), _ptr(k, C.c_int),
                                   _ptr(svx, ty), _ptr(svy, ty), _ptr(g[0], ty), _ptr(g[1], ty), _ptr(g[2], ty),
                                   _ptr(mis, C.c_double))
    assert rc == 0, rc
    return dict(seis_vx=svx, seis_vy=svy, grad_vp=g[0], grad_vs=g[1], grad_rho=g[2], misfit=mis)


def ref_gradient(p, model, srcpos, recpos, diff_vx, diff_vy):
    vp, vs, rho = (_f32(model[k]) for k in ("vp", "vs", "rho"))
    taus = _f32(model["taus"]) if p.L else None
    taup = _f32(model["taup"]) if p.L else None
    srcpos = _f32(srcpos).reshape(-1, 8)
    recpos = np.ascontiguousarray(recpos, dtype=np.int32).reshape(-1, 2)
    ntr = recpos.shape[0]
    svx = np.zeros((ntr, p.NT), np.float32)
    svy = np.zeros((ntr, p.NT), np.float32)
    g = np.zeros((3, p.NY, p.NX), np.float32)
    ip, fp, fl = _ip_fp(p)
    rc = ref().ref_gradient(ip, fp, fl, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                            C.c_int(srcpos.shape[0]), _ptr(srcpos), C.c_int(ntr), _ptr(recpos, C.c_int),
                            _ptr(_f32(diff_vx)), _ptr(_f32(diff_vy)), _ptr(svx), _ptr(svy),
                            _ptr(g[0]), _ptr(g[1]), _ptr(g[2]))
    assert rc == 0, rc
    return dict(seis_vx=svx, seis_vy=svy, grad_vp=g[0], grad_vs=g[1], grad_rho=g[2])


def f32_to_16(x, prec):
    """16-bit snapshot codes of the oracle (prec 1 BF16, 2 FP16)."""
    L = oracle()
    L.orc_f32_to_16.restype = C.c_ushort
    L.orc_f32_to_16.argtypes = [C.c_float, C.c_int]
    return np.array([L.orc_f32_to_16(float(v), prec) for v in np.asarray(x, np.float32).ravel()], np.uint16)


def f16_to_32(h, prec):
    L = oracle()
    L.orc_16_to_f32.restype = C.c_float
    L.orc_16_to_f32.argtypes = [C.c_ushort, C.c_int]
    return np.array([L.orc_16_to_f32(int(v), prec) for v in np.asarray(h, np.uint16).ravel()], np.float32)


# ---------------------------------------------------------------- optimizer --
class UpdateParams(C.Structure):
    _fields_ = ([(n, C.c_int) for n in ("iter", "INV_VP_ITER", "INV_VS_ITER", "INV_RHO_ITER", "S", "GRAD_METHOD",
                                        "LBFGS_STEP_LENGTH", "LBFGS_iter_start")]
                + [(n, C.c_float) for n in ("eps_scale", "VPUPPERLIM", "VPLOWERLIM", "VSUPPERLIM", "VSLOWERLIM",
                                            "RHOUPPERLIM", "RHOLOWERLIM", "VP_VS_RATIO", "S_VP", "S_VS", "S_RHO")])


def update_params(**kw):
    d = dict(iter=1, INV_VP_ITER=0, INV_VS_ITER=0, INV_RHO_ITER=0, S=0, GRAD_METHOD=1, LBFGS_STEP_LENGTH=1,
             LBFGS_iter_start=1, eps_scale=0.01, VPUPPERLIM=1e9, VPLOWERLIM=0.0, VSUPPERLIM=1e9, VSLOWERLIM=0.0,
             RHOUPPERLIM=1e9, RHOLOWERLIM=0.0, VP_VS_RATIO=0.0, S_VP=0.0, S_VS=0.0, S_RHO=0.0)
    d.update(kw)
    u = UpdateParams()
    for k, v in d.items():
        setattr(u, k, v)
    return u


def pcg(grad, C_, p_last, c_last, it, start, which="oracle"):
    ny, nx = grad.shape
    gp = np.zeros((ny, nx), np.float32)
    d = np.zeros((ny, nx), np.float32)
    L = oracle()
    L.orc_pcg.restype = C.c_float
    beta = L.orc_pcg(nx, ny, _ptr(_f32(grad)), C.c_float(C_), _ptr(_f32(p_last)), _ptr(_f32(c_last)), it, start,
                     _ptr(gp), _ptr(d))
    return beta, gp, d


def opt_step(L2t, epst, which="oracle"):
    a, b = _f32(np.asarray(L2t)), _f32(np.asarray(epst))
    if which == "ref":
        L = ref()
        L.ref_calc_opt_step.restype = C.c_float
        return L.ref_calc_opt_step(_ptr(a), _ptr(b))
    L = oracle()
    L.orc_opt_step.restype = C.c_float
    return L.orc_opt_step(_ptr(a), _ptr(b))


def mat_change(u, vp, vs, rho, g_vp, g_vs, g_rho, vp0, vs0, rho0, which="oracle"):
    ny, nx = vp.shape
    arrs = [_f32(a) for a in (vp, vs, rho, g_vp, g_vs, g_rho, vp0, vs0, rho0)]
    out = [np.zeros((ny, nx), np.float32) for _ in range(3)]
    eps = np.zeros(3, np.float32)
    if which == "ref":
        ip = (C.c_int * 8)(*[getattr(u, k) for k, _ in u._fields_[:8]])
        fp = (C.c_float * 11)(*[getattr(u, k) for k, _ in u._fields_[8:]])
        ref().ref_mat_change(nx, ny, ip, fp, *[_ptr(a) for a in arrs], *[_ptr(o) for o in out], _ptr(eps))
    else:
        oracle().orc_mat_change(nx, ny, C.byref(u), *[_ptr(a) for a in arrs], *[_ptr(o) for o in out], _ptr(eps))
    return out[0], out[1], out[2], eps


def lbfgs_core(iteration, s, y, rho, q, alpha, which="oracle"):
    N, n = s.shape
    s, y = _f32(s), _f32(y)
    rho, q, alpha = _f32(rho).copy(), _f32(q).copy(), _f32(alpha).copy()
    r = np.zeros(n, np.float32)
    f = ref().ref_lbfgs_core if which == "ref" else oracle().orc_lbfgs_core
    f(iteration, N, n, _ptr(s), _ptr(y), _ptr(rho), _ptr(q), _ptr(alpha), _ptr(r))
    return r, rho, alpha, q


def gradient_we(p, model, srcpos, recpos, obs_vx, obs_vy, fp64=False):
    """Oracle gradient with the energy preconditioner set in p; returns
    (gradient dict, We [NY][NX]); fp64: the FP64 build of the scheme."""
    vp, vs, rho = (_f32(model[k]) for k in ("vp", "vs", "rho"))
    taus = _f32(model["taus"]) if p.L else None
    taup = _f32(model["taup"]) if p.L else None
    srcpos = _f32(srcpos).reshape(-1, 8)
    recpos = np.ascontiguousarray(recpos, dtype=np.int32).reshape(-1, 2)
    ntr = recpos.shape[0]
    dt = np.float64 if fp64 else np.float32
    ty = C.c_double if fp64 else C.c_float
    svx = np.zeros((ntr, p.NT), dt)
    svy = np.zeros((ntr, p.NT), dt)
    g = np.zeros((3, p.NY, p.NX), dt)
    we = np.zeros((p.NY, p.NX), dt)
    mis = np.zeros(5, np.float64)
    rc = oracle(fp64).orc_gradient_we(C.byref(p), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                                      C.c_int(srcpos.shape[0]), _ptr(srcpos), C.c_int(ntr), _ptr(recpos, C.c_int),
                                      _ptr(_f32(obs_vx)), _ptr(_f32(obs_vy)), None, _ptr(svx, ty), _ptr(svy, ty),
                                      _ptr(g[0], ty), _ptr(g[1], ty), _ptr(g[2], ty), _ptr(mis, C.c_double),
                                      _ptr(we, ty))
    assert rc == 0, rc
    return dict(seis_vx=svx, seis_vy=svy, grad_vp=g[0], grad_vs=g[1], grad_rho=g[2], misfit=mis), we


def ref_gradient_we(p, model, srcpos, recpos, diff_vx, diff_vy):
    vp, vs, rho = (_f32(model[k]) for k in ("vp", "vs", "rho"))
    taus = _f32(model["taus"]) if p.L else None
    taup = _f32(model["taup"]) if p.L else None
    srcpos = _f32(srcpos).reshape(-1, 8)
    recpos = np.ascontiguousarray(recpos, dtype=np.int32).reshape(-1, 2)
    ntr = recpos.shape[0]
    svx = np.zeros((ntr, p.NT), np.float32)
    svy = np.zeros((ntr, p.NT), np.float32)
    g = np.zeros((3, p.NY, p.NX), np.float32)
    we = np.zeros((p.NY, p.NX), np.float32)
    ip, fp, fl = _ip_fp(p)
    rc = ref().ref_gradient_we(ip, fp, fl, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                               C.c_int(srcpos.shape[0]), _ptr(srcpos), C.c_int(ntr), _ptr(recpos, C.c_int),
                               _ptr(_f32(diff_vx)), _ptr(_f32(diff_vy)), _ptr(svx), _ptr(svy), _ptr(g[0]), _ptr(g[1]),
                               _ptr(g[2]), C.c_int(p.EPRECOND), C.c_float(p.EPSILON_WE),
                               C.c_int(p.EPRECOND_PER_SHOT), _ptr(we))
    assert rc == 0, rc
    return dict(grad_vp=g[0], grad_vs=g[1], grad_rho=g[2]), we


# ---------------------------------------------------------------------------
# gradient post-processing (oracle/post_oracle.c; reference taper_grad_shot,
# taper_grad, smooth through oracle/_ref)
def taper_sources(g, DH, SRTRADIUS, SRTSHAPE, FILTSIZE, isrc, jsrc, which="oracle", ishot=1):
    """Circular source taper + box zeroing.  isrc/jsrc: grid indices of all
    shots.  which="ref" with ishot > 0 calls taper_grad_shot(..., ishot, 1);
    ishot = 0 calls taper_grad(..., 3) over all shots."""
    ny, nx = g.shape
    out = _f32(g).copy()
    isrc = np.ascontiguousarray(isrc, np.int32)
    jsrc = np.ascontiguousarray(jsrc, np.int32)
    if which == "ref":
        # positions in metres that iround(x/DH) maps back to the indices
        xs = _f32(isrc * np.float32(DH))
        ys = _f32(jsrc * np.float32(DH))
        if ishot > 0:
            ref().ref_taper_grad_shot(nx, ny, C.c_float(DH), C.c_float(SRTRADIUS), SRTSHAPE, FILTSIZE, len(isrc),
                                      _ptr(xs), _ptr(ys), ishot, _ptr(out))
        else:
            gt = np.zeros(4, np.int32)
            ref().ref_taper_grad(nx, ny, C.c_float(DH), _ptr(gt, C.c_int), C.c_float(SRTRADIUS), SRTSHAPE,
                                 FILTSIZE, len(isrc), _ptr(xs), _ptr(ys), 0, None, 3, _ptr(out))
        return out
    if ishot > 0:
        isrc, jsrc = isrc[ishot - 1:ishot].copy(), jsrc[ishot - 1:ishot].copy()
    oracle().orc_taper_sources(nx, ny, C.c_float(DH), C.c_float(SRTRADIUS), SRTSHAPE, FILTSIZE, len(isrc),
                               _ptr(isrc, C.c_int), _ptr(jsrc, C.c_int), _ptr(out), None)
    return out


def taper_vertical(g, gradt, isrc, jsrc, recpos, DH=1.0, which="oracle"):
    ny, nx = g.shape
    out = _f32(g).copy()
    isrc = np.ascontiguousarray(isrc, np.int32)
    jsrc = np.ascontiguousarray(jsrc, np.int32)
    rp = np.ascontiguousarray(recpos, np.int32).reshape(-1, 2)
    if which == "ref":
        gt = np.asarray(gradt, np.int32)
        xs = _f32(isrc * np.float32(DH))
        ys = _f32(jsrc * np.float32(DH))
        ref().ref_taper_grad(nx, ny, C.c_float(DH), _ptr(gt, C.c_int), C.c_float(0), 1, 0, len(isrc), _ptr(xs),
                             _ptr(ys), len(rp), _ptr(rp, C.c_int), 1, _ptr(out))
        return out
    oracle().orc_taper_vertical(nx, ny, *[int(v) for v in gradt], len(isrc), _ptr(isrc, C.c_int),
                                _ptr(jsrc, C.c_int), len(rp), _ptr(rp, C.c_int), _ptr(out), None)
    return out


def taper_horizontal(g, gradt, which="oracle"):
    ny, nx = g.shape
    out = _f32(g).copy()
    if which == "ref":
        gt = np.asarray(gradt, np.int32)
        one = np.ones(1, np.float32)
        ref().ref_taper_grad(nx, ny, C.c_float(1.0), _ptr(gt, C.c_int), C.c_float(0), 1, 0, 1, _ptr(one),
                             _ptr(one), 0, None, 2, _ptr(out))
        return out
    oracle().orc_taper_horizontal(nx, ny, int(gradt[0]), int(gradt[1]), _ptr(out), None)
    return out


def median_smooth(g, filtsize, which="oracle"):
    ny, nx = g.shape
    if which == "ref":
        out = _f32(g).copy()
        ref().ref_smooth(nx, ny, filtsize, _ptr(out))
        return out
    out = np.empty((ny, nx), np.float32)
    oracle().orc_median_smooth(nx, ny, filtsize, _ptr(_f32(g)), _ptr(out))
    return out


# ---------------------------------------------------------------------------
# time-domain Butterworth filter (oracle/filt_oracle.c; reference
# timedomain_filt + contrib/libcseife through oracle/_ref)
def time_filter(data, DT, fc, order, method=1, fhp=0.0, which="oracle"):
    out = _f32(data).copy()
    ntr, ns = out.shape
    f = ref().ref_time_filter if which == "ref" else oracle().orc_time_filter
    f(_ptr(out), ntr, ns, C.c_float(DT), C.c_float(fc), order, method, C.c_float(fhp))
    return out


def spat_filt(g, size, first_row, which="oracle"):
    ny, nx = g.shape
    if which == "ref":
        # the reference builds "<JACOBIAN>_g.old" in a 74-byte buffer
        # (STRING_SIZE, src/fd.h:22; src/spat_filt.c:45-62): keep the path short
        import shutil
        import tempfile
        d = tempfile.mkdtemp(prefix="sf", dir="/tmp")
        try:
            prefix = os.path.join(d, "jac")
            assert len(prefix) + 10 < 74, prefix
            out = _f32(g).copy()
            ref().ref_spat_filt(nx, ny, size, first_row, prefix.encode(), _ptr(out))
        finally:
            shutil.rmtree(d)
        return out
    out = np.empty((ny, nx), np.float32)
    oracle().orc_spat_filt(nx, ny, size, first_row, _ptr(_f32(g)), _ptr(out))
    return out


# --------------------------------------------------- single-operator probes --
def probe_shape(p):
    nd = p.FDORDER // 2 + 1
    return nd, (p.NY + 2 * nd, p.NX + 2 * nd)


def probe_derivs(p, fields5, which="oracle", fp64=False):
    """The eight unscaled staggered difference sums (sxx_x, sxy_x, sxy_y,
    syy_y of the velocity update; vxx, vyx, vxy, vyy of the stress update) of
    padded fields vx, vy, sxx, syy, sxy; [8][NY][NX]."""
    nd, shp = probe_shape(p)
    if which == "oracle":
        dt = np.float64 if fp64 else np.float32
        ty = C.c_double if fp64 else C.c_float
        f = np.ascontiguousarray(fields5, dtype=dt).reshape(5, *shp)
        out = np.zeros((8, p.NY, p.NX), dt)
        rc = oracle(fp64).orc_probe_derivs(C.byref(p), _ptr(f, ty), _ptr(out, ty))
        assert rc == 0
        return out
    f = _f32(fields5).reshape(5, *shp)
    out = np.zeros((8, p.NY, p.NX), np.float32)
    ip, fp, _ = _ip_fp(p)
    ref().ref_probe_derivs(ip, fp, _ptr(f), _ptr(out))
    return out


def probe_surface(p, model, fields8, prow, which="oracle"):
    nd, shp = probe_shape(p)
    vp, vs, rho = (_f32(model[k]) for k in ("vp", "vs", "rho"))
    taus = _f32(model["taus"]) if p.L else None
    taup = _f32(model["taup"]) if p.L else None
    f = _f32(fields8).reshape(8, *shp).copy()
    pr = _f32(prow).reshape(max(p.L, 1), p.NX).copy()
    if which == "oracle":
        rc = oracle().orc_probe_surface(C.byref(p), _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup),
                                        _ptr(f), _ptr(pr))
        assert rc == 0
    else:
        ip, fp, fl = _ip_fp(p)
        ref().ref_probe_surface(ip, fp, fl, _ptr(vp), _ptr(vs), _ptr(rho), _ptr(taus), _ptr(taup), _ptr(f), _ptr(pr))
    return f, pr


def ref_pcg(grads, C3, start, prefix):
    """The reference's PCG() over iterations 1..len(grads); grads
    [niter][3][NY][NX].  Returns (directions, scaled gradients), same shape."""
    g = _f32(grads)
    n, _, ny, nx = g.shape
    dirs = np.zeros_like(g)
    gps = np.zeros_like(g)
    c3 = _f32(np.asarray(C3))
    ref().ref_pcg(C.c_int(nx), C.c_int(ny), C.c_int(n), C.c_int(start), _ptr(c3), prefix.encode(), _ptr(g),
                  _ptr(dirs), _ptr(gps))
    return dirs, gps


# ------------------------------------------------------------- file readers --
def ref_sources(path, source_type=1, max_src=1000):
    out = np.zeros((max_src, 8), np.float32)
    n, fc = C.c_int(), C.c_float()
    rc = ref().ref_sources(path.encode(), source_type, max_src, _ptr(out), C.byref(n), C.byref(fc))
    assert rc == 0
    return out[:n.value].copy(), fc.value


def ref_splitsrc(srcpos_m, DH, NX, NY):
    sp = _f32(srcpos_m).reshape(-1, 8)
    out = np.zeros_like(sp)
    n = C.c_int()
    ref().ref_splitsrc(NX, NY, C.c_float(DH), len(sp), _ptr(sp), _ptr(out), C.byref(n))
    return out[:n.value].copy()


def ref_receivers(readrec, path, DH, refrec=(0.0, 0.0), line=(0.0, 0.0, 0.0, 0.0), ngeoph=1, max_tr=10000):
    out = np.zeros((max_tr, 2), np.int32)
    n = C.c_int()
    rc = ref().ref_receivers(readrec, path.encode() if path else None, C.c_float(DH), C.c_float(refrec[0]),
                             C.c_float(refrec[1]), *[C.c_float(v) for v in line], ngeoph, max_tr,
                             _ptr(out, C.c_int), C.byref(n))
    assert rc == 0
    return out[:n.value].copy()


def ref_inseis(prefix, shot, ntr, ns):
    out = np.zeros((ntr, ns), np.float32)
    ref().ref_inseis(prefix.encode(), shot, ntr, ns, _ptr(out))
    return out


def ref_stf_taper(trace, fc, dt):
    """taper() of src/taper.c (TAPER_STF) applied in place to a float32 trace."""
    L = ref()
    L.ref_stf_taper.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float]
    L.ref_stf_taper.restype = None
    L.ref_stf_taper(trace.ctypes.data_as(C.c_void_p), trace.shape[0], fc, dt)
