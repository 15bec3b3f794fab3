"""ctypes mirror of include/ifos2d_opt.h: the model-update side of an FWI
iteration on the device (reductions, PCG direction, step-length parabola,
bounded model update, L-BFGS) and the gradient post-processing (tapers,
smoothing).  Arrays are DeviceBuffer objects holding
interior [NY][NX] float32 grids; the compute runs in the gfx950 library."""
import ctypes as C

import numpy as np

from . import DeviceBuffer, IfosError, _check, lib

_VOIDP = C.c_void_p


class UpdateParams(C.Structure):
    """calc_mat_change_test parameters (src/calc_mat_change_test.c:26)."""
    _fields_ = ([(n, C.c_int) for n in ("iter", "INV_VP_ITER", "INV_VS_ITER", "INV_RHO_ITER", "S", "GRAD_METHOD",
                                        "LBFGS_STEP_LENGTH", "LBFGS_iter_start")]
                + [(n, C.c_float) for n in ("eps_scale", "VPUPPERLIM", "VPLOWERLIM", "VSUPPERLIM", "VSLOWERLIM",
                                            "RHOUPPERLIM", "RHOLOWERLIM", "VP_VS_RATIO", "S_VP", "S_VS", "S_RHO")])


def update_params(**kw):
    d = dict(iter=1, INV_VP_ITER=0, INV_VS_ITER=0, INV_RHO_ITER=0, S=0, GRAD_METHOD=1, LBFGS_STEP_LENGTH=1,
             LBFGS_iter_start=1, eps_scale=0.01, VPUPPERLIM=1e9, VPLOWERLIM=0.0, VSUPPERLIM=1e9, VSLOWERLIM=0.0,
             RHOUPPERLIM=1e9, RHOLOWERLIM=0.0, VP_VS_RATIO=0.0, S_VP=0.0, S_VS=0.0, S_RHO=0.0)
    unknown = set(kw) - set(d)
    if unknown:
        raise IfosError(f"unknown update parameters {sorted(unknown)}")
    d.update(kw)
    u = UpdateParams()
    for k, v in d.items():
        setattr(u, k, v)
    return u


def _p(b):
    return _VOIDP(b.ptr if isinstance(b, DeviceBuffer) else b)


TAPER_VERTICAL, TAPER_HORIZONTAL, TAPER_SOURCES = 1, 2, 3


class Taper(C.Structure):
    """ifos_taper: a whole-gradient taper of taper_grad (src/taper_grad.c:28)."""
    _fields_ = [("kind", C.c_int), ("GRADT1", C.c_int), ("GRADT2", C.c_int), ("GRADT3", C.c_int),
                ("GRADT4", C.c_int), ("DH", C.c_float), ("SRTRADIUS", C.c_float), ("SRTSHAPE", C.c_int),
                ("FILTSIZE", C.c_int), ("nsrc", C.c_int), ("srcpos", C.POINTER(C.c_int)), ("ntr", C.c_int),
                ("recpos", C.POINTER(C.c_int))]


def taper_coeff(kind, NX, NY, gradt=(0, 0, 0, 0), DH=1.0, SRTRADIUS=0.0, SRTSHAPE=1, FILTSIZE=0, srcpos=None,
                recpos=None):
    """Host: coefficients [NY][NX] and zero mask of a taper (ifos_taper_coeff).
    srcpos / recpos: [n][2] grid indices."""
    L = lib()
    L.ifos_taper_coeff.argtypes = [C.POINTER(Taper), C.c_int, C.c_int, C.POINTER(C.c_float),
                                   C.POINTER(C.c_ubyte)]
    sp = np.ascontiguousarray(srcpos if srcpos is not None else np.zeros((0, 2)), np.int32).reshape(-1, 2)
    rp = np.ascontiguousarray(recpos if recpos is not None else np.zeros((0, 2)), np.int32).reshape(-1, 2)
    t = Taper(kind, *[int(v) for v in gradt], DH, SRTRADIUS, SRTSHAPE, FILTSIZE, len(sp),
              sp.ctypes.data_as(C.POINTER(C.c_int)), len(rp), rp.ctypes.data_as(C.POINTER(C.c_int)))
    coef = np.empty((NY, NX), np.float32)
    zero = np.empty((NY, NX), np.uint8)
    _check(L.ifos_taper_coeff(C.byref(t), NX, NY, coef.ctypes.data_as(C.POINTER(C.c_float)),
                              zero.ctypes.data_as(C.POINTER(C.c_ubyte))))
    return coef, zero


class Optimizer:
    def __init__(self, device=0):
        L = lib()
        L.ifos_opt_create.argtypes = [C.c_int, C.POINTER(_VOIDP)]
        L.ifos_opt_destroy.argtypes = [_VOIDP]
        L.ifos_opt_destroy.restype = None
        L.ifos_opt_dot.argtypes = [_VOIDP, _VOIDP, _VOIDP, C.c_long, C.POINTER(C.c_double)]
        L.ifos_opt_max.argtypes = [_VOIDP, _VOIDP, C.c_long, C.c_int, C.POINTER(C.c_float)]
        L.ifos_opt_pcg.argtypes = [_VOIDP, C.c_long, _VOIDP, C.c_float, _VOIDP, _VOIDP, C.c_int, C.c_int, _VOIDP,
                                   C.POINTER(C.c_double)]
        L.ifos_opt_step_length.argtypes = [_VOIDP, C.POINTER(C.c_float), C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.ifos_opt_model_update.argtypes = [_VOIDP, C.c_long, C.POINTER(UpdateParams)] + [_VOIDP] * 12 + \
            [C.POINTER(C.c_float)]
        L.ifos_opt_lbfgs_store_s.argtypes = [_VOIDP, C.c_long, C.c_int, C.POINTER(_VOIDP), C.POINTER(_VOIDP),
                                             C.POINTER(C.c_float), _VOIDP, C.c_int, C.c_int]
        L.ifos_opt_lbfgs.argtypes = [_VOIDP, C.c_long, C.c_int, C.POINTER(_VOIDP), C.POINTER(C.c_float), _VOIDP,
                                     _VOIDP, C.POINTER(C.c_double), _VOIDP, C.c_int, C.c_int, C.c_int]
        L.ifos_opt_taper.argtypes = [_VOIDP, C.c_long, C.POINTER(C.c_float), C.POINTER(C.c_ubyte), _VOIDP]
        L.ifos_opt_smooth.argtypes = [_VOIDP, C.c_int, C.c_int, C.c_int, _VOIDP, C.POINTER(C.c_int)]
        L.ifos_opt_spat_filt.argtypes = [_VOIDP, C.c_int, C.c_int, C.c_int, C.c_int, _VOIDP]
        h = _VOIDP()
        _check(L.ifos_opt_create(device, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().ifos_opt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dot(self, a, b, n):
        out = C.c_double()
        _check(lib().ifos_opt_dot(self.h, _p(a), _p(b), n, C.byref(out)))
        return out.value

    def max(self, a, n, absolute=False):
        out = C.c_float()
        _check(lib().ifos_opt_max(self.h, _p(a), n, int(absolute), C.byref(out)))
        return out.value

    def pcg(self, n, grad, C_, p_last, c_last, it, pcg_iter_start, dir_out):
        beta = C.c_double()
        _check(lib().ifos_opt_pcg(self.h, n, _p(grad), C_, _p(p_last), _p(c_last), it, pcg_iter_start,
                                  _p(dir_out), C.byref(beta)))
        return beta.value

    def step_length(self, L2t, epst):
        a = (C.c_float * 3)(*[float(v) for v in L2t])
        b = (C.c_float * 3)(*[float(v) for v in epst])
        out = C.c_float()
        _check(lib().ifos_opt_step_length(self.h, a, b, C.byref(out)))
        return out.value

    def model_update(self, n, u, vp, vs, rho, g_vp, g_vs, g_rho, vp0, vs0, rho0, vp1, vs1, rho1):
        eps = (C.c_float * 3)()
        _check(lib().ifos_opt_model_update(self.h, n, C.byref(u), *[_p(b) for b in (vp, vs, rho, g_vp, g_vs, g_rho,
                                                                                    vp0, vs0, rho0, vp1, vs1, rho1)],
                                           eps))
        return np.array(eps[:], np.float32)

    def lbfgs_store_s(self, ncell, m1, m0, avg, s_hist, N, it):
        npar = len(m1)
        a = (_VOIDP * npar)(*[_p(b) for b in m1])
        b = (_VOIDP * npar)(*[_p(x) for x in m0])
        av = (C.c_float * npar)(*[float(v) for v in avg])
        _check(lib().ifos_opt_lbfgs_store_s(self.h, ncell, npar, a, b, av, _p(s_hist), N, it))

    def lbfgs(self, ncell, grads, avg, s_hist, y_hist, rho_hist, work, N, it, start):
        npar = len(grads)
        g = (_VOIDP * npar)(*[_p(b) for b in grads])
        av = (C.c_float * npar)(*[float(v) for v in avg])
        r = rho_hist.ctypes.data_as(C.POINTER(C.c_double))
        _check(lib().ifos_opt_lbfgs(self.h, ncell, npar, g, av, _p(s_hist), _p(y_hist), r, _p(work), N, it, start))

    def taper(self, grad, coef, zero=None):
        """grad (device) = zero ? 0 : grad * coef (coef, zero host arrays)."""
        coef = np.ascontiguousarray(coef, np.float32)
        z = None if zero is None else np.ascontiguousarray(zero, np.uint8)
        _check(lib().ifos_opt_taper(self.h, coef.size, coef.ctypes.data_as(C.POINTER(C.c_float)),
                                    None if z is None else z.ctypes.data_as(C.POINTER(C.c_ubyte)), _p(grad)))

    def smooth(self, NX, NY, filtsize, grad):
        """smooth() of src/smooth.c on a device gradient, in place; returns
        the window side used."""
        side = C.c_int()
        _check(lib().ifos_opt_smooth(self.h, NX, NY, filtsize, _p(grad), C.byref(side)))
        return side.value

    def spat_filt(self, NX, NY, size, first_row, grad):
        """spat_filt() of src/spat_filt.c on a device gradient, in place."""
        _check(lib().ifos_opt_spat_filt(self.h, NX, NY, size, first_row, _p(grad)))
