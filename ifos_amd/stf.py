"""ctypes mirror of include/ifos2d_stf.h: source-time-function inversion
(INV_STF, src/stf.c:24) on the host, and the sequence stf() runs: the
correction filter from observed and synthetic vertical-component traces of
a shot (the fdlsq engine of contrib/stfinv), the source wavelet convolved
with it (conv_FD), the optional taper (TAPER_STF), ready for
Engine.set_wavelets."""
import ctypes as C

import numpy as np

from . import _check, lib


class StfParams(C.Structure):
    _fields_ = [("waterlevel", C.c_double), ("fpad", C.c_double), ("fpow2", C.c_int), ("fdiv", C.c_int),
                ("use_tshift", C.c_int), ("tshift", C.c_double), ("use_exp", C.c_int), ("exp", C.c_double),
                ("use_irtap", C.c_int), ("irtap", C.c_double * 4)]


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _p(a, ty=C.c_float):
    return None if a is None else a.ctypes.data_as(C.POINTER(ty))


def parse(para):
    """Parameters of an engine string such as "fdlsq:tshift=0.0"."""
    p = StfParams()
    L = lib()
    L.ifos_stf_parse.argtypes = [C.c_char_p, C.POINTER(StfParams)]
    _check(L.ifos_stf_parse(para.encode(), C.byref(p)))
    return p


def fdlsq(para, obs, syn, dt, offset=None):
    """Correction filter (impulse response, [ns]) and the synthetics
    convolved with it ([ntr][ns]) for observed and synthetic traces [ntr][ns]."""
    p = parse(para) if isinstance(para, str) else para
    obs, syn = _f32(obs), _f32(syn)
    ntr, ns = syn.shape
    off = None if offset is None else np.ascontiguousarray(offset, np.float64)
    stf = np.zeros(ns, np.float32)
    conv = np.zeros((ntr, ns), np.float32)
    L = lib()
    L.ifos_stf_fdlsq.argtypes = [C.POINTER(StfParams), C.c_int, C.c_int, C.c_float, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_void_p, C.c_void_p]
    _check(L.ifos_stf_fdlsq(C.byref(p), ntr, ns, dt, _p(obs), _p(syn), _p(off, C.c_double), _p(stf), _p(conv)))
    return stf, conv


def conv_fd(a, b):
    """First len(a) samples of the linear convolution of a and b (conv_FD)."""
    a, b = _f32(a), _f32(b)
    out = np.zeros_like(a)
    L = lib()
    L.ifos_conv_fd.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    _check(L.ifos_conv_fd(_p(a), _p(b), _p(out), a.shape[0]))
    return out


def taper(trace, fc, dt):
    """The TAPER_STF window of src/taper.c applied to a copy of one trace."""
    t = _f32(trace).copy()
    L = lib()
    L.ifos_stf_taper.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_float]
    L.ifos_stf_taper.restype = None
    L.ifos_stf_taper(_p(t), t.shape[0], fc, dt)
    return t


def dft(x, sign=-1):
    """Complex DFT of length len(x), unscaled, sign -1 or +1."""
    x = np.asarray(x, np.complex128)
    re = np.ascontiguousarray(x.real, np.float64)
    im = np.ascontiguousarray(x.imag, np.float64)
    L = lib()
    L.ifos_dft.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_int]
    _check(L.ifos_dft(_p(re, C.c_double), _p(im, C.c_double), len(x), sign))
    return re + 1j * im


def invert(para, obs_vy, syn_vy, wavelet, dt, offsets=None, fc=None, taper_stf=False):
    """stf() of src/stf.c:24 for one shot: the correction filter from the
    vertical-component traces, the wavelet convolved with it, and with
    taper_stf the TAPER_STF window at fc; returns (new wavelet, filter,
    convolved synthetics)."""
    stf, conv = fdlsq(para, obs_vy, syn_vy, dt, offsets)
    w = conv_fd(wavelet, stf)
    if taper_stf:
        w = taper(w, fc, dt)
    return w, stf, conv
