"""ctypes mirror of include/ifos2d_formats.h: the reference's parameter
(quasi-JSON), model (binary) and seismogram (SU) files."""
import ctypes as C

import numpy as np

from . import Config, _check, lib


def _L():
    L = lib()
    L.ifos_json_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    L.ifos_json_free.argtypes = [C.c_void_p]
    L.ifos_json_free.restype = None
    L.ifos_json_count.argtypes = [C.c_void_p]
    L.ifos_json_get.argtypes = [C.c_void_p, C.c_char_p]
    L.ifos_json_get.restype = C.c_char_p
    L.ifos_config_from_json.argtypes = [C.c_char_p, C.POINTER(Config)]
    L.ifos_model_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.ifos_model_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float)]
    L.ifos_su_write.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int), C.c_float,
                                C.c_float, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
    L.ifos_su_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int)]
    return L


def read_json(path):
    """All objects of a parameter file, in file order (name, value)."""
    L = _L()
    h = C.c_void_p()
    _check(L.ifos_json_load(path.encode(), C.byref(h)))
    try:
        return {k.decode(): v.decode() for k, v in _pairs(L, h)}
    finally:
        L.ifos_json_free(h)


def _pairs(L, h):
    # the C side exposes lookup by name; walk the file order through the names it holds
    n = L.ifos_json_count(h)
    L.ifos_json_name.argtypes = [C.c_void_p, C.c_int]
    L.ifos_json_name.restype = C.c_char_p
    L.ifos_json_value.argtypes = [C.c_void_p, C.c_int]
    L.ifos_json_value.restype = C.c_char_p
    return [(L.ifos_json_name(h, k), L.ifos_json_value(h, k)) for k in range(n)]


def config_from_json(path, base=None):
    c = Config() if base is None else base
    _check(_L().ifos_config_from_json(path.encode(), C.byref(c)))
    return c


def read_model(path, NX, NY):
    m = np.zeros((NY, NX), np.float32)
    _check(_L().ifos_model_read(path.encode(), NX, NY, m.ctypes.data_as(C.POINTER(C.c_float))))
    return m


def write_model(path, m):
    m = np.ascontiguousarray(m, np.float32)
    _check(_L().ifos_model_write(path.encode(), m.shape[1], m.shape[0], m.ctypes.data_as(C.POINTER(C.c_float))))


def write_su(path, data, recpos, DH, DT, xs, ys, ndt=1, refrec=(0.0, 0.0)):
    data = np.ascontiguousarray(data, np.float32)
    rp = np.ascontiguousarray(recpos, np.int32)
    _check(_L().ifos_su_write(path.encode(), data.shape[0], data.shape[1], data.ctypes.data_as(C.POINTER(C.c_float)),
                              rp.ctypes.data_as(C.POINTER(C.c_int)), DH, DT, ndt, xs, ys, refrec[0], refrec[1]))


def read_su(path, ns, max_tr=100000):
    buf = np.zeros((max_tr, ns), np.float32)
    n = C.c_int()
    _check(_L().ifos_su_read(path.encode(), max_tr, ns, buf.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)))
    return buf[:n.value].copy()


def read_trkill(path, ntr_glob, nsrc_glob):
    """Trace kill table [ntr_glob][nsrc_glob] from a TRKILL_FILE.dat."""
    L = _L()
    L.ifos_trkill_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int)]
    t = np.zeros((ntr_glob, nsrc_glob), np.int32)
    _check(L.ifos_trkill_read(path.encode(), ntr_glob, nsrc_glob, t.ctypes.data_as(C.POINTER(C.c_int))))
    return t


def trkill_offsets(recpos, srcpos, DH, lower, upper, table=None):
    """Offset-based trace kill (create_trkill_table): recpos [ntr][2] grid
    indices, srcpos [nsrc][2] metres; marks 1 where lower < offset < upper,
    added to `table` when given."""
    L = _L()
    L.ifos_trkill_offsets.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_float), C.c_float,
                                      C.c_float, C.c_float, C.POINTER(C.c_int)]
    rp = np.ascontiguousarray(recpos, np.int32).reshape(-1, 2)
    sp = np.ascontiguousarray(srcpos, np.float32).reshape(-1, 2)
    t = np.zeros((len(rp), len(sp)), np.int32) if table is None else np.ascontiguousarray(table, np.int32).copy()
    _check(L.ifos_trkill_offsets(len(rp), rp.ctypes.data_as(C.POINTER(C.c_int)), len(sp),
                                 sp.ctypes.data_as(C.POINTER(C.c_float)), DH, lower, upper,
                                 t.ctypes.data_as(C.POINTER(C.c_int))))
    return t


def read_su_observed(path, ntr, ns):
    """Observed traces as inseis reads them: sample 1 of each trace zeroed."""
    L = _L()
    L.ifos_su_read_observed.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float)]
    buf = np.zeros((ntr, ns), np.float32)
    _check(L.ifos_su_read_observed(path.encode(), ntr, ns, buf.ctypes.data_as(C.POINTER(C.c_float))))
    return buf


def read_sources(path, source_type=1, max_src=100000):
    """SOURCE_FILE rows [nsrc][8] (x, y in metres, 0, tshift, fc, amp,
    azimuth, type) and the largest fc."""
    L = _L()
    L.ifos_sources_read.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(C.c_int),
                                    C.POINTER(C.c_float)]
    buf = np.zeros((max_src, 8), np.float32)
    n, fc = C.c_int(), C.c_float()
    _check(L.ifos_sources_read(path.encode(), source_type, max_src, buf.ctypes.data_as(C.POINTER(C.c_float)),
                               C.byref(n), C.byref(fc)))
    return buf[:n.value].copy(), fc.value


def sources_to_grid(srcpos_m, DH, NX, NY):
    """Grid-index source rows (splitsrc on one rank)."""
    L = _L()
    L.ifos_sources_to_grid.argtypes = [C.c_int, C.POINTER(C.c_float), C.c_float, C.c_int, C.c_int,
                                       C.POINTER(C.c_float), C.POINTER(C.c_int)]
    sp = np.ascontiguousarray(srcpos_m, np.float32).reshape(-1, 8)
    out = np.zeros_like(sp)
    n = C.c_int()
    _check(L.ifos_sources_to_grid(len(sp), sp.ctypes.data_as(C.POINTER(C.c_float)), DH, NX, NY,
                                  out.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)))
    return out[:n.value].copy()


def read_receivers(path, DH, refrec=(0.0, 0.0), max_tr=100000):
    """REC_FILE receiver grid indices [ntr][2] (READREC 1)."""
    L = _L()
    L.ifos_receivers_read.argtypes = [C.c_char_p, C.c_float, C.c_float, C.c_float, C.c_int, C.POINTER(C.c_int),
                                      C.POINTER(C.c_int)]
    buf = np.zeros((max_tr, 2), np.int32)
    n = C.c_int()
    _check(L.ifos_receivers_read(path.encode(), DH, refrec[0], refrec[1], max_tr,
                                 buf.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n)))
    return buf[:n.value].copy()


def receiver_line(DH, XREC1, YREC1, XREC2, YREC2, NGEOPH, max_tr=100000):
    """A straight receiver line (READREC 0), grid indices [ntr][2]."""
    L = _L()
    L.ifos_receivers_line.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int,
                                      C.POINTER(C.c_int), C.POINTER(C.c_int)]
    buf = np.zeros((max_tr, 2), np.int32)
    n = C.c_int()
    _check(L.ifos_receivers_line(DH, XREC1, YREC1, XREC2, YREC2, NGEOPH, max_tr,
                                 buf.ctypes.data_as(C.POINTER(C.c_int)), C.byref(n)))
    return buf[:n.value].copy()


def dump_matrix(path_name, m):
    """write_matrix_disk: m [NY][NX] to <path_name>.bin, x-outer."""
    L = _L()
    L.ifos_matrix_dump.argtypes = [C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_float)]
    m = np.ascontiguousarray(m, np.float32)
    _check(L.ifos_matrix_dump(path_name.encode(), m.shape[1], m.shape[0], m.ctypes.data_as(C.POINTER(C.c_float))))
