"""ctypes mirror of include/ifos2d_serve.h: the request server over the
native engine (admission by an HBM budget, resident engines, worker threads
whose requests overlap on the device).  Inputs are host arrays; the server
copies them at submission."""
import ctypes as C

import numpy as np

from . import Config, SRC_STRIDE, _check, lib

FORWARD, GRADIENT = 0, 1
_PREC = {"fp32": 0, "bf16": 1, "fp16": 2}


class _Request(C.Structure):
    _fields_ = [("cfg", Config), ("kind", C.c_int), ("snap_prec", C.c_int), ("nshots", C.c_int),
                ("nsrc", C.POINTER(C.c_int)), ("srcpos", C.POINTER(C.c_float)), ("ntr", C.POINTER(C.c_int)),
                ("recpos", C.POINTER(C.c_int)), ("vp", C.POINTER(C.c_float)), ("vs", C.POINTER(C.c_float)),
                ("rho", C.POINTER(C.c_float)), ("taus", C.POINTER(C.c_float)), ("taup", C.POINTER(C.c_float)),
                ("model_id", C.c_longlong), ("obs_vx", C.POINTER(C.c_float)), ("obs_vy", C.POINTER(C.c_float)),
                ("kill", C.POINTER(C.c_int))]


class _Result(C.Structure):
    _fields_ = [("seis_vx", C.POINTER(C.c_float)), ("seis_vy", C.POINTER(C.c_float)),
                ("g_vp", C.POINTER(C.c_float)), ("g_vs", C.POINTER(C.c_float)), ("g_rho", C.POINTER(C.c_float)),
                ("misfit", C.POINTER(C.c_double)), ("times", C.c_double * 3), ("slots", C.c_int),
                ("reused", C.c_int), ("shared", C.c_int)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_longlong) for n in ("submitted", "finished", "failed", "engines_created", "engines_reused",
                                            "engines_evicted")] + \
               [(n, C.c_size_t) for n in ("budget", "bytes_in_use", "peak_bytes")] + \
               [(n, C.c_longlong) for n in ("launches", "shared_launches")]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _f(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)


def _fp(a, ty=C.c_float):
    return None if a is None else a.ctypes.data_as(C.POINTER(ty))


class Server:
    def __init__(self, hbm_budget, workers=2, max_slots=4, device=0):
        L = lib()
        L.ifos_server_create.argtypes = [C.c_int, C.c_size_t, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.ifos_server_destroy.argtypes = [C.c_void_p]
        L.ifos_server_destroy.restype = None
        L.ifos_server_submit.argtypes = [C.c_void_p, C.POINTER(_Request), C.POINTER(C.c_longlong)]
        L.ifos_server_wait.argtypes = [C.c_void_p, C.c_longlong, C.c_double]
        L.ifos_server_result.argtypes = [C.c_void_p, C.c_longlong, C.POINTER(_Result)]
        L.ifos_server_stats_get.argtypes = [C.c_void_p, C.POINTER(Stats)]
        h = C.c_void_p()
        _check(L.ifos_server_create(device, int(hbm_budget), workers, max_slots, C.byref(h)))
        self.h = h
        self.meta = {}

    def close(self):
        if self.h:
            lib().ifos_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, kind, cfg, model, shots, obs=None, kill=None, snap="fp32", model_id=0):
        """shots: list of (srcpos [nsrc][8], recpos [ntr][2]); obs: (vx, vy)
        [sum ntr][NT] for gradient requests.  Returns the request id."""
        srcs = [np.asarray(sp, np.float32).reshape(-1, SRC_STRIDE) for sp, _ in shots]
        recs = [np.asarray(rp, np.int32).reshape(-1, 2) for _, rp in shots]
        keep = dict(nsrc=np.array([a.shape[0] for a in srcs], np.int32),
                    ntr=np.array([a.shape[0] for a in recs], np.int32),
                    srcpos=_f(np.concatenate(srcs)), recpos=np.ascontiguousarray(np.concatenate(recs), np.int32),
                    vp=_f(model["vp"]), vs=_f(model["vs"]), rho=_f(model["rho"]),
                    taus=_f(model.get("taus")), taup=_f(model.get("taup")))
        r = _Request()
        r.cfg = cfg
        r.kind = kind
        r.snap_prec = _PREC.get(snap, snap)
        r.nshots = len(shots)
        r.nsrc = _fp(keep["nsrc"], C.c_int)
        r.ntr = _fp(keep["ntr"], C.c_int)
        r.srcpos = _fp(keep["srcpos"])
        r.recpos = _fp(keep["recpos"], C.c_int)
        for k in ("vp", "vs", "rho", "taus", "taup"):
            setattr(r, k, _fp(keep[k]))
        r.model_id = model_id
        if kind == GRADIENT:
            keep["ovx"], keep["ovy"] = _f(obs[0]), _f(obs[1])
            r.obs_vx, r.obs_vy = _fp(keep["ovx"]), _fp(keep["ovy"])
            if kill is not None:
                keep["kill"] = np.ascontiguousarray(kill, np.int32)
                r.kill = _fp(keep["kill"], C.c_int)
        rid = C.c_longlong()
        _check(lib().ifos_server_submit(self.h, C.byref(r), C.byref(rid)))
        self.meta[rid.value] = (kind, cfg.NX, cfg.NY, cfg.NT, int(keep["ntr"].sum()), len(shots))
        return rid.value

    def wait(self, rid, timeout=-1.0):
        rc = lib().ifos_server_wait(self.h, rid, timeout)
        if rc == 1:
            return False
        _check(rc)
        return True

    def result(self, rid):
        kind, nx, ny, nt, ntr, nshots = self.meta.pop(rid)
        res = _Result()
        out = {}
        if kind == FORWARD:
            out["vx"] = np.zeros((ntr, nt), np.float32)
            out["vy"] = np.zeros((ntr, nt), np.float32)
            res.seis_vx, res.seis_vy = _fp(out["vx"]), _fp(out["vy"])
        else:
            out["grad"] = np.zeros((3, ny, nx), np.float32)
            out["misfit"] = np.zeros((nshots, 2), np.float64)
            res.g_vp, res.g_vs, res.g_rho = (_fp(out["grad"][k]) for k in range(3))
            res.misfit = _fp(out["misfit"], C.c_double)
        _check(lib().ifos_server_result(self.h, rid, C.byref(res)))
        out["times"] = tuple(res.times)
        out["slots"] = res.slots
        out["reused"] = bool(res.reused)
        out["shared"] = res.shared
        return out

    def stats(self):
        st = Stats()
        _check(lib().ifos_server_stats_get(self.h, C.byref(st)))
        return st.as_dict()
