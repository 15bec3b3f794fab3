"""ifos_amd -- MI355X engine for IFOS2D's P-SV forward solver and adjoint
gradient.

The compute path is the gfx950 library lib/libifos2d_hip.so (C ABI declared
in include/ifos2d_hip.h).  This module is a thin ctypes binding to it; there
is no CPU fallback: loading fails loudly when the library has not been built.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# IFOS_LIB names another build of the library (for comparing builds)
LIB_PATH = os.environ.get("IFOS_LIB") or os.path.join(HERE, "lib", "libifos2d_hip.so")

_INT_FIELDS = ["NX", "NY", "FDORDER", "MAXRELERROR", "FW", "FREE_SURF", "BOUNDARY", "L",
               "VELOCITY", "ADJOINT_TYPE", "PARAMETERIZATION", "DTINV", "NT",
               "SOURCE_SHAPE", "SOURCE_TYPE", "TRKILL"]
_FLT_FIELDS = ["DH", "DT", "VPPML", "FPML", "npower", "k_max_PML", "F_REF", "TAU"]

SRC_STRIDE = 8


class Config(C.Structure):
    """Mirror of ifos_config (include/ifos2d_hip.h); names of src/globvar.h."""
    _fields_ = ([(n, C.c_int) for n in _INT_FIELDS] + [(n, C.c_float) for n in _FLT_FIELDS]
                + [("FL", C.c_float * 8)])

    def as_dict(self):
        d = {n: getattr(self, n) for n in _INT_FIELDS + _FLT_FIELDS}
        d["FL"] = [self.FL[i] for i in range(1, self.L + 1)]
        return d


DEFAULTS = dict(MAXRELERROR=1, FW=10, FREE_SURF=1, BOUNDARY=0, L=0, VELOCITY=0, ADJOINT_TYPE=1,
                PARAMETERIZATION=1, DTINV=1, SOURCE_SHAPE=1, SOURCE_TYPE=1, TRKILL=0,
                VPPML=3500.0, FPML=30.0, npower=2.0, k_max_PML=1.0, F_REF=0.0, TAU=0.0)


def make_config(**kw):
    c = Config()
    d = dict(DEFAULTS)
    d.update(kw)
    fl = d.pop("FL", None)
    for k, v in d.items():
        setattr(c, k, v)
    if fl is not None:
        for i, v in enumerate(fl):
            c.FL[i + 1] = v
    return c


_LIB = None


class IfosError(RuntimeError):
    pass


def lib():
    """Load the gfx950 library (raises if it is missing)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise IfosError(f"{LIB_PATH} is missing: build it with `python -m ifos_amd.build`")
        L = C.CDLL(LIB_PATH)
        L.ifos_last_error.restype = C.c_char_p
        L.ifos_dev_alloc.restype = C.c_void_p
        L.ifos_dev_alloc.argtypes = [C.c_size_t]
        L.ifos_dev_free.argtypes = [C.c_void_p]
        L.ifos_memcpy_h2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.ifos_memcpy_d2h.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.ifos_memcpy_d2d.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t]
        L.ifos_engine_create.argtypes = [C.POINTER(Config), C.c_int, C.POINTER(C.c_void_p)]
        L.ifos_engine_create_batch.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.ifos_engine_destroy.argtypes = [C.c_void_p]
        L.ifos_engine_destroy.restype = None
        L.ifos_engine_kernel_bytes_per_cell.restype = C.c_double
        L.ifos_engine_kernel_bytes_per_cell.argtypes = [C.c_void_p, C.c_int]
        for n in ["set_model", "set_shot", "forward", "seismograms", "fields", "residual", "adjoint",
                  "shot_gradient", "gradient", "reset_gradient", "time_kernel", "nan_step",
                  "set_snapshot_precision", "set_shots", "shot_misfit", "set_launch", "get_launch", "autotune",
                  "set_eprecond", "apply_eprecond", "energy_weight"]:
            getattr(L, "ifos_engine_" + n).restype = C.c_int
        L.ifos_snapshot_bytes.restype = C.c_size_t
        L.ifos_engine_time_copy.restype = C.c_int
        L.ifos_engine_time_copy.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_double)]
        L.ifos_engine_set_eprecond.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_int]
        L.ifos_engine_bytes.restype = C.c_size_t
        L.ifos_engine_bytes.argtypes = [C.POINTER(Config), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.ifos_snapshot_bytes.argtypes = [C.POINTER(Config), C.c_int]
        _LIB = L
    return _LIB


def _check(rc):
    if rc != 0:
        raise IfosError(lib().ifos_last_error().decode() + f" (code {rc})")


def _fp(a, ty=C.c_float):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ty))


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def time_filter(buf, ntr, ns, DT, fc, order, method=1, f_high_pass=0.0):
    """timedomain_filt on device traces buf [ntr][ns] (a DeviceBuffer), in place."""
    L = lib()
    L.ifos_time_filter.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, C.c_float]
    _check(L.ifos_time_filter(C.c_void_p(buf.ptr), ntr, ns, DT, fc, order, method, f_high_pass))


class DeviceBuffer:
    """Device allocation owned by the library's allocator (plain bytes)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = lib().ifos_dev_alloc(self.nbytes)
        if not self.ptr:
            raise IfosError(lib().ifos_last_error().decode())

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().ifos_memcpy_h2d(C.c_void_p(b.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))
        return b

    def to_array(self, shape, dtype=np.float32):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        _check(lib().ifos_memcpy_d2h(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes))
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        _check(lib().ifos_memcpy_h2d(C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes))

    def free(self):
        if self.ptr:
            lib().ifos_dev_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Engine:
    """One-GPU P-SV engine; mirrors the reference's per-shot flow
    (src/IFOS2D.c:1076-2724): set_model -> set_shot -> forward / residual /
    adjoint -> gradient."""

    def __init__(self, config, device=0, slots=1):
        self.cfg = config
        self.slots = slots
        h = C.c_void_p()
        if slots == 1:
            _check(lib().ifos_engine_create(C.byref(config), device, C.byref(h)))
        else:
            _check(lib().ifos_engine_create_batch(C.byref(config), device, slots, C.byref(h)))
        self.h = h
        self.ntr = 0
        self.nsrc = 0
        self.nshots = 0

    def close(self):
        if self.h:
            lib().ifos_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_model(self, vp, vs, rho, taus=None, taup=None):
        vp, vs, rho, taus, taup = (_f32(a) for a in (vp, vs, rho, taus, taup))
        _check(lib().ifos_engine_set_model(self.h, _fp(vp), _fp(vs), _fp(rho), _fp(taus), _fp(taup)))

    def set_shot(self, srcpos, recpos):
        srcpos = _f32(np.asarray(srcpos, np.float32).reshape(-1, SRC_STRIDE))
        recpos = np.ascontiguousarray(np.asarray(recpos, np.int32).reshape(-1, 2))
        self.nsrc, self.ntr = srcpos.shape[0], recpos.shape[0]
        self.nshots = 1
        self.trace_counts = [self.ntr]
        _check(lib().ifos_engine_set_shot(self.h, self.nsrc, _fp(srcpos), self.ntr, _fp(recpos, C.c_int)))

    def set_shots(self, shots):
        """Several shots at once, one per slot: a list of (srcpos, recpos).
        Seismograms, observed traces and kill masks are then the slots'
        traces concatenated in slot order."""
        srcs = [np.asarray(sp, np.float32).reshape(-1, SRC_STRIDE) for sp, _ in shots]
        recs = [np.asarray(rp, np.int32).reshape(-1, 2) for _, rp in shots]
        ns = np.array([a.shape[0] for a in srcs], np.int32)
        nr = np.array([a.shape[0] for a in recs], np.int32)
        srcpos = _f32(np.concatenate(srcs) if len(srcs) else np.zeros((0, SRC_STRIDE), np.float32))
        recpos = np.ascontiguousarray(np.concatenate(recs), np.int32)
        self.nsrc, self.ntr, self.nshots = int(ns.sum()), int(nr.sum()), len(shots)
        self.trace_counts = [int(v) for v in nr]
        _check(lib().ifos_engine_set_shots(self.h, len(shots), _fp(ns, C.c_int), _fp(srcpos), _fp(nr, C.c_int),
                                           _fp(recpos, C.c_int)))

    def set_wavelets(self, sig):
        """Source signals [nsrc][NT] of the current shots' sources instead of the generated wavelets."""
        sig = _f32(np.asarray(sig, np.float32).reshape(self.nsrc, self.cfg.NT))
        _check(lib().ifos_engine_set_wavelets(self.h, _fp(sig)))

    def shot_misfit(self):
        """(L2, energy) per shot of the last residual, FP64."""
        out = np.zeros((self.nshots, 2), np.float64)
        _check(lib().ifos_engine_shot_misfit(self.h, _fp(out, C.c_double)))
        return out

    def set_graphs(self, mode):
        """Time loop: 1 captured graphs, 0 eager launches, -1 by grid size."""
        _check(lib().ifos_engine_set_graphs(self.h, int(mode)))

    def get_graphs(self):
        on = C.c_int()
        _check(lib().ifos_engine_get_graphs(self.h, C.byref(on)))
        return bool(on.value)

    def set_launch(self, spb=1, grid=0):
        _check(lib().ifos_engine_set_launch(self.h, spb, grid))

    def get_launch(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().ifos_engine_get_launch(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def get_launch_grids(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().ifos_engine_get_launch_grids(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_velocity_rows(self, rows):
        _check(lib().ifos_engine_set_velocity_rows(self.h, int(rows)))

    def get_velocity_rows(self):
        v = C.c_int()
        _check(lib().ifos_engine_get_velocity_rows(self.h, C.byref(v)))
        return v.value

    def set_fused(self, on):
        """One kernel per time step (velocity and stress) instead of two; same bits."""
        _check(lib().ifos_engine_set_fused(self.h, int(bool(on))))

    def get_fused(self):
        v = C.c_int()
        _check(lib().ifos_engine_get_fused(self.h, C.byref(v)))
        return v.value

    def set_velocity_lds(self, on):
        _check(lib().ifos_engine_set_velocity_lds(self.h, int(bool(on))))

    def get_velocity_lds(self):
        v = C.c_int()
        _check(lib().ifos_engine_get_velocity_lds(self.h, C.byref(v)))
        return v.value

    def set_work_order(self, slot_major):
        _check(lib().ifos_engine_set_work_order(self.h, int(bool(slot_major))))

    def get_work_order(self):
        v = C.c_int()
        _check(lib().ifos_engine_get_work_order(self.h, C.byref(v)))
        return v.value

    def autotune(self, reps=5):
        """Pick shots per workgroup and the persistent grid by timing; returns (spb, grid, ms per step)."""
        a, b, ms = C.c_int(), C.c_int(), C.c_double()
        _check(lib().ifos_engine_autotune(self.h, reps, C.byref(a), C.byref(b), C.byref(ms)))
        return a.value, b.value, ms.value

    def forward(self, snapshots=False):
        _check(lib().ifos_engine_forward(self.h, int(bool(snapshots))))

    def seismograms(self):
        vx = np.zeros((self.ntr, self.cfg.NT), np.float32)
        vy = np.zeros_like(vx)
        _check(lib().ifos_engine_seismograms(self.h, _fp(vx), _fp(vy)))
        return vx, vy

    def fields(self):
        out = np.zeros((5, self.cfg.NY, self.cfg.NX), np.float32)
        _check(lib().ifos_engine_fields(self.h, _fp(out)))
        return out

    def _obs_args(self, obs_vx, obs_vy):
        if isinstance(obs_vx, DeviceBuffer):
            return C.c_void_p(obs_vx.ptr), C.c_void_p(obs_vy.ptr), 1, None
        ovx, ovy = _f32(obs_vx), _f32(obs_vy)
        return ovx.ctypes.data_as(C.c_void_p), ovy.ctypes.data_as(C.c_void_p), 0, (ovx, ovy)

    def residual(self, obs_vx, obs_vy, kill=None):
        a, b, dev, keep = self._obs_args(obs_vx, obs_vy)
        k = None if kill is None else np.ascontiguousarray(kill, np.int32)
        mis = np.zeros(2, np.float64)
        _check(lib().ifos_engine_residual(self.h, a, b, _fp(k, C.c_int), dev, _fp(mis, C.c_double)))
        return mis

    def adjoint(self):
        _check(lib().ifos_engine_adjoint(self.h))

    def shot_gradient(self, obs_vx, obs_vy, kill=None):
        a, b, dev, keep = self._obs_args(obs_vx, obs_vy)
        k = None if kill is None else np.ascontiguousarray(kill, np.int32)
        mis = np.zeros(2, np.float64)
        _check(lib().ifos_engine_shot_gradient(self.h, a, b, _fp(k, C.c_int), dev, _fp(mis, C.c_double)))
        return mis

    def gradient(self, total=False):
        g = np.zeros((3, self.cfg.NY, self.cfg.NX), np.float32)
        _check(lib().ifos_engine_gradient(self.h, int(total), _fp(g[0]), _fp(g[1]), _fp(g[2])))
        return g

    def reset_gradient(self):
        _check(lib().ifos_engine_reset_gradient(self.h))

    def slot_gradient(self, slot):
        """The gradient of one slot's shot from the last adjoint run (before the shot sum)."""
        g = np.zeros((3, self.cfg.NY, self.cfg.NX), np.float32)
        _check(lib().ifos_engine_slot_gradient(self.h, int(slot), _fp(g[0]), _fp(g[1]), _fp(g[2])))
        return g

    def time_kernel(self, kernel, reps=20):
        ms = C.c_double()
        _check(lib().ifos_engine_time_kernel(self.h, kernel, reps, C.byref(ms)))
        return ms.value

    def time_copy(self, nbytes, reps=20):
        """Average ms of a device copy reading and writing nbytes/2 each,
        back to back on the engine's stream."""
        ms = C.c_double()
        _check(lib().ifos_engine_time_copy(self.h, int(nbytes), reps, C.byref(ms)))
        return ms.value

    def kernel_bytes_per_cell(self, kernel):
        return lib().ifos_engine_kernel_bytes_per_cell(self.h, kernel)

    def set_eprecond(self, mode, epsilon_we=0.005, per_shot=False):
        """Energy preconditioning of the gradient (EPRECOND 1 or 3; 0 off)."""
        _check(lib().ifos_engine_set_eprecond(self.h, mode, epsilon_we, int(per_shot)))

    def set_eprecond_fp64(self, on):
        """Energy sums of the preconditioner in FP64 (True) or FP32 as the reference (False)."""
        _check(lib().ifos_engine_set_eprecond_fp64(self.h, int(bool(on))))

    def apply_eprecond(self):
        _check(lib().ifos_engine_apply_eprecond(self.h))

    def set_shot_taper_file(self, slot, vp=None, vs=None, rho=None):
        """Per-shot taper grids (SWS_TAPER_FILE_PER_SHOT) of one slot; no grids turns it off."""
        L = lib()
        L.ifos_engine_set_shot_taper_file.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        arrs = [None if a is None else _f32(a) for a in (vp, vs, rho)]
        ptrs = [None if a is None else a.ctypes.data_as(C.c_void_p) for a in arrs]
        _check(L.ifos_engine_set_shot_taper_file(self.h, slot, *ptrs))

    def set_time_filter(self, f_low_pass, order=4, f_high_pass=0.0):
        """Butterworth filter of wavelets and observed traces (TIME_FILT); 0 turns it off."""
        L = lib()
        L.ifos_engine_set_time_filter.argtypes = [C.c_void_p, C.c_float, C.c_int, C.c_float]
        _check(L.ifos_engine_set_time_filter(self.h, f_low_pass, order, f_high_pass))

    def set_shot_taper(self, on, srtradius=0.0, srtshape=1, filtsize=0):
        """Circular taper of each shot's gradient around its source
        (SWS_TAPER_CIRCULAR_PER_SHOT); on=False turns it off."""
        L = lib()
        L.ifos_engine_set_shot_taper.argtypes = [C.c_void_p, C.c_int, C.c_float, C.c_int, C.c_int]
        _check(L.ifos_engine_set_shot_taper(self.h, int(bool(on)), srtradius, srtshape, filtsize))

    def energy_weight(self, which=0):
        out = np.zeros((self.cfg.NY, self.cfg.NX), np.float32)
        _check(lib().ifos_engine_energy_weight(self.h, which, _fp(out)))
        return out

    def set_snapshot_precision(self, prec):
        """0 / "fp32", 1 / "bf16" or 2 / "fp16" storage of the forward snapshots."""
        _check(lib().ifos_engine_set_snapshot_precision(self.h, SNAP_PRECISIONS.get(prec, prec)))


SNAP_PRECISIONS = {"fp32": 0, "bf16": 1, "fp16": 2}


def mall_slots(config):
    """Shots per launch whose per-step working set fits the MALL (ifos_engine_mall_slots)."""
    L = lib()
    L.ifos_engine_mall_slots.argtypes = [C.POINTER(Config)]
    n = L.ifos_engine_mall_slots(C.byref(config))
    if n < 1:
        raise IfosError("invalid config")
    return n


def snapshot_bytes(config, prec=0):
    """Device bytes of one shot's snapshot storage at that precision."""
    return int(lib().ifos_snapshot_bytes(C.byref(config), SNAP_PRECISIONS.get(prec, prec)))


class Fp64Engine:
    """The scheme in FP64 on the device with checkpointed forward
    recomputation (include/ifos2d_hip.h, ifos_fp64_*): the yardstick that
    16-bit snapshot runs are measured against.  Same call order as Engine:
    set_model -> set_shot -> shot_gradient -> gradient."""

    def __init__(self, config, device=0, segment_steps=0):
        L = lib()
        L.ifos_fp64_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        L.ifos_fp64_destroy.argtypes = [C.c_void_p]
        L.ifos_fp64_destroy.restype = None
        for name in ("ifos_fp64_set_model", "ifos_fp64_set_shot", "ifos_fp64_shot_gradient", "ifos_fp64_seismograms",
                     "ifos_fp64_gradient", "ifos_fp64_reset_gradient", "ifos_fp64_segment_steps"):
            getattr(L, name).restype = C.c_int
        L.ifos_fp64_set_model.argtypes = [C.c_void_p] + [C.c_void_p] * 5
        L.ifos_fp64_set_shot.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
        L.ifos_fp64_shot_gradient.argtypes = [C.c_void_p] + [C.c_void_p] * 4
        L.ifos_fp64_seismograms.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.ifos_fp64_gradient.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ifos_fp64_reset_gradient.argtypes = [C.c_void_p]
        L.ifos_fp64_segment_steps.argtypes = [C.c_void_p]
        self.cfg = config
        h = C.c_void_p()
        _check(L.ifos_fp64_create(C.byref(config), device, int(segment_steps), C.byref(h)))
        self.h = h
        self.ntr = 0

    def close(self):
        if self.h:
            lib().ifos_fp64_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def segment_steps(self):
        return lib().ifos_fp64_segment_steps(self.h)

    def set_model(self, vp, vs, rho, taus=None, taup=None):
        arrs = [None if a is None else _f32(a) for a in (vp, vs, rho, taus, taup)]
        _check(lib().ifos_fp64_set_model(self.h, *[None if a is None else a.ctypes.data_as(C.c_void_p)
                                                   for a in arrs]))

    def set_shot(self, srcpos, recpos):
        srcpos = _f32(np.asarray(srcpos, np.float32).reshape(-1, SRC_STRIDE))
        recpos = np.ascontiguousarray(np.asarray(recpos, np.int32).reshape(-1, 2))
        self.ntr = recpos.shape[0]
        _check(lib().ifos_fp64_set_shot(self.h, srcpos.shape[0], srcpos.ctypes.data_as(C.c_void_p), self.ntr,
                                        recpos.ctypes.data_as(C.c_void_p)))

    def shot_gradient(self, obs_vx, obs_vy, kill=None):
        ovx, ovy = _f32(obs_vx), _f32(obs_vy)
        k = None if kill is None else np.ascontiguousarray(kill, np.int32)
        mis = np.zeros(2, np.float64)
        _check(lib().ifos_fp64_shot_gradient(self.h, ovx.ctypes.data_as(C.c_void_p), ovy.ctypes.data_as(C.c_void_p),
                                             None if k is None else k.ctypes.data_as(C.c_void_p),
                                             mis.ctypes.data_as(C.c_void_p)))
        return mis

    def seismograms(self):
        vx = np.zeros((self.ntr, self.cfg.NT), np.float64)
        vy = np.zeros_like(vx)
        _check(lib().ifos_fp64_seismograms(self.h, vx.ctypes.data_as(C.c_void_p), vy.ctypes.data_as(C.c_void_p)))
        return vx, vy

    def gradient(self, total=False):
        g = np.zeros((3, self.cfg.NY, self.cfg.NX), np.float64)
        _check(lib().ifos_fp64_gradient(self.h, int(total), *[g[k].ctypes.data_as(C.c_void_p) for k in range(3)]))
        return g

    def reset_gradient(self):
        _check(lib().ifos_fp64_reset_gradient(self.h))


def fp64_bytes(config, segment_steps=0, ntr=0, nsrc=0):
    """Device bytes of an Fp64Engine for that configuration."""
    L = lib()
    L.ifos_fp64_bytes.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.ifos_fp64_bytes.restype = C.c_size_t
    return int(L.ifos_fp64_bytes(C.byref(config), segment_steps, ntr, nsrc))
