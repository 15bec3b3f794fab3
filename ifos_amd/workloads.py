"""Seeded synthetic workloads for the P-SV hot path.

Each builder returns a Workload: the run Config, the true and start models
([NY][NX] float32, x fastest, the orientation of the engine's arrays), the
shots (source rows in grid indices, receiver grid indices, optional kill
masks).  The model files of the reference are x-outer/y-inner float32
(src/readmod_elastic.c:94-125); ifos_amd.formats converts between the two.

Workloads (sizes as named by the benchmark definitions):
  toy_elastic        400x200, dh 1 m, FDORDER 4, one explosive 30 Hz shot
  layered_visc       2000x800, dh 2 m, FDORDER 8, L=2, Q=20, 64 shots
  batched_aperture   1700x350, dh 8 m, FDORDER 8, L=2, 256 shots with
                     per-shot offset windows (trace-kill masks)
  precision_large    4000x1000, dh 2 m, FDORDER 8, L=3, 16 shots
"""
import dataclasses
import math

import numpy as np

from . import make_config, SRC_STRIDE


@dataclasses.dataclass
class Shot:
    srcpos: np.ndarray          # [nsrc][8]
    recpos: np.ndarray          # [ntr][2] int32
    kill: np.ndarray = None     # [ntr] int32 or None


@dataclasses.dataclass
class Workload:
    name: str
    cfg: object
    true: dict
    start: dict
    shots: list
    note: str = ""

    @property
    def cells(self):
        return self.cfg.NX * self.cfg.NY


def holberg_gamma(fdorder, maxrelerror):
    from . import lib
    import ctypes as C
    hc = (C.c_float * 7)()
    lib().ifos_holberg(fdorder, maxrelerror, hc)
    return sum(abs(hc[k]) for k in range(1, 7))


def stable_dt(dh, vmax, fdorder, maxrelerror, fraction=0.5):
    """fraction x the CFL limit DH/(sqrt(2)*gamma*vmax) of
    src/checkfd_ssg_elastic.c:105-106."""
    return np.float32(fraction * dh / (math.sqrt(2.0) * holberg_gamma(fdorder, maxrelerror) * vmax))


def source_row(ix, iy, fc, amp=1.0e10, tshift=0.0, stype=1, azimuth=0.0):
    return np.array([ix, iy, 0.0, tshift, fc, amp, azimuth, stype], np.float32)


def toy_elastic(seed=0, NT=2000, NX=400, NY=200):
    """Plain elastic FWI iteration, FP32: homogeneous background with one
    seeded Gaussian anomaly of +/-10 %, radius 15 cells."""
    rng = np.random.default_rng(seed)
    FW = 20
    vp0, vs0, rho0 = 3500.0, 2000.0, 2000.0
    cx = rng.uniform(FW + 15, NX - FW - 15)
    cy = rng.uniform(15, NY - FW - 15)
    sign = 1.0 if rng.random() < 0.5 else -1.0
    y, x = np.mgrid[1:NY + 1, 1:NX + 1].astype(np.float64)
    gauss = np.exp(-((x - cx) ** 2 + (y - cy) ** 2) / (2 * 15.0 ** 2))
    fac = (1.0 + sign * 0.1 * gauss)
    true = dict(vp=(vp0 * fac).astype(np.float32), vs=(vs0 * fac).astype(np.float32),
                rho=(rho0 * fac).astype(np.float32))
    start = dict(vp=np.full((NY, NX), vp0, np.float32), vs=np.full((NY, NX), vs0, np.float32),
                 rho=np.full((NY, NX), rho0, np.float32))
    vmax = float(true["vp"].max())
    DT = stable_dt(1.0, vmax, 4, 1)
    cfg = make_config(NX=NX, NY=NY, DH=1.0, DT=float(DT), NT=NT, FDORDER=4, MAXRELERROR=1, FW=FW,
                      FREE_SURF=1, VPPML=vp0, FPML=30.0, DTINV=2, VELOCITY=0, L=0, npower=2.0, k_max_PML=1.0)
    rec = np.array([[21 + 2 * k, 5] for k in range(min(180, (NX - 22) // 2))], np.int32)
    shots = [Shot(srcpos=source_row(min(100, NX // 4), min(10, NY // 4), 30.0)[None, :], recpos=rec)]
    return Workload("toy_elastic", cfg, true, start, shots,
                    note=f"anomaly centre ({cx:.1f},{cy:.1f}) sign {sign:+.0f}")


def _layered(rng, NX, NY, nlayers=12, vmin=1800.0, vmax=4500.0, nfaults=3):
    """12 layers with vp increasing with depth, vs = vp/1.8, Gardner density
    (rho = 310 vp^0.25, SI), offset along 3 seeded dipping faults."""
    depth = np.sort(rng.uniform(0.05, 0.95, nlayers - 1)) * NY
    vps = np.sort(rng.uniform(vmin, vmax, nlayers))
    vps[0], vps[-1] = vmin, vmax
    y, x = np.mgrid[1:NY + 1, 1:NX + 1].astype(np.float64)
    shift = np.zeros_like(y)
    for _ in range(nfaults):
        x0 = rng.uniform(0.15, 0.85) * NX
        dip = np.tan(np.radians(rng.uniform(50, 80))) * (1 if rng.random() < 0.5 else -1)
        throw = rng.uniform(0.02, 0.06) * NY
        side = (x - x0) * dip - y > 0
        shift += np.where(side, throw, 0.0)
    idx = np.searchsorted(depth, y - shift)
    vp = vps[np.clip(idx, 0, nlayers - 1)]
    vs = vp / 1.8
    rho = 310.0 * vp ** 0.25
    return vp.astype(np.float32), vs.astype(np.float32), rho.astype(np.float32)


def _smooth(a, r=8):
    k = np.ones(2 * r + 1) / (2 * r + 1)
    b = np.apply_along_axis(lambda v: np.convolve(np.pad(v, r, mode="edge"), k, "valid"), 0, a)
    b = np.apply_along_axis(lambda v: np.convolve(np.pad(v, r, mode="edge"), k, "valid"), 1, b)
    return b.astype(np.float32)


def layered_visc(seed=1, NX=2000, NY=800, DH=2.0, NT=4500, DTINV=3, nshots=64, nrec=400, rec_step=5,
                 FDORDER=8, L=2, Q=20.0, fc=20.0, FW=30):
    """Viscoelastic P-SV gradient: layered model with faults, L relaxation
    mechanisms (FL = 10, 40 Hz), tau = 2/Q (src/readmod.c:194)."""
    rng = np.random.default_rng(seed)
    vp, vs, rho = _layered(rng, NX, NY)
    tau = np.float32(2.0 / Q)
    true = dict(vp=vp, vs=vs, rho=rho, taus=np.full((NY, NX), tau, np.float32),
                taup=np.full((NY, NX), tau, np.float32))
    start = dict(vp=_smooth(vp), vs=_smooth(vs), rho=_smooth(rho), taus=true["taus"], taup=true["taup"])
    vmax = float(vp.max()) * math.sqrt(1.0 + L * float(tau))
    DT = stable_dt(DH, vmax, FDORDER, 1)
    FL = [10.0, 40.0, 80.0][:L]
    # CPML exponent and coordinate stretch as in the reference's own
    # parameter file (par/in_and_out/IFOS2D_FW.json: npower 4, k_max_PML 8)
    cfg = make_config(NX=NX, NY=NY, DH=DH, DT=float(DT), NT=NT, FDORDER=FDORDER, MAXRELERROR=1, FW=FW,
                      FREE_SURF=1, VPPML=float(vp.max()), FPML=fc, DTINV=DTINV, VELOCITY=1, L=L, FL=FL,
                      F_REF=fc, npower=4.0, k_max_PML=8.0)
    rx = np.array([FW + 1 + rec_step * k for k in range(nrec) if FW + 1 + rec_step * k <= NX - FW], np.int32)
    rec = np.stack([rx, np.full_like(rx, 2)], 1).astype(np.int32)
    sx = np.linspace(FW + 5, NX - FW - 5, nshots).round().astype(int)
    shots = [Shot(srcpos=source_row(int(x), 4, fc)[None, :], recpos=rec) for x in sx]
    return Workload("layered_visc", cfg, true, start, shots)


def batched_aperture(seed=2, NX=1700, NY=350, DH=8.0, NT=3000, DTINV=4, nshots=256, FDORDER=8, L=2, fc=20.0,
                     FW=30):
    """Same generator with per-shot receiver apertures: each shot keeps only
    the receivers within a seeded offset of 150-900 cells (as a trace-kill
    mask, src/create_trkill_table.c:21-43)."""
    w = layered_visc(seed=seed, NX=NX, NY=NY, DH=DH, NT=NT, DTINV=DTINV, nshots=nshots, nrec=NX, rec_step=5,
                     FDORDER=FDORDER, L=L, fc=fc, FW=FW)
    rng = np.random.default_rng(seed + 100)
    w.cfg.TRKILL = 1
    shots = []
    for s in w.shots:
        off = rng.uniform(150, 900)
        sx = s.srcpos[0, 0]
        keep = np.abs(s.recpos[:, 0] - sx) <= off
        shots.append(Shot(srcpos=s.srcpos, recpos=s.recpos[keep], kill=None))
    w.shots = shots
    w.name = "batched_aperture"
    return w


def precision_large(seed=3, NX=4000, NY=1000, DH=2.0, NT=8000, DTINV=2, nshots=16, fc=15.0):
    w = layered_visc(seed=seed, NX=NX, NY=NY, DH=DH, NT=NT, DTINV=DTINV, nshots=nshots, nrec=NX // 5, fc=fc, L=3,
                     FW=30)
    w.name = "precision_large"
    return w


def small_case(seed=0, NX=64, NY=48, NT=200, FDORDER=4, L=0, VELOCITY=0, DTINV=2, FREE_SURF=1, stype=1, FW=10,
               nrec=None, amp=1.0e10):
    """Small randomised case for parity tests (fits the oracle in seconds)."""
    rng = np.random.default_rng(seed)
    vp = (3000.0 * (1 + 0.05 * rng.standard_normal((NY, NX)))).astype(np.float32)
    vs = (vp / 1.7 * (1 + 0.03 * rng.standard_normal((NY, NX)))).astype(np.float32)
    rho = (2100.0 * (1 + 0.05 * rng.standard_normal((NY, NX)))).astype(np.float32)
    m = dict(vp=vp, vs=vs, rho=rho)
    if L:
        m["taus"] = (0.1 * (1 + 0.2 * rng.random((NY, NX)))).astype(np.float32)
        m["taup"] = (0.08 * (1 + 0.2 * rng.random((NY, NX)))).astype(np.float32)
    start = {k: v.copy() for k, v in m.items()}
    start["vp"] = np.full_like(vp, 3000.0)
    start["vs"] = np.full_like(vs, 3000.0 / 1.7)
    DH = 2.0
    vmax = float(vp.max()) * math.sqrt(1.0 + L * 0.12)
    DT = stable_dt(DH, vmax, FDORDER, 1)
    cfg = make_config(NX=NX, NY=NY, DH=DH, DT=float(DT), NT=NT, FDORDER=FDORDER, MAXRELERROR=1, FW=FW,
                      FREE_SURF=FREE_SURF, VPPML=3000.0, FPML=25.0, DTINV=DTINV, VELOCITY=VELOCITY, L=L,
                      FL=[10.0, 40.0, 70.0][:max(L, 1)], F_REF=25.0, SOURCE_TYPE=stype)
    nrec = nrec or max(4, (NX - 2 * FW) // 3)
    xs = np.linspace(FW + 2, NX - FW - 2, nrec).round().astype(np.int32)
    xs = np.unique(xs)
    rec = np.stack([xs, np.full_like(xs, 3)], 1).astype(np.int32)
    src = source_row(NX // 3, max(2, NY // 6), 25.0, amp=amp, stype=stype, azimuth=30.0)[None, :]
    return Workload(f"small_{seed}", cfg, m, start, [Shot(srcpos=src, recpos=rec)])
