"""Write the golden vectors under tests/golden/ from the reference's own
hot-path kernels (oracle/_ref/libifosref.so, compiled from the reference
sources by oracle/Makefile).

For each case: the forward run on the true model (seismograms vx/vy at every
step, final vx, vy, sxx, syy, sxy) and the per-shot Vp/Vs/rho gradient of the
start model, whose adjoint sources are the oracle's calc_res restatement of
the observed-minus-synthetic residual (the reference's calc_res.c needs FFTW
at link time and is not built).  Observed traces have their first sample
zeroed, as src/inseis.c:75-83 delivers them.

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as O  # noqa: E402
from ifos_amd import workloads  # noqa: E402

CASES = {
    "el4_fs_v0": dict(seed=0, FDORDER=4, L=0, VELOCITY=0, DTINV=2, FREE_SURF=1, stype=1, NT=200),
    "visc8_fs_v1": dict(seed=1, FDORDER=8, L=2, VELOCITY=1, DTINV=3, FREE_SURF=1, stype=1, NT=180),
    "el2_nofs_rot": dict(seed=2, FDORDER=2, L=0, VELOCITY=0, DTINV=2, FREE_SURF=0, stype=4, NT=200),
    "visc6_fs_fx_v1": dict(seed=3, FDORDER=6, L=1, VELOCITY=1, DTINV=1, FREE_SURF=1, stype=2, NT=160),
}


def params_from(cfg):
    d = {k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"}
    d["FL"] = [cfg.FL[i] for i in range(1, 8)]
    return d


def to_oracle(cfgd):
    d = dict(cfgd)
    fl = d.pop("FL")
    return O.make_params(FL=fl, **d)


def make_case(name, kw):
    w = workloads.small_case(**kw)
    cfgd = params_from(w.cfg)
    p = to_oracle(cfgd)
    shot = w.shots[0]
    svx, svy, fin = O.forward(p, w.true, shot.srcpos, shot.recpos, which="ref")
    obs_vx, obs_vy = svx.copy(), svy.copy()
    obs_vx[:, 0] = 0.0
    obs_vy[:, 0] = 0.0
    # synthetic seismograms of the start model -> adjoint sources (oracle calc_res)
    s0vx, s0vy, _ = O.forward(p, w.start, shot.srcpos, shot.recpos, which="ref")
    dvx, _ = O.calc_res(p, s0vx, obs_vx)
    dvy, _ = O.calc_res(p, s0vy, obs_vy)
    if p.ADJOINT_TYPE == 2:
        dvx[:] = 0
    if p.ADJOINT_TYPE == 3:
        dvy[:] = 0
    g = O.ref_gradient(p, w.start, shot.srcpos, shot.recpos, dvx, dvy)
    out = dict(srcpos=shot.srcpos, recpos=shot.recpos, seis_vx=svx, seis_vy=svy, final5=fin,
               obs_vx=obs_vx, obs_vy=obs_vy, grad_vp=g["grad_vp"], grad_vs=g["grad_vs"], grad_rho=g["grad_rho"],
               start_seis_vx=g["seis_vx"], start_seis_vy=g["seis_vy"])
    for k, v in w.true.items():
        out["true_" + k] = v
    for k, v in w.start.items():
        out["start_" + k] = v
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)
    with open(os.path.join(HERE, name + ".json"), "w") as f:
        json.dump(dict(config=cfgd, case=kw, source="oracle/_ref (reference kernels via ref_harness.c)"), f,
                  indent=1, sort_keys=True)
    return out


if __name__ == "__main__":
    if not O.ref_available():
        raise SystemExit("oracle/_ref not built (make -C oracle)")
    for name, kw in CASES.items():
        make_case(name, kw)
        print("wrote", name)
