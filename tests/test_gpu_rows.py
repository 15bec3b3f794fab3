"""Velocity step with two or four cells per thread (adjacent rows of the
64x8 tile per thread of a 64x4 or 64x2 workgroup), with two rows per
thread reading the tile's stresses from an LDS copy, and the one-kernel
step (velocity of the tile and its rim, then its stresses, wavefield sets
alternating between steps): the same bits as one cell per thread
and as the oracle, for every part of the velocity step -- CPML strips on all
sides, force sources and adjoint injection on either row of a thread, grids
whose last tile row is partly outside the model, acceleration snapshots in
32 and 16 bits, the integrated adjoint velocity (VELOCITY 0), the energy
sums of the preconditioner and several shots per workgroup."""
import numpy as np
import pytest

import ifos_amd
from ifos_amd import workloads
import oracle_lib as O
from test_precision import params_of

pytestmark = pytest.mark.gpu

# NY not a multiple of 8 (the last tile row is partly outside the model),
# odd, and 1 or 3 mod 4 (the last thread row has cells outside)
CASES = [dict(seed=61, NX=75, NY=51, NT=150, FDORDER=8, L=2, VELOCITY=1, DTINV=3, stype=1),
         dict(seed=62, NX=70, NY=45, NT=140, FDORDER=4, L=0, VELOCITY=0, DTINV=2, stype=4, FREE_SURF=0),
         dict(seed=63, NX=66, NY=50, NT=130, FDORDER=12, L=1, VELOCITY=1, DTINV=2, stype=2),
         dict(seed=64, NX=64, NY=43, NT=130, FDORDER=2, L=0, VELOCITY=1, DTINV=2, stype=3)]


def run(w, rows, obs=None, slots=1, shots=None, spb=1, grid=0, prec=0, eprecond=0):
    """rows 1, 2, 4: cells per thread; "lds": two rows per thread from LDS;
    "fused": one kernel per time step."""
    eng = ifos_amd.Engine(w.cfg, slots=slots)
    shots = shots or [(w.shots[0].srcpos, w.shots[0].recpos)]
    eng.set_model(**(w.start if obs is not None else w.true))
    eng.set_shots(shots)
    if rows == "fused":
        eng.set_fused(True)
        assert eng.get_fused() == 1
    elif rows == "lds":
        eng.set_velocity_lds(True)
        assert eng.get_velocity_lds() == 1 and eng.get_velocity_rows() == 2
    else:
        eng.set_velocity_rows(rows)
        assert eng.get_velocity_rows() == rows and eng.get_velocity_lds() == 0
    if spb != 1 or grid:
        eng.set_launch(spb, grid)
    if prec:
        eng.set_snapshot_precision(prec)
    if eprecond:
        eng.set_eprecond(eprecond, 0.005, True)
    if obs is None:
        eng.forward()
        out = (eng.seismograms(), eng.fields())
    else:
        eng.shot_gradient(*obs)
        out = (eng.seismograms(), eng.gradient(), eng.energy_weight(0) if eprecond else None)
    eng.close()
    return out


@pytest.mark.parametrize("rows", [2, 4, "lds", "fused"])
@pytest.mark.parametrize("case", CASES, ids=["visc8_odd", "el4_v0_pml_top", "visc12_force", "el2_force"])
def test_rows_match_oracle(gpu_lib, case, rows):
    w = workloads.small_case(FW=10, **case)
    print(f"[rows {rows}] {case}", flush=True)
    p = params_of(w)
    s = w.shots[0]
    ovx, ovy, ofin = O.forward(p, w.true, s.srcpos, s.recpos)
    (vx, vy), fin = run(w, rows)
    np.testing.assert_array_equal(vx, ovx)
    np.testing.assert_array_equal(vy, ovy)
    np.testing.assert_array_equal(fin, ofin)
    ovx[:, 0] = 0
    ovy[:, 0] = 0
    ref = O.gradient(p, w.start, s.srcpos, s.recpos, ovx, ovy)
    _, g, _ = run(w, rows, obs=(ovx, ovy))
    for k, n in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(g[k], ref["grad_" + n], err_msg=n)


@pytest.mark.parametrize("rows", [2, 4, "lds", "fused"])
@pytest.mark.parametrize("prec", [1, 2])
def test_rows_snapshot16(gpu_lib, prec, rows):
    w = workloads.small_case(seed=65, NX=80, NY=47, NT=150, FDORDER=8, L=2, VELOCITY=1, DTINV=3, FW=10)
    print(f"[rows {rows}] 16-bit snapshots prec {prec}", flush=True)
    s = w.shots[0]
    ovx, ovy, _ = O.forward(params_of(w), w.true, s.srcpos, s.recpos)
    want = O.gradient(params_of(w, SNAP_PREC=prec), w.start, s.srcpos, s.recpos, ovx, ovy)
    _, g, _ = run(w, rows, obs=(ovx, ovy), prec=prec)
    for k, n in enumerate(("vp", "vs", "rho")):
        np.testing.assert_array_equal(g[k], want["grad_" + n], err_msg=n)


@pytest.mark.parametrize("rows", [2, 4, "lds", "fused"])
def test_rows_energy_and_batches(gpu_lib, rows):
    w = workloads.small_case(seed=66, NX=90, NY=53, NT=160, FDORDER=6, L=1, VELOCITY=1, DTINV=2, FW=10)
    print(f"[rows {rows}] energy sums, 3 shots, 2 per workgroup, persistent grid", flush=True)
    base = np.array(w.shots[0].srcpos, np.float32).reshape(-1, 8)
    shots = []
    for x, t in ((20, 1), (44, 2), (70, 4)):
        sp = base.copy()
        sp[:, 0] = x
        sp[:, 7] = t
        shots.append((sp, w.shots[0].recpos))
    p = params_of(w)
    obs = [O.forward(p, w.true, sp, rp)[:2] for sp, rp in shots]
    ovx = np.concatenate([o[0] for o in obs])
    ovy = np.concatenate([o[1] for o in obs])
    one = run(w, 1, obs=(ovx, ovy), slots=3, shots=shots, eprecond=1)
    for spb, grid in ((1, 0), (2, 0), (2, 16)):
        two = run(w, rows, obs=(ovx, ovy), slots=3, shots=shots, spb=spb, grid=grid, eprecond=1)
        np.testing.assert_array_equal(two[0][0], one[0][0], err_msg=f"vx spb {spb} grid {grid}")
        np.testing.assert_array_equal(two[1], one[1], err_msg=f"gradient spb {spb} grid {grid}")
        np.testing.assert_array_equal(two[2], one[2], err_msg=f"energy spb {spb} grid {grid}")
