"""16-bit snapshot storage in the oracle: the rounding helpers against
independent implementations, and the gradient error of BF16/FP16 snapshots
against an FP64 build of the same scheme (oracle/liboracle64.so, snapshots
kept in FP64).

The bar for 16-bit snapshots is a relative L2 gradient error of at most 1e-3
per parameter (Vp, Vs, rho) against the FP64 build; the FP32 path's own
error is printed beside it.  The snapshot storage has no reference
counterpart (the reference keeps FP32 snapshots), so this precision claim is
pinned only by the FP64 build and the rounding checks here.
"""
import numpy as np
import pytest

import oracle_lib as O
from ifos_amd import workloads


def params_of(w, **over):
    cfg = w.cfg
    kw = {k: getattr(cfg, k) for k, _ in cfg._fields_ if k != "FL"}
    kw.update(over)
    return O.make_params(FL=[cfg.FL[i] for i in range(1, 8)], **kw)


def test_fp16_rounding_matches_numpy():
    rng = np.random.default_rng(3)
    x = np.concatenate([
        rng.standard_normal(3000).astype(np.float32),
        (rng.standard_normal(3000) * 1e-5).astype(np.float32),       # half subnormals
        (rng.standard_normal(1000) * 1e-8).astype(np.float32),       # below half range
        np.array([0.0, -0.0, 1.0, -1.0, 2.0 ** -24, 2.0 ** -25, 3 * 2.0 ** -26, 2.0 ** -14,
                  65504.0, 0.999999, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 2.0 ** -15 * 1.5], np.float32),
    ])
    # exact ties between two halves, both parities of the lower neighbour
    halves = rng.integers(1, 0x7BFF, 2000).astype(np.uint16)
    lo = halves.view(np.float16).astype(np.float32)
    hi = (halves + 1).view(np.float16).astype(np.float32)
    x = np.concatenate([x, ((lo.astype(np.float64) + hi) / 2).astype(np.float32)])
    codes = O.f32_to_16(x, 2)
    want = x.astype(np.float16).view(np.uint16)
    np.testing.assert_array_equal(codes, want)
    np.testing.assert_array_equal(O.f16_to_32(codes, 2), want.view(np.float16).astype(np.float32))


def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.standard_normal(4000) * 10.0 ** rng.integers(-30, 30, 4000),
                        np.array([1.0, -1.0, 0.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8])]).astype(np.float32)
    codes = O.f32_to_16(x, 1)
    bits = x.view(np.uint32)
    lo = (bits >> 16).astype(np.uint32)
    cand_lo = (lo << 16).view(np.float32).astype(np.float64)
    cand_hi = ((lo + 1) << 16).view(np.float32).astype(np.float64)
    xv = x.astype(np.float64)
    dlo, dhi = np.abs(xv - cand_lo), np.abs(cand_hi - xv)
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (lo % 2 == 1))
    want = np.where(pick_hi, lo + 1, lo).astype(np.uint16)
    np.testing.assert_array_equal(codes, want)
    np.testing.assert_array_equal(O.f16_to_32(codes, 1), (codes.astype(np.uint32) << 16).view(np.float32))


def rel(a, b):
    return float(np.linalg.norm((np.asarray(a, np.float64) - b).ravel()) / np.linalg.norm(np.asarray(b).ravel()))


PREC_CASES = [
    dict(seed=31, NX=96, NY=64, NT=402, FDORDER=8, L=2, VELOCITY=1, DTINV=3),
    dict(seed=32, NX=80, NY=60, NT=360, FDORDER=4, L=0, VELOCITY=0, DTINV=2),
]


@pytest.mark.parametrize("case", PREC_CASES, ids=["visc8_v1", "el4_v0"])
def test_snapshot16_gradient_error_vs_fp64(case):
    w = workloads.small_case(**case)
    s = w.shots[0]
    p64 = params_of(w)
    ovx, ovy, _ = O.forward(p64, w.true, s.srcpos, s.recpos, fp64=True)
    ovx, ovy = ovx.astype(np.float32), ovy.astype(np.float32)
    g64 = O.gradient(p64, w.start, s.srcpos, s.recpos, ovx, ovy, fp64=True)
    errs = {}
    for prec in (0, 1, 2):
        g = O.gradient(params_of(w, SNAP_PREC=prec), w.start, s.srcpos, s.recpos, ovx, ovy)
        errs[prec] = [rel(g["grad_" + k], g64["grad_" + k]) for k in ("vp", "vs", "rho")]
    print("rel L2 gradient error vs FP64 (vp, vs, rho):",
          {("fp32", "bf16", "fp16")[k]: [f"{e:.2e}" for e in v] for k, v in errs.items()})
    assert max(errs[0]) < 1e-4
    assert max(errs[1]) <= 1e-3
    assert max(errs[2]) <= 1e-3
    # FP16 keeps 3 more mantissa bits than BF16
    assert max(errs[2]) < max(errs[1])
