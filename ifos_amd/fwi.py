"""One FWI iteration on the GPU, in the order of the reference's driver
(src/IFOS2D.c:1076-3876, PARAMETERIZATION 1, LNORM 2):

1. gradient: forward with snapshots, residual, adjoint and chain rule for
   every shot, summed in shot order (the engine; several shots per launch
   when it has more than one slot); the misfit of the test shots gives
   L2t[1] = L2 / energy (src/IFOS2D.c:2829-2888);
2. direction: PCG per parameter (src/PCG.c) or L-BFGS (src/LBFGS.c) on the
   device (ifos_amd.opt);
3. step length: test models m - eps*dir (calc_mat_change_test, itest = 1),
   forward runs of the test shots, the step search of
   src/IFOS2D.c:3206-3740 and the parabola of calc_opt_step;
4. update: calc_mat_change_test with the chosen step (itest = 0), and for
   L-BFGS the stored model difference.

The host side only sequences calls: every array operation runs in the gfx950
library.  Misfits are FP64 sums (the reference's are FP32).
"""
from dataclasses import dataclass, field

import numpy as np

from . import DeviceBuffer, Engine, mall_slots
from .opt import Optimizer, update_params


@dataclass
class FWISettings:
    EPS_SCALE: float = 0.01      # first test step: 1 % of the largest model value
    STEPMAX: int = 4             # test forwards before giving up (src/IFOS2D.c:531)
    SCALEFAC: float = 2.0        # step scale factor (src/IFOS2D.c:532)
    TESTSHOT_START: int = 1      # test shots, 1-based (src/IFOS2D.c:1064)
    TESTSHOT_END: int = 1
    TESTSHOT_INCR: int = 1
    GRAD_METHOD: int = 1         # 1 PCG, 2 L-BFGS
    PCG_iter_start: int = 1
    N_LBFGS: int = 5
    LBFGS_iter_start: int = 1
    MIN_ITER: int = 0
    C: tuple = (1.0, 1.0, 1.0)   # PCG scaling of the Vp, Vs, rho gradients
    limits: dict = field(default_factory=dict)  # calc_mat_change_test bounds (VPUPPERLIM, ...)


class FWI:
    def __init__(self, workload, settings=None, slots=0, snap="fp32", device=0):
        """slots: shots advanced per launch; 0 sizes it to the MALL
        (ifos_engine_mall_slots, at least 2, at most the number of shots).
        Gradients are the same bits for every count."""
        self.w = workload
        self.s = settings or FWISettings()
        cfg = workload.cfg
        if not slots:
            slots = max(1, min(len(workload.shots), max(2, mall_slots(cfg))))
        self.cfg = cfg
        self.n = cfg.NX * cfg.NY
        self.eng = Engine(cfg, device=device, slots=slots)
        self.eng.set_snapshot_precision(snap)
        self.slots = slots
        self.opt = Optimizer(device)
        n4 = 4 * self.n
        self.model = {k: np.ascontiguousarray(v, np.float32) for k, v in workload.start.items()}
        self.start = {k: v.copy() for k, v in self.model.items()}
        # device state of the optimizer: p_last / c_last per parameter, L-BFGS memory
        self.p_last = [DeviceBuffer(n4) for _ in range(3)]
        self.c_last = [DeviceBuffer(n4) for _ in range(3)]
        self.dir = [DeviceBuffer(n4) for _ in range(3)]
        N = self.s.N_LBFGS
        self.s_hist = DeviceBuffer(4 * N * 3 * self.n)
        self.y_hist = DeviceBuffer(4 * N * 3 * self.n)
        self.s_hist.upload(np.zeros(N * 3 * self.n, np.float32))
        self.y_hist.upload(np.zeros(N * 3 * self.n, np.float32))
        self.rho_hist = np.zeros(N, np.float64)
        self.work = DeviceBuffer(4 * 2 * 3 * self.n)
        self.obs = None
        self.history = []

    def close(self):
        self.eng.close()
        self.opt.close()

    # ------------------------------------------------------------------ data
    def observe(self):
        """Observed traces: forward runs on the true model, first sample
        zeroed as src/inseis.c:75-83 delivers them."""
        self.eng.set_model(**self.w.true)
        self.obs = []
        for batch in self._batches(range(len(self.w.shots))):
            self.eng.set_shots([(self.w.shots[k].srcpos, self.w.shots[k].recpos) for k in batch])
            self.eng.forward(snapshots=False)
            vx, vy = self.eng.seismograms()
            vx[:, 0] = 0.0
            vy[:, 0] = 0.0
            t0 = 0
            for k in batch:
                n = len(self.w.shots[k].recpos)
                self.obs.append((vx[t0:t0 + n], vy[t0:t0 + n]))
                t0 += n

    def _batches(self, shots):
        shots = list(shots)
        return [shots[i:i + self.slots] for i in range(0, len(shots), self.slots)]

    def _kill(self, batch):
        if not self.cfg.TRKILL:
            return None
        return np.concatenate([self.w.shots[k].kill if self.w.shots[k].kill is not None
                               else np.zeros(len(self.w.shots[k].recpos), np.int32) for k in batch])

    def _obs(self, batch):
        return (np.concatenate([self.obs[k][0] for k in batch]), np.concatenate([self.obs[k][1] for k in batch]))

    def test_shots(self):
        s = self.s
        return list(range(s.TESTSHOT_START - 1, min(s.TESTSHOT_END, len(self.w.shots)), s.TESTSHOT_INCR))

    # -------------------------------------------------------------- gradient
    def gradient(self, model):
        """Total gradient over all shots and the test-shot misfit L2 / energy."""
        self.eng.set_model(**model)
        self.eng.reset_gradient()
        tests = set(self.test_shots())
        l2 = en = 0.0
        for batch in self._batches(range(len(self.w.shots))):
            self.eng.set_shots([(self.w.shots[k].srcpos, self.w.shots[k].recpos) for k in batch])
            ovx, ovy = self._obs(batch)
            self.eng.shot_gradient(ovx, ovy, kill=self._kill(batch))
            per = self.eng.shot_misfit()
            for j, k in enumerate(batch):
                if k in tests:
                    l2 += per[j, 0]
                    en += per[j, 1]
        return self.eng.gradient(total=True), l2 / en

    def test_misfit(self, model):
        """L2 / energy of the test shots for a trial model (forward only)."""
        self.eng.set_model(**model)
        l2 = en = 0.0
        for batch in self._batches(self.test_shots()):
            self.eng.set_shots([(self.w.shots[k].srcpos, self.w.shots[k].recpos) for k in batch])
            self.eng.forward(snapshots=False)
            ovx, ovy = self._obs(batch)
            m = self.eng.residual(ovx, ovy, kill=self._kill(batch))
            l2 += m[0]
            en += m[1]
        return l2 / en

    # ------------------------------------------------------------- direction
    def direction(self, grads, it):
        """PCG (per parameter) or L-BFGS directions in self.dir (device)."""
        s = self.s
        for k in range(3):
            self.dir[k].upload(np.ascontiguousarray(grads[k], np.float32))
        if s.GRAD_METHOD == 1:
            betas = []
            for k in range(3):
                g = DeviceBuffer.from_array(np.ascontiguousarray(grads[k], np.float32))
                betas.append(self.opt.pcg(self.n, g, s.C[k], self.p_last[k], self.c_last[k], it, s.PCG_iter_start,
                                          self.dir[k]))
                g.free()
            return betas
        # L-BFGS works on vs, rho, vp normalised by their averages (src/LBFGS.c)
        avg = self._averages()
        order = [1, 2, 0]
        self.opt.lbfgs(self.n, [self.dir[k] for k in order], [avg[k] for k in order], self.s_hist, self.y_hist,
                       self.rho_hist, self.work, s.N_LBFGS, it, s.LBFGS_iter_start)
        return None

    def _averages(self):
        m = self.model
        return [float(np.mean(m["vp"])), float(np.mean(m["vs"])), float(np.mean(m["rho"]))]

    # ---------------------------------------------------------------- update
    def trial(self, eps_scale, it, itest):
        """calc_mat_change_test: the model for step eps_scale (itest 1) or the
        final update (itest 0)."""
        u = update_params(iter=it, eps_scale=eps_scale, GRAD_METHOD=self.s.GRAD_METHOD,
                          LBFGS_iter_start=self.s.LBFGS_iter_start, **self.s.limits)
        dev = {k: DeviceBuffer.from_array(self.model[k]) for k in ("vp", "vs", "rho")}
        d0 = {k: DeviceBuffer.from_array(self.start[k]) for k in ("vp", "vs", "rho")}
        out = {k: DeviceBuffer(4 * self.n) for k in ("vp", "vs", "rho")}
        eps = self.opt.model_update(self.n, u, dev["vp"], dev["vs"], dev["rho"], self.dir[0], self.dir[1],
                                    self.dir[2], d0["vp"], d0["vs"], d0["rho"], out["vp"], out["vs"], out["rho"])
        shape = (self.cfg.NY, self.cfg.NX)
        new = dict(self.model)
        for k in ("vp", "vs", "rho"):
            new[k] = out[k].to_array(shape)
        if itest == 0 and self.s.GRAD_METHOD == 2:
            avg = self._averages()
            order = [1, 2, 0]
            names = ["vp", "vs", "rho"]
            self.opt.lbfgs_store_s(self.n, [out[names[k]] for k in order], [dev[names[k]] for k in order],
                                   [avg[k] for k in order], self.s_hist, self.s.N_LBFGS, it)
        return new, eps

    def step_search(self, L2_1, it):
        """src/IFOS2D.c:3206-3740 (PCG branch): three misfits around the
        minimum, then calc_opt_step."""
        s = self.s
        L2t = [L2_1, 0.0, 0.0]
        epst = [0.0, 0.0, 0.0]
        step1 = step2 = 0
        eps_scale = s.EPS_SCALE
        countstep = 0
        itest = 2
        forwards = 0
        while step2 != 1 or step1 != 1:
            trial, _ = self.trial(eps_scale, it, 1)
            L2t[itest - 1] = self.test_misfit(trial)
            forwards += 1
            if step1 == 0 and L2t[0] <= L2t[1]:
                eps_scale = eps_scale / s.SCALEFAC
                countstep += 1
            if step1 == 1 and L2t[1] < L2t[2]:
                epst[2] = eps_scale
                step2 = 1
            if step1 == 1 and L2t[1] >= L2t[2]:
                epst[2] = eps_scale
                eps_scale = eps_scale + eps_scale / s.SCALEFAC
                countstep += 1
            if step1 == 0 and L2t[0] > L2t[1]:
                epst[1] = eps_scale
                step1 = 1
                itest = 3
                countstep = 0
                eps_scale = eps_scale + eps_scale / s.SCALEFAC
            if step1 == 0 and countstep > s.STEPMAX:
                break
            if step1 == 1 and countstep > s.STEPMAX:
                step1 = step2 = 1
        if step1 == 1:
            opteps = self.opt.step_length(np.array(L2t, np.float32), np.array(epst, np.float32))
        else:
            opteps = s.EPS_SCALE if it < s.MIN_ITER else 0.0
        return opteps, L2t, epst, forwards

    def iteration(self, it):
        """One FWI iteration (it is 1-based); returns a summary dict."""
        if self.obs is None:
            self.observe()
        grads, L2_1 = self.gradient(self.model)
        betas = self.direction(grads, it)
        opteps, L2t, epst, forwards = self.step_search(L2_1, it)
        new, eps = self.trial(opteps, it, 0)
        self.model = new
        rec = dict(iter=it, misfit=L2_1, L2t=L2t, epst=epst, opteps=opteps, eps=eps.tolist(), betas=betas,
                   test_forwards=forwards)
        self.history.append(rec)
        return rec
