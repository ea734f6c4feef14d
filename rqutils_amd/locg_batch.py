r"""Batched single-vector LOCG: every problem of a ``BatchPlan`` iterates in lock step.

No counterpart in the reference (``ground_locg`` solves one problem per call,
``rqutils/ground_locg.py:128-162``).  The step is the mixed-precision LOCG of
``locg_mixed`` with every per-problem scalar kept on the device:

* one batched matvec per iteration (one launch for all problems, persistent
  tile queue), complex64 / float32 in the bulk phase, the vector dtype in the
  refinement phase;
* the step is ``rqx_locg_step``: eight fused passes over the vectors (each a
  linear update with per-problem coefficients plus the inner products the
  next stage needs), small per-problem scalar kernels between them (the 3x3
  Rayleigh-Ritz eigenpairs among them) and the matvec;
* converged problems (residual ``|A x - theta x| <= tol max(|theta|, |A x0|)``,
  ``x0`` the normalised start vector) are frozen: their coefficients become
  (1, 0, 0) so their vectors stop changing;
* no host synchronisation inside a step, so the step is captured once in a
  HIP graph (``torch.cuda.CUDAGraph``) and replayed; convergence is polled
  every ``check_every`` steps.
"""
import ctypes
import time
import warnings

import torch

from rqutils_amd import _capi
from rqutils_amd.plan import Plan, eigenpair_batch

_LO = {torch.float64: torch.float32, torch.complex128: torch.complex64}


class BatchLOCG:
    """Persistent buffers + one LOCG step for all problems of a plan (``BatchPlan`` or single ``Plan``)."""

    def __init__(self, plan: Plan, x0: torch.Tensor, tol=1e-7, switch_tol=1e-6, mixed=True, lo_storage=True, orth=()):
        if x0.dtype not in _LO:
            raise ValueError('x0 must be float64 or complex128')
        if x0.shape != (plan.n_total,) or x0.device != plan.device:
            raise ValueError('x0 must be a device vector over all problems')
        self.plan = plan
        self.K = plan.n_problems
        self.hi = x0.dtype
        self.lo = _LO[x0.dtype]
        self.tol = tol
        self.switch_tol = switch_tol
        dev = plan.device
        z = lambda: torch.zeros_like(x0)  # noqa: E731
        self.x, self.ax, self.y, self.ay, self.p, self.ap, self.r = (z() for _ in range(7))
        self.p_lo = torch.zeros(plan.n_total, dtype=self.lo, device=dev)
        self.ap_lo = torch.zeros_like(self.p_lo)
        self.active = torch.ones(self.K, dtype=torch.bool, device=dev)
        self.theta = torch.zeros(self.K, dtype=torch.float64, device=dev)
        self.rel = torch.full((self.K,), float('inf'), dtype=torch.float64, device=dev)
        self.one = torch.ones(self.K, dtype=torch.complex128, device=dev)
        self.zero = torch.zeros(self.K, dtype=torch.complex128, device=dev)
        self.work = torch.zeros(self.K * _capi.RQX_LOCG_WORK, dtype=torch.float64, device=dev)
        self._vec_ptrs = {}
        self.phase = 'lo' if mixed else 'hi'
        # lo_storage (default): the bulk phase keeps its vectors in low precision
        # (the step's inner products and coefficients stay FP64) and x, y are
        # widened and re-orthonormalised at the switch; otherwise the vectors stay
        # FP64 with low copies of p and A p around the low-precision matvec
        # (measured: the same time to solution at 1M rows, 8% slower at 25M)
        self.lo_storage = bool(mixed and lo_storage)
        if self.lo_storage:
            zl = lambda: torch.zeros(plan.n_total, dtype=self.lo, device=dev)  # noqa: E731
            self.x_lo, self.ax_lo, self.y_lo, self.ay_lo = (zl() for _ in range(4))
        # deflation vectors (each over all rows: per problem, its rows are that problem's
        # vector), normalised per problem, with their images in the vector dtype
        self.orth, self.orth_img = [], []
        for o in orth:
            o = torch.as_tensor(o, device=dev).to(self.hi).contiguous()
            if o.shape != (plan.n_total,):
                raise ValueError('orth vectors must be device vectors over all rows')
            o = o.clone()
            self._normalize(o, None)
            self.orth.append(o)
            self.orth_img.append(plan.matvec(o))
        if self.lo_storage and self.orth:
            self.orth_lo = [o.to(self.lo) for o in self.orth]
            self.orth_img_lo = [a.to(self.lo) for a in self.orth_img]
        self.x.copy_(x0)
        self._deflate_vectors(self.x)
        self._normalize(self.x, None)
        self._apply(self.x, self.ax)
        # residual scale: max(|theta|, |A x0|) (|A x0| sizes the operator, so a problem
        # whose ground energy is 0 still converges); read by the fused step from its work slot
        self.scale = torch.sqrt(plan.dots([(self.ax, self.ax)])[:, 0].real)
        self.work.view(self.K, _capi.RQX_LOCG_WORK)[:, _capi.RQX_LOCG_SCALE] = self.scale
        self._first_step()
        if self.lo_storage:
            for lo_v, hi_v in ((self.x_lo, self.x), (self.ax_lo, self.ax), (self.y_lo, self.y), (self.ay_lo, self.ay)):
                lo_v.copy_(hi_v)

    # ------------------------------------------------------------------ helpers
    def _deflate_vectors(self, v, av=None, orth=None, imgs=None):
        """Remove the orth components of v (per problem), and the same combination of the
        orth images from its image av."""
        orth = self.orth if orth is None else orth
        imgs = self.orth_img if imgs is None else imgs
        for o, ao in zip(orth, imgs):
            c = self.plan.dots([(o, v)])[:, 0]
            coef = torch.stack([self.one, -c], 1)
            self.plan.combine([v, o], coef, v)
            if av is not None:
                self.plan.combine([av, ao], coef, av)

    def _deflate(self, lo_copies=None):
        """x and y (with their images) orthogonal to the orth vectors again, in the storage
        of the current phase (no host synchronisation: part of the captured step)."""
        if not self.orth:
            return
        if lo_copies is None:
            lo_copies = self.phase == 'lo' and self.lo_storage
        if lo_copies:
            vecs, orth, imgs = ((self.x_lo, self.ax_lo), (self.y_lo, self.ay_lo)), self.orth_lo, self.orth_img_lo
        else:
            vecs, orth, imgs = ((self.x, self.ax), (self.y, self.ay)), self.orth, self.orth_img
        for v, av in vecs:
            self._deflate_vectors(v, av, orth, imgs)

    def _apply(self, v, out):
        if self.phase == 'lo':
            self.p_lo.copy_(v)
            self.plan.matvec(self.p_lo, out=self.ap_lo)
            out.copy_(self.ap_lo)
        else:
            self.plan.matvec(v, out=out)

    def _normalize(self, v, av):
        n2 = self.plan.dots([(v, v)])[:, 0].real
        inv = torch.where(n2 > 0, 1.0 / torch.sqrt(n2), torch.zeros_like(n2)).to(torch.complex128)
        self.plan.combine([v], inv[:, None], v)
        if av is not None:
            self.plan.combine([av], inv[:, None], av)

    def _residual(self):
        """theta, r = Ax - theta x, relative residual, p = normalised r orthogonal to x and y."""
        plan = self.plan
        th = plan.dots([(self.x, self.ax)])[:, 0].real
        self.theta.copy_(th)
        plan.combine([self.ax, self.x], torch.stack([self.one, -th.to(torch.complex128)], 1), self.r)
        rn = torch.sqrt(plan.dots([(self.r, self.r)])[:, 0].real)
        self.rel.copy_(rn / torch.maximum(th.abs(), self.scale).clamp_min(1e-300))
        self.active &= self.rel > self.tol
        self.p.copy_(self.r)
        for _ in range(2):
            d = plan.dots([(self.x, self.p), (self.y, self.p)])
            plan.combine([self.p, self.x, self.y], torch.stack([self.one, -d[:, 0], -d[:, 1]], 1), self.p)
        self._normalize(self.p, None)

    def _update(self, c):
        """x <- c0 x + c1 y + c2 p (and images); y <- normalised (c1 y + c2 p) orthogonal to x."""
        plan = self.plan
        keep = self.active[:, None]
        ident = torch.stack([self.one, self.zero, self.zero], 1)
        c = torch.where(keep, c, ident)
        plan.combine([self.y, self.p], c[:, 1:], self.y)
        plan.combine([self.ay, self.ap], c[:, 1:], self.ay)
        plan.combine([self.x, self.y], torch.stack([c[:, 0], self.one], 1), self.x)
        plan.combine([self.ax, self.ay], torch.stack([c[:, 0], self.one], 1), self.ax)
        n2 = plan.dots([(self.x, self.x)])[:, 0].real
        inv = (1.0 / torch.sqrt(n2)).to(torch.complex128)
        plan.combine([self.x], inv[:, None], self.x)
        plan.combine([self.ax], inv[:, None], self.ax)
        pr = plan.dots([(self.x, self.y)])[:, 0]
        plan.combine([self.y, self.x], torch.stack([self.one, -pr], 1), self.y)
        plan.combine([self.ay, self.ax], torch.stack([self.one, -pr], 1), self.ay)
        self._normalize(self.y, self.ay)

    def _first_step(self):
        """Rayleigh-Ritz on [x, p] (no previous direction yet)."""
        self._residual()
        self._apply(self.p, self.ap)
        d = self.plan.dots([(self.x, self.ax), (self.x, self.ap), (self.p, self.ax), (self.p, self.ap)])
        m = torch.empty((self.K, 2, 2), dtype=torch.complex128, device=self.plan.device)
        m[:, 0, 0] = d[:, 0].real
        m[:, 1, 1] = d[:, 3].real
        m[:, 0, 1] = 0.5 * (d[:, 1] + d[:, 2].conj())
        m[:, 1, 0] = m[:, 0, 1].conj()
        _, c2 = eigenpair_batch(m)
        if not self.x.is_complex():
            c2 = c2.real.to(torch.complex128)
        c = torch.stack([c2[:, 0], self.zero, c2[:, 1]], 1)
        self._update(c)
        self._deflate(lo_copies=False)        # the low copies are made after this step

    def step(self):
        """One LOCG iteration for every problem (no host synchronisation): ``rqx_locg_step``.

        Eight fused passes over the vectors, scalar stages on the device and one
        matvec; the same sequence as ``step_reference``.
        """
        lo_store = self.phase == 'lo' and self.lo_storage
        key = (self.phase, lo_store)
        if key not in self._vec_ptrs:
            vecs = ((self.x_lo, self.ax_lo, self.y_lo, self.ay_lo, self.p_lo, self.ap_lo) if lo_store else
                    (self.x, self.ax, self.y, self.ay, self.p, self.ap, self.p_lo, self.ap_lo))
            self._vec_ptrs[key] = (ctypes.c_void_p * 8)(*[t.data_ptr() for t in vecs])
        dt = self.lo if lo_store else self.hi
        dtype = {torch.complex128: _capi.RQX_C128, torch.float64: _capi.RQX_F64,
                 torch.complex64: _capi.RQX_C64, torch.float32: _capi.RQX_F32}[dt]
        low = int(self.phase == 'lo' and not lo_store)
        stream = _capi.current_stream(self.plan.device)
        _capi.check(_capi.lib().rqx_locg_step(self.plan._handle, dtype, low, self._vec_ptrs[key],
                                              self.theta.data_ptr(), self.rel.data_ptr(), self.active.data_ptr(),
                                              self.work.data_ptr(), float(self.tol), stream))
        self._deflate()

    def step_reference(self):
        """The same iteration from the per-problem vector kernels and torch (one launch per operation)."""
        if self.lo_storage and self.phase == 'lo':
            raise ValueError('step_reference runs on FP64 vectors: build the solver with lo_storage=False')
        plan = self.plan
        self._residual()
        self._apply(self.p, self.ap)
        d = plan.dots([(self.x, self.ax), (self.x, self.ay), (self.x, self.ap),
                       (self.y, self.ax), (self.y, self.ay), (self.y, self.ap),
                       (self.p, self.ax), (self.p, self.ay), (self.p, self.ap)]).reshape(self.K, 3, 3)
        m = 0.5 * (d + d.conj().transpose(1, 2))
        _, c = eigenpair_batch(m)
        if not self.x.is_complex():
            c = c.real.to(torch.complex128)
        self._update(c)
        self._deflate()

    def probe(self):
        """FP64 Rayleigh quotient and relative residual of every problem's current x, from one
        matvec in the vector dtype (the solver's state is unchanged).

        In the bulk phase the step's own theta comes from a low-precision image (its error is
        ~1e-8..1e-7 relative); this one is exact for the vector held, so it certifies the
        energy.  With low-precision storage the FP64 ``x`` and ``Ax`` buffers are idle in the
        bulk phase and hold the widened, normalised ``x`` and its image; otherwise two
        scratch vectors are used.
        """
        if self.lo_storage and self.phase == 'lo':
            q, aq = self.x, self.ax
            q.copy_(self.x_lo)
        else:
            if not hasattr(self, '_probe_q'):
                self._probe_q = torch.zeros_like(self.x)
                self._probe_aq = torch.zeros_like(self.x)
            q, aq = self._probe_q, self._probe_aq
            q.copy_(self.x)
        self._normalize(q, None)
        self.plan.matvec(q, out=aq)
        th = self.plan.dots([(q, aq)])[:, 0].real
        self.plan.combine([aq, q], torch.stack([self.one, -th.to(torch.complex128)], 1), aq)
        rn = torch.sqrt(self.plan.dots([(aq, aq)])[:, 0].real)
        return th, rn / torch.maximum(th.abs(), self.scale).clamp_min(1e-300)

    def to_low_precision(self):
        """Back to the bulk precision (after an escape from an excited-state plateau): x, y
        and their FP64 images rounded into the low-precision storage."""
        if self.lo_storage:
            for lo_v, hi_v in ((self.x_lo, self.x), (self.ax_lo, self.ax), (self.y_lo, self.y), (self.ay_lo, self.ay)):
                lo_v.copy_(hi_v)
        self.phase = 'lo'

    def to_high_precision(self):
        """Switch to the vector dtype: x and y widened and re-orthonormalised in FP64, images recomputed."""
        if self.lo_storage and self.phase == 'lo':
            self.x.copy_(self.x_lo)
            self.y.copy_(self.y_lo)
            self._normalize(self.x, None)
            pr = self.plan.dots([(self.x, self.y)])[:, 0]
            self.plan.combine([self.y, self.x], torch.stack([self.one, -pr], 1), self.y)
            self._normalize(self.y, None)
        self.phase = 'hi'
        self._deflate(lo_copies=False)
        self._apply(self.x, self.ax)
        self._apply(self.y, self.ay)


def batch_ground_locg(plan: Plan, x0, tol=1e-7, switch_tol=1e-6, maxiter=1000, check_every=10,
                      use_graph=True, mixed=True, record=False, energy_tol=None, energy_checks=2,
                      energy_rel_residual=1e-7, floor_polls=3, lo_storage=True, check_every_hi=1, orth=(),
                      energy_probe=None, floor_gate=1e-4):
    """Ground eigenpairs of every problem of ``plan`` (a ``BatchPlan``, or a ``Plan`` as one problem).

    The bulk phase (low-precision matvec) runs until every active problem's
    relative residual, polled every ``check_every`` steps, is at most
    ``switch_tol`` or has reached its low-precision floor: its best value is at
    most ``floor_gate`` and has not improved by 1.5x at ``floor_polls``
    consecutive polls (a residual 100x above the best so far, a descent after a
    stall near an excited eigenvector, starts a new best; a stall above
    ``floor_gate`` is the iteration's, not the precision's).  The FP64 phase then polls every ``check_every_hi`` steps (its
    start vector is converged to the FP32 floor, so a few steps finish it).

    Near an excited eigenvector the residual can fall to the switch level and then
    rise again as the ground state's component grows (an excited-state plateau of
    LOCG): a problem whose FP64-phase residual rises 100x above its best and above
    ``floor_gate`` sends the solve back to the bulk phase (at most four times).

    A problem stops (its vectors freeze) at relative residual ``tol``, or, with
    ``energy_tol``, in the FP64 phase once its theta has settled -- the last
    decrease and the geometric (Aitken) estimate of the remaining decrease both
    below ``energy_tol * |theta|`` at ``energy_checks`` consecutive polls -- and
    its relative residual is at most ``energy_rel_residual`` (near an excited
    eigenvector theta can sit still for tens of iterations with a residual of
    ~1e-6 before it falls to the ground energy).  Differences of theta at the
    rounding level count as settled (|decrease| <= ``energy_tol * |theta|``).

    ``energy_probe``: a relative residual below which every bulk-phase poll also
    computes each problem's FP64 Rayleigh quotient and residual of its current
    ``x`` (``BatchLOCG.probe``, one FP64 matvec): the bulk phase converges the
    vector as fast as an FP64 run does, but its own theta carries the low-precision
    image's error, so these probes are what certify its energy along the way.
    They are returned in ``info['probes']`` as ``(seconds, iteration, theta64,
    rel64)`` and do not change the iteration.

    ``orth``: vectors over all rows that every problem's solution is kept
    orthogonal to (per problem, its own rows of each vector; deflation as in
    ``rqutils/ground_locg.py:221-222,283-284,304-306``): the start vector is
    projected out of them, and ``x`` and ``y`` again after every step.

    Returns ``(theta (K,), x, iterations, info)`` where ``info`` holds per-problem
    relative residuals, the iteration of the precision switch and timings.
    """
    t0 = time.perf_counter()
    solver = BatchLOCG(plan, x0, tol=tol, switch_tol=switch_tol, mixed=mixed, lo_storage=lo_storage, orth=orth)
    graphs = {}

    def run_step():
        if not use_graph:
            solver.step()
            return
        g = graphs.get(solver.phase)
        if g is None:
            # warm up on a side stream (allocator), then capture one step
            s = torch.cuda.Stream(device=plan.device)
            s.wait_stream(torch.cuda.current_stream(plan.device))
            with torch.cuda.stream(s):
                solver.step()
            torch.cuda.current_stream(plan.device).wait_stream(s)
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                solver.step()
            graphs[solver.phase] = g
            return
        g.replay()

    K = solver.K
    it = 1
    switch_at = None
    history = []
    thetas = []
    probes = []
    calm = torch.zeros(K, dtype=torch.int64)
    best = torch.full((K,), float('inf'), dtype=torch.float64)
    best_hi = torch.full((K,), float('inf'), dtype=torch.float64)
    back = 0
    since = torch.zeros(K, dtype=torch.int64)
    while it < maxiter:
        for _ in range(check_every if switch_at is None else check_every_hi):
            run_step()
            it += 1
        # one device -> host copy per poll: active flags, theta, relative residuals
        status = torch.cat([solver.active.to(torch.float64), solver.theta, solver.rel]).cpu()
        act = status[:K] > 0.5
        th = status[K:2 * K]
        rel = status[2 * K:]
        active = int(act.sum())
        worst = float(rel[act].max()) if active else 0.0
        if record:
            history.append((time.perf_counter() - t0, it, active, worst, solver.phase))
        if solver.phase == 'lo':
            restart = rel > 100 * best
            better = rel < best / 1.5
            since = torch.where(restart | better, torch.zeros_like(since), since + 1)
            best = torch.where(restart | better, rel, best)
            if energy_probe is not None and bool((act & (rel <= energy_probe)).any()):
                th64, rel64 = solver.probe()
                probes.append((time.perf_counter() - t0, it, th64.cpu(), rel64.cpu()))
            ready = ~act | (rel <= solver.switch_tol) | ((since >= floor_polls) & (best <= floor_gate))
            if bool(ready.all()) or it >= maxiter // 2:
                solver.to_high_precision()
                solver.active.fill_(True)
                switch_at = it
            continue
        best_hi = torch.minimum(best_hi, torch.where(act, rel, torch.full_like(rel, float('inf'))))
        escaped = act & (rel > 100 * best_hi) & (rel > floor_gate)
        if mixed and back < 4 and bool(escaped.any()):
            solver.to_low_precision()
            back += 1
            switch_at = None
            since.zero_()
            best = rel.clone()
            best_hi.fill_(float('inf'))
            thetas.clear()
            calm.zero_()
            continue
        if energy_tol is None:
            if active == 0:
                break
            continue
        thetas.append(th)
        settled = torch.zeros(K, dtype=torch.bool)
        if len(thetas) >= 3:
            d0, d1 = thetas[-3] - thetas[-2], thetas[-2] - thetas[-1]
            lim = energy_tol * th.abs()
            # a rise at the rounding level is settled too (tail 0); a decrease that is not
            # shrinking has no geometric tail (inf)
            tail = torch.where((d0 > d1) & (d1 > 0), d1 * d1 / (d0 - d1).clamp_min(1e-300),
                               torch.where(d1 <= 0, torch.zeros_like(d1), torch.full_like(d1, float('inf'))))
            ok = (d1.abs() <= lim) & (tail <= lim)
            calm = torch.where(ok, calm + 1, torch.zeros_like(calm))
            settled = (calm >= energy_checks) & (rel <= energy_rel_residual)
        if bool((settled & act).any()):
            solver.active.copy_((act & ~settled).to(plan.device))
            active = int((act & ~settled).sum())
        if active == 0:
            break
    if solver.phase == 'lo' and solver.lo_storage:
        solver.x.copy_(solver.x_lo)
    torch.cuda.synchronize(plan.device)
    converged = ~solver.active.cpu()        # stopped by the residual or the energy test
    if not bool(converged.all()):
        warnings.warn(f'batch_ground_locg: {int((~converged).sum())} of {solver.K} problems not converged '
                      f'after {it} iterations', RuntimeWarning, stacklevel=2)
    info = {'rel_residual': solver.rel.clone(), 'switch_iteration': switch_at, 'back_to_bulk': back, 'iterations': it,
            'seconds': time.perf_counter() - t0, 'history': history, 'graph': use_graph,
            'converged': converged, 'probes': probes}
    return solver.theta.clone(), solver.x, it, info
