r"""Mixed-precision single-vector LOCG: low-precision matvec, FP64 everything else.

No counterpart in the reference (``rqutils/ground_locg.py`` runs every step in
the vector dtype and spends four matvecs per iteration, ``:256,346``).  Here:

* the vectors ``x`` (Ritz vector), ``y`` (previous direction) and ``p``
  (residual direction) are FP64 / complex128, and so are their images
  ``Ax, Ay, Ap``, which are carried through every linear combination, so an
  iteration needs exactly one new matvec (``A p``);
* that matvec runs in FP32 / complex64 (``matvec_lo``) in the first phase;
* ``[x, y, p]`` is orthonormalised in FP64 (classical Gram-Schmidt, two
  passes), the 3x3 Rayleigh-Ritz matrix ``V^H A V`` is formed from the carried
  images and solved by the same analytic eigen routine as the plain path;
* once the FP64 residual of the low phase reaches ``switch_tol`` or stops
  improving (the low-precision floor), ``Ax`` and ``Ay`` are recomputed with
  the FP64 matvec (``matvec_hi``) and the iteration continues in FP64 until
  ``tol`` (or the energy test): the returned energy is a Rayleigh quotient of
  FP64 images.

Per iteration: 1 matvec, about 25 FP64 vector passes, 1 host synchronisation
every ``check_every`` iterations.
"""
import time

import torch

from rqutils_amd.plan import eigenpair_small


def _dot(a, b):
    return torch.vdot(a, b) if a.is_complex() else torch.dot(a, b)


def _nrm(v):
    return torch.linalg.vector_norm(v)


class MixedLOCGResult:
    def __init__(self, theta, x, niter, matvecs_lo, matvecs_hi, history):
        self.theta = theta
        self.x = x
        self.niter = niter
        self.matvecs_lo = matvecs_lo
        self.matvecs_hi = matvecs_hi
        self.history = history     # list of (seconds, theta, rel_residual, phase)

    def __iter__(self):
        return iter((self.theta, self.x, self.niter))


def ground_locg_mixed(matvec_hi, matvec_lo, xinit, maxiter=1000, tol=1e-6, switch_tol=1e-6,
                      check_every=1, stagnation=15, record=False, energy_tol=None, energy_checks=3,
                      energy_rel_residual=1e-7, orth=()):
    """Ground eigenpair with a low-precision bulk matvec and FP64 refinement.

    Args:
        matvec_hi: ``v -> A v`` in the vector dtype (float64 / complex128).
        matvec_lo: ``v -> A v`` in low precision (float32 / complex64), or None
            to run the FP64 phase only.
        xinit: start vector (float64 / complex128 tensor).
        maxiter: total iteration limit.
        tol: stop when ``|A x - theta x| <= tol max(|theta|, |A x0|)`` with FP64
            images (``x0`` the normalised start vector).
        switch_tol: leave the low-precision phase at this relative residual.
        check_every: iterations between host-side convergence checks.
        stagnation: also leave the low phase after this many checks without
            improving the best residual by 1.5x (its low-precision floor); a
            residual 100x above the best, a descent after a stall near an
            excited eigenvector, starts a new best.
        record: keep ``(seconds, theta, rel_residual, phase)`` per check.
        energy_tol: if set, also stop in the FP64 phase once the remaining
            decrease of theta, estimated from the last three decreases as a
            geometric tail (Aitken), and the last decrease are both below
            ``energy_tol * |theta|`` for ``energy_checks`` consecutive checks and
            the relative residual is at most ``energy_rel_residual`` (a stalled
            theta near an excited eigenvector is not converged).
        orth: vectors the solution is kept orthogonal to (deflation, as in
            ``rqutils/ground_locg.py:221-222,283-284,304-306``): the start vector
            and every residual direction are projected out of them, and ``x``
            and ``y`` again after each update (their carried images corrected
            with the images of the orth vectors, computed once in FP64).

    Returns:
        ``MixedLOCGResult`` (iterable as ``theta, x, niter``).
    """
    hi = xinit.dtype
    if hi not in (torch.float64, torch.complex128):
        raise ValueError('xinit must be float64 or complex128')
    lo = {torch.float64: torch.float32, torch.complex128: torch.complex64}[hi]
    t_start = time.perf_counter()
    counts = {'lo': 0, 'hi': 0}

    def mv(v, phase):
        counts[phase] += 1
        if phase == 'lo':
            return matvec_lo(v.to(lo)).to(hi)
        return matvec_hi(v)

    phase = 'lo' if matvec_lo is not None else 'hi'
    orth = [o.to(hi) / _nrm(o.to(hi)) for o in orth]
    orth_img = [matvec_hi(o) for o in orth]

    def deflate(v, av=None):
        """v (and its image av) with the orth components removed (two passes)."""
        for _ in range(2):
            for o, ao in zip(orth, orth_img):
                c = _dot(o, v)
                v = v - c * o
                if av is not None:
                    av = av - c * ao
        return v, av

    x = xinit.to(hi)
    if orth:
        x, _ = deflate(x)
    x = x / _nrm(x)
    ax = mv(x, phase)
    scale = float(_nrm(ax))
    y = ay = None
    history = []
    best = float('inf')
    since_best = 0
    theta = None
    niter = 0
    hi_thetas = []
    calm = 0
    while True:
        theta = _dot(x, ax).real
        r = ax - theta.to(hi) * x
        if niter % check_every == 0 or niter >= maxiter:
            # relative to max(|theta|, |A x0|): defined also for a zero eigenvalue
            rel = float(_nrm(r)) / max(abs(float(theta)), scale, 1e-300)
            if record:
                history.append((time.perf_counter() - t_start, float(theta), rel, phase))
            if phase == 'lo':
                if rel > 100 * best or rel < best / 1.5:
                    best, since_best = rel, 0
                else:
                    since_best += 1
                if rel <= switch_tol or since_best >= stagnation or niter >= maxiter:
                    phase = 'hi'
                    ax = mv(x, 'hi')
                    if y is not None:
                        ay = mv(y, 'hi')
                    continue
            else:
                if rel <= tol or niter >= maxiter:
                    break
                if energy_tol is not None:
                    th = float(theta)
                    hi_thetas.append(th)
                    if len(hi_thetas) >= 3:
                        d0 = hi_thetas[-3] - hi_thetas[-2]
                        d1 = hi_thetas[-2] - hi_thetas[-1]
                        lim = energy_tol * abs(th)
                        tail = d1 * d1 / (d0 - d1) if d0 > d1 >= 0 else float('inf')
                        calm = calm + 1 if (0 <= d1 <= lim and tail <= lim) else 0
                        if calm >= energy_checks and rel <= energy_rel_residual:
                            break
        # residual direction, orthonormal to x and y (and the orth vectors), two CGS passes
        p = r
        for _ in range(2):
            for o in orth:
                p = p - o * _dot(o, p)
            p = p - x * _dot(x, p)
            if y is not None:
                p = p - y * _dot(y, p)
        pn = _nrm(p)
        p = p / pn
        ap = mv(p, phase)
        # Rayleigh-Ritz on [x, y, p] from the carried images
        basis = (x, p) if y is None else (x, y, p)
        images = (ax, ap) if y is None else (ax, ay, ap)
        n = len(basis)
        k = torch.empty((n, n), dtype=torch.complex128, device=x.device)
        for a in range(n):
            for b in range(a, n):
                val = 0.5 * (_dot(basis[a], images[b]) + _dot(images[a], basis[b]))
                k[a, b] = val
                k[b, a] = val.conj() if torch.is_tensor(val) and val.is_complex() else val
        _, c = eigenpair_small(k)
        if not x.is_complex():
            c = c.real
        c = c.to(hi)
        xn = sum(ci * v for ci, v in zip(c, basis))
        axn = sum(ci * v for ci, v in zip(c, images))
        d = sum(ci * v for ci, v in zip(c[1:], basis[1:]))
        ad = sum(ci * v for ci, v in zip(c[1:], images[1:]))
        if orth:
            xn, axn = deflate(xn, axn)
            d, ad = deflate(d, ad)
        s = _nrm(xn)
        x, ax = xn / s, axn / s
        proj = _dot(x, d)
        d = d - x * proj
        ad = ad - ax * proj
        dn = _nrm(d)
        if float(dn) > 1e-14:
            y, ay = d / dn, ad / dn
        else:                      # x is already an eigenvector of the subspace
            y = ay = None
        niter += 1
    return MixedLOCGResult(theta, x, niter, counts['lo'], counts['hi'], history)
