r"""Single-vector LOBPCG (LOCG) on the GPU.

Drop-in for ``rqutils.ground_locg`` (``rqutils/ground_locg.py:128-514``): the
same signature, iteration sequence, stopping rule and analytic Rayleigh-Ritz.
Vectors are torch tensors on the GPU; the small eigenpairs run in one HIP
thread (``rqx_eigenpair_small``), so the only host synchronisation per
iteration is the convergence flag.

Iteration (``ground_locg.py:44-81``), with :math:`R_A` the Rayleigh-Ritz step::

    p <- project_out((x, y), r)
    theta, kappa <- R_A[x, y, p]              (3 matvecs)
    s <- kappa1 y + kappa2 p
    t <- (kappa0/|s|) s - |s| x ;  u <- kappa0 x + s
    x <- u/|u| ; y <- t/|t| ; r <- A x - theta x   (1 matvec)

and stop when ``|r| < tol * 10 N (|Ax| - theta)`` (``:371-376``).
"""
from collections.abc import Callable
import logging
from typing import Optional

import numpy as np
import torch

from rqutils_amd.plan import eigenpair_small

LOG = logging.getLogger(__name__)

_NP_TO_TORCH = {
    np.dtype(np.float64): torch.float64,
    np.dtype(np.complex128): torch.complex128,
    np.dtype(np.float32): torch.float32,
    np.dtype(np.complex64): torch.complex64,
}


def torch_dtype(dtype):
    if isinstance(dtype, torch.dtype):
        return dtype
    return _NP_TO_TORCH[np.dtype(dtype)]


def _dot(a, b):
    """``sum(conj(a) * b)`` as a 0-dim device tensor."""
    return torch.vdot(a, b) if a.is_complex() else torch.dot(a, b)


def _norm(v):
    return torch.linalg.vector_norm(v)


def _project_out(basis, vector, normalize=True, rounds=2):
    """Classical Gram-Schmidt, ``ground_locg.py:426-464``: ``rounds`` normalised
    rounds, ``rounds`` plain rounds, and a zeroed result if its norm is < 0.99."""
    for iround in range(2 * rounds):
        ips = [_dot(vb, vector) for vb in basis]
        for vb, ip in zip(basis, ips):
            vector = vector - vb * ip
        if normalize and iround < rounds:
            norm = _norm(vector)
            vector = vector / torch.where(norm == 0., torch.ones_like(norm), norm)
    if normalize:
        vector = vector * (_norm(vector) >= 0.99).to(vector.dtype)
    return vector


def eigenpair_2x2(mat):
    """Lowest eigenpair of a 2x2 Hermitian matrix (``ground_locg.py:467-478``)."""
    theta, kappa = eigenpair_small(mat)
    return theta, kappa


def eigenpair_3x3(mat):
    """Lowest eigenpair of a 3x3 Hermitian matrix via Kopp's Cardano method
    (``ground_locg.py:481-514``)."""
    theta, kappa = eigenpair_small(mat)
    return theta, kappa


def _default_device():
    if not torch.cuda.is_available():
        raise RuntimeError('ground_locg needs a ROCm GPU (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


def ground_locg(
    mat: Callable | torch.Tensor | np.ndarray,
    xinit,
    args: tuple = (),
    orth=(),
    maxiter: int = 1000,
    tol: Optional[float] = None,
    vspace=None,
    log_level: int = logging.WARNING,
    *,
    callback=None
):
    r"""Single-vector LOBPCG (``ground_locg.py:128-162``).

    Args:
        mat: Matrix :math:`A` (tensor/array) or a callable :math:`x \mapsto Ax`
            (``mat(x, *args)`` must keep the shape and dtype of ``x``).
        xinit: Initial vector, or an integer index of a one-hot start vector
            (requires ``vspace`` if ``mat`` is callable).
        args: Additional arguments to callable ``mat``.
        orth: Vectors to orthogonalize the result against.
        maxiter: Maximum number of gradient descent iterations.
        tol: Convergence condition (default: machine epsilon of the vector dtype).
        vspace: ``(dimension, dtype[, device])`` of the vector space.
        log_level: Verbosity level.
        callback: optional ``callback(niter, theta)`` after every main-loop
            iteration (``theta`` a Python float; keyword-only extension).

    Returns:
        The smallest eigenvalue (0-dim tensor), its eigenvector and the number
        of main-loop iterations.
    """
    if not isinstance(orth, tuple):
        orth = (orth,)
    if callable(mat) and not isinstance(mat, (torch.Tensor, np.ndarray)):
        return _ground_locg_callable(mat, xinit, args, orth, maxiter, tol, vspace, log_level, callback)
    return _ground_locg_matrix(mat, xinit, orth, maxiter, tol, log_level, callback)


def _ground_locg_matrix(mat, xinit, orth, maxiter, tol, log_level, callback=None):
    """Dense-matrix path (``ground_locg.py:165-185``)."""
    if isinstance(xinit, torch.Tensor) and xinit.is_cuda:
        device = xinit.device
    elif isinstance(mat, torch.Tensor) and mat.is_cuda:
        device = mat.device
    else:
        device = _default_device()
    m = torch.as_tensor(mat).to(device)
    vspace = None
    if _is_index(xinit):
        vspace = (m.shape[1], m.dtype, device)

    def matvec(x):
        return torch.mv(m, x)

    return _ground_locg_callable(matvec, xinit, (), orth, maxiter, tol, vspace, log_level, callback)


def _is_index(xinit):
    if isinstance(xinit, (int, np.integer)):
        return True
    if isinstance(xinit, (torch.Tensor, np.ndarray)):
        return xinit.ndim == 0 and not (xinit.dtype.is_floating_point if isinstance(xinit, torch.Tensor)
                                        else np.issubdtype(xinit.dtype, np.inexact)) and \
            not (isinstance(xinit, torch.Tensor) and xinit.is_complex())
    return False


def _ground_locg_callable(matvec, xinit, args, orth, maxiter, tol, vspace, log_level, callback=None):
    """``ground_locg.py:197-423``."""
    if _is_index(xinit):
        ndim, dtype = vspace[0], torch_dtype(vspace[1])
        device = torch.device(vspace[2]) if len(vspace) == 3 else _default_device()
        idx = torch.as_tensor(xinit, device=device)
        xinit = (torch.arange(ndim, device=device) == idx).to(dtype)
        orth = tuple(torch.as_tensor(v, device=device) for v in orth)
        orth = tuple(v / _norm(v) for v in orth)
        if orth:
            xinit = _project_out(orth, xinit)
    else:
        if not isinstance(xinit, torch.Tensor):
            xinit = torch.as_tensor(np.asarray(xinit), device=_default_device())
        device = xinit.device
        orth = tuple(torch.as_tensor(v, device=device) for v in orth)
        orth = tuple(v / _norm(v) for v in orth)

    if tol is None:
        tol = float(torch.finfo(xinit.dtype).eps)
    is_complex = xinit.is_complex()

    def mv(v):
        return matvec(v, *args)

    def compute_sas(*vectors):
        nv = len(vectors)
        sas = torch.zeros((nv, nv), dtype=torch.complex128, device=device)
        mvs = []
        for iv1, v1 in enumerate(vectors):
            mv1 = mv(v1)
            mvs.append(mv1)
            for iv2 in range(iv1 + 1, nv):
                sas[iv2, iv1] = _dot(vectors[iv2], mv1)
        sas = sas + sas.conj().T
        for iv1, (v1, mv1) in enumerate(zip(vectors, mvs)):
            sas[iv1, iv1] = _dot(v1, mv1)
        return sas

    def rayleigh_ritz(*vectors):
        sas = compute_sas(*vectors)
        theta, kappa = eigenpair_small(sas)
        if not is_complex:
            kappa = kappa.real
        return theta.to(torch.float64), kappa.to(xinit.dtype)

    if LOG.isEnabledFor(logging.DEBUG) or log_level <= logging.DEBUG:
        LOG.debug('Performing first LOBPCG steps')

    xcurr = xinit / _norm(xinit)

    # iteration 0 (ground_locg.py:280-291)
    ax = mv(xcurr)
    if orth:
        ax = _project_out(orth, ax, normalize=False)
    theta = _dot(xcurr, ax).real
    rcurr = ax - theta.to(xcurr.dtype) * xcurr

    # iteration 1 (ground_locg.py:293-322)
    norm_r = _norm(rcurr)
    if float(norm_r) == 0.:
        ycurr = torch.zeros_like(xcurr)
    else:
        tmp_p = rcurr / norm_r
        theta, kappa = rayleigh_ritz(xcurr, tmp_p)
        tmp_t = tmp_p * kappa[0] - xcurr * kappa[1]
        tmp_u = xcurr * kappa[0] + tmp_p * kappa[1]
        xnext = tmp_u / _norm(tmp_u)
        ynext = tmp_t / _norm(tmp_t)
        if orth:
            xnext = _project_out(orth, xnext)
        rnext = mv(xnext) - theta.to(xnext.dtype) * xnext
        if orth:
            rnext = _project_out(orth, rnext, normalize=False)
        xcurr, ycurr, rcurr = xnext, ynext, rnext

    # main loop (ground_locg.py:324-383,412-416)
    niter = 0
    converged = False
    n = xcurr.shape[0]
    while niter < maxiter and not converged:
        if log_level <= logging.DEBUG:
            LOG.debug('LOCG iteration %d', niter)
        tmp_p = _project_out((xcurr, ycurr), rcurr)
        theta, kappa = rayleigh_ritz(xcurr, ycurr, tmp_p)
        tmp_s = ycurr * kappa[1] + tmp_p * kappa[2]
        norm_s = _norm(tmp_s)
        tmp_t = tmp_s * (kappa[0] / norm_s) - xcurr * norm_s
        tmp_u = xcurr * kappa[0] + tmp_s
        xnext = tmp_u / _norm(tmp_u)
        ynext = tmp_t / _norm(tmp_t)
        if orth:
            xnext = _project_out(orth, xnext)
        axnext = mv(xnext)
        rnext = axnext - xnext * theta.to(xnext.dtype)
        if orth:
            rnext = _project_out(orth, rnext, normalize=False)
        reltol = (_norm(axnext) - theta) * n * 10
        norm_rnext = _norm(rnext)
        conv_t = norm_rnext < tol * reltol
        xcurr, ycurr, rcurr = xnext, ynext, rnext
        niter += 1
        converged = bool(conv_t)
        if callback is not None:
            callback(niter, float(theta))
        if log_level <= logging.DEBUG:
            LOG.debug('Residual %s, reltol %s, converged: %s', float(norm_rnext), float(reltol), converged)

    return theta, xcurr, niter
