"""Oracle: NumPy FP64 restatement of ``rqutils/ground_locg.py`` (test infrastructure only).

Follows ``_ground_locg_callable`` (``ground_locg.py:197-423``) step by step:
iteration 0 (``:280-291``), iteration 1 with a 2x2 Rayleigh-Ritz (``:293-322``),
the main body with ``_project_out`` and a 3x3 Rayleigh-Ritz (``:324-383``), the
relative stopping rule (``:371-376``), ``_project_out`` (``:426-464``) and the
analytic eigenpairs (``:467-514``).
"""
import numpy as np


def _dot(a, b):
    return np.sum(np.conjugate(a) * b)


def project_out(basis, vector, normalize=True, rounds=2):
    """``ground_locg.py:426-464``."""
    vector = vector.copy()
    for iround in range(2 * rounds):
        ips = [_dot(vb, vector) for vb in basis]
        for vb, ip in zip(basis, ips):
            vector = vector - vb * ip
        if normalize and iround < rounds:
            norm = np.linalg.norm(vector)
            vector = vector / (1. if norm == 0. else norm)
    if normalize:
        vector = vector * float(np.linalg.norm(vector) >= 0.99)
    return vector


def eigenpair_2x2(mat):
    """``ground_locg.py:467-478``: lowest eigenpair, mediant-form eigenvector."""
    d = np.diagonal(mat).real
    det = np.prod(d) - np.abs(mat[1, 0]) ** 2
    tr = np.sum(d)
    eigval = (tr - np.sqrt(tr * tr - 4. * det)) * 0.5
    eigvec = np.array([(d[1] - eigval + np.conjugate(mat[1, 0])) / (d[0] - eigval + mat[1, 0]),
                       -1.], dtype=np.result_type(mat.dtype, np.float64))
    eigvec = eigvec / np.sqrt(np.abs(eigvec[0]) ** 2 + 1.)
    return eigval, eigvec


def eigenpair_3x3(mat):
    """``ground_locg.py:481-514``: Kopp's Cardano eigenvalue + cross-product eigenvector."""
    d = np.diagonal(mat).real
    modod = np.abs(mat[[1, 2, 2], [0, 0, 1]]) ** 2
    c2 = -np.sum(d)
    c1 = np.sum(d * np.roll(d, 1)) - np.sum(modod)
    c0 = np.sum(d * modod[::-1])
    c0 -= np.prod(d)
    c0 -= 2. * (mat[0, 2] * mat[1, 0] * mat[2, 1]).real
    p = c2 * c2 - 3. * c1
    q = -13.5 * c0 - c2 * c2 * c2 + 4.5 * c2 * c1
    phi = np.arctan2(np.sqrt(27. * (0.25 * c1 * c1 * (p - c1) + c0 * (q + 6.75 * c0))), q) / 3.
    cphi = np.cos(phi)
    sphi = np.sin(phi)
    xmin = min(2. * cphi, -cphi - np.sqrt(3.) * sphi, -cphi + np.sqrt(3.) * sphi)
    eigval = np.sqrt(p) / 3. * xmin - c2 / 3.
    v0 = np.array(mat[:, 1], dtype=np.complex128)
    v0[1] -= eigval
    v1 = np.array(mat[:, 2], dtype=np.complex128)
    v1[2] -= eigval
    eigvec = np.conjugate(np.cross(v0, v1))
    eigvec = eigvec / np.sqrt(np.sum(np.abs(eigvec) ** 2))
    if not np.iscomplexobj(mat):
        eigvec = eigvec.real
    return eigval, eigvec


def ground_locg(matvec, xinit, orth=(), maxiter=1000, tol=None, dtype=None, ndim=None,
                trace=None):
    """``ground_locg.py:128-162,197-423`` for a callable ``matvec``; returns (theta, x, niter).

    ``trace`` (a list) receives ``theta`` after every main-body iteration.
    """
    if not isinstance(orth, tuple):
        orth = (orth,)
    orth = tuple(v / np.linalg.norm(v) for v in orth)
    if isinstance(xinit, (int, np.integer)):
        xinit = (np.arange(ndim) == xinit).astype(dtype)
        if orth:
            xinit = project_out(orth, xinit)
    xinit = np.array(xinit)
    if tol is None:
        tol = float(np.finfo(xinit.dtype).eps)

    def compute_sas(*vectors):
        nv = len(vectors)
        sas = np.zeros((nv, nv), dtype=vectors[0].dtype)
        mvs = []
        for iv1, v1 in enumerate(vectors):
            mv1 = matvec(v1)
            mvs.append(mv1)
            for iv2 in range(iv1 + 1, nv):
                sas[iv2, iv1] = _dot(vectors[iv2], mv1)
        sas = sas + np.conjugate(sas).T
        for iv1, (v1, mv1) in enumerate(zip(vectors, mvs)):
            sas[iv1, iv1] = _dot(v1, mv1)
        return sas

    def rayleigh_ritz(*vectors):
        sas = compute_sas(*vectors)
        return eigenpair_2x2(sas) if len(vectors) == 2 else eigenpair_3x3(sas)

    xcurr = xinit / np.linalg.norm(xinit)
    # iteration 0
    ax = matvec(xcurr)
    if orth:
        ax = project_out(orth, ax, normalize=False)
    theta = _dot(xcurr, ax).real
    rcurr = ax - theta * xcurr
    # iteration 1
    norm_r = np.linalg.norm(rcurr)
    if norm_r == 0.:
        ycurr = np.zeros_like(xcurr)
    else:
        tmp_p = rcurr / norm_r
        theta, kappa = rayleigh_ritz(xcurr, tmp_p)
        tmp_t = tmp_p * kappa[0] - xcurr * kappa[1]
        tmp_u = xcurr * kappa[0] + tmp_p * kappa[1]
        xnext = tmp_u / np.linalg.norm(tmp_u)
        ynext = tmp_t / np.linalg.norm(tmp_t)
        if orth:
            xnext = project_out(orth, xnext)
        rnext = matvec(xnext) - theta * xnext
        if orth:
            rnext = project_out(orth, rnext, normalize=False)
        xcurr, ycurr, rcurr = xnext, ynext, rnext
    niter = 0
    converged = False
    while niter < maxiter and not converged:
        tmp_p = project_out((xcurr, ycurr), rcurr)
        theta, kappa = rayleigh_ritz(xcurr, ycurr, tmp_p)
        tmp_s = ycurr * kappa[1] + tmp_p * kappa[2]
        norm_s = np.linalg.norm(tmp_s)
        tmp_t = tmp_s * (kappa[0] / norm_s) - xcurr * norm_s
        tmp_u = xcurr * kappa[0] + tmp_s
        xnext = tmp_u / np.linalg.norm(tmp_u)
        ynext = tmp_t / np.linalg.norm(tmp_t)
        if orth:
            xnext = project_out(orth, xnext)
        axnext = matvec(xnext)
        rnext = axnext - xnext * theta
        if orth:
            rnext = project_out(orth, rnext, normalize=False)
        reltol = (np.linalg.norm(axnext) - theta) * xcurr.shape[0] * 10
        converged = bool(np.linalg.norm(rnext) < tol * reltol)
        xcurr, ycurr, rcurr = xnext, ynext, rnext
        niter += 1
        if trace is not None:
            trace.append(float(theta))
    return float(theta), xcurr, niter
