"""Seeded synthetic SQD workloads (random Pauli sums, Jordan-Wigner electronic models).

No counterpart in the reference (its example draws unseeded random Pauli
strings, ``examples/sqd.py:131-142``).  Every draw uses
``numpy.random.default_rng(1000 * stream + role)`` (one stream per workload) with role 0 = Hamiltonian,
1 = states, 2 = start vector, 3 = gradient probes.

Jordan-Wigner convention: spin-orbital ``(p, alpha)`` is qubit ``p`` and
``(p, beta)`` is qubit ``M + p`` (blocked spin order, ``M`` spatial orbitals);
occupied = bit 1; ``a_j = Z_0 ... Z_{j-1} |0><1|_j``.  In the operator basis
``Z^z X^x`` (bit masks) a product is
``(Z^z1 X^x1)(Z^z2 X^x2) = (-1)^{popcount(x1 & z2)} Z^{z1^z2} X^{x1^x2}``, and
``Z^z X^x = i^{popcount(x & z)} Q(x, z)`` for the Pauli string ``Q``.
"""
from dataclasses import dataclass
from itertools import combinations
import math

import numpy as np

from rqutils_amd.paulis.symplectic import TermTables

P_SPARSE = (0.7, 0.1, 0.1, 0.1)
P_DENSE36 = (0.6, 0.133, 0.133, 0.134)


def rng_for(stream, role):
    return np.random.default_rng(1000 * stream + role)


def random_pauli_codes(num_qubits, num_terms, probs, rng):
    """(T, n) codes 0=I 1=X 2=Y 3=Z (column c = qubit n-1-c); all-identity rows redrawn."""
    codes = rng.choice(4, size=(num_terms, num_qubits), p=np.asarray(probs) / np.sum(probs))
    while True:
        bad = np.all(codes == 0, axis=1)
        if not bad.any():
            return codes
        codes[bad] = rng.choice(4, size=(int(bad.sum()), num_qubits), p=np.asarray(probs) / np.sum(probs))


def codes_to_strings(codes):
    lut = np.frombuffer(b'IXYZ', dtype=np.uint8)
    return [row.tobytes().decode('ascii') for row in lut[codes]]


def codes_to_masks(codes):
    n = codes.shape[1]
    w = np.uint64(1) << np.arange(n - 1, -1, -1, dtype=np.uint64)
    xm = np.bitwise_or.reduce(np.where((codes == 1) | (codes == 2), w, np.uint64(0)), axis=1)
    zm = np.bitwise_or.reduce(np.where((codes == 2) | (codes == 3), w, np.uint64(0)), axis=1)
    return xm.astype(np.uint64), zm.astype(np.uint64)


def random_subspace(num_qubits, size, rng):
    """``size`` distinct uniform uint64 states in ``[0, 2^n)``, sorted."""
    if size > (1 << num_qubits):
        raise ValueError('subspace larger than the Hilbert space')
    if num_qubits <= 24:
        return np.sort(rng.choice(1 << num_qubits, size=size, replace=False).astype(np.uint64))
    out = np.zeros(0, dtype=np.uint64)
    while out.shape[0] < size:
        draw = rng.integers(0, 1 << num_qubits, size=int(1.1 * (size - out.shape[0])) + 16, dtype=np.uint64)
        out = np.unique(np.concatenate([out, draw]))
    keep = np.sort(rng.choice(out.shape[0], size=size, replace=False))
    return out[keep]


# ---------------------------------------------------------------- Jordan-Wigner

def _ladder_terms(modes, dagger):
    """ZX-basis terms of a_j (dagger=False) / a_j^+ for each mode: list of (z, x, coef) arrays."""
    modes = np.asarray(modes, dtype=np.uint64)
    below = (np.uint64(1) << modes) - np.uint64(1)
    e = np.uint64(1) << modes
    s = -0.5 if dagger else 0.5
    return [(below, e, np.full(modes.shape, 0.5)), (below | e, e, np.full(modes.shape, s))]


def _mul_terms(a, b):
    """Products of every term of ``a`` with every term of ``b`` (elementwise over the op axis)."""
    out = []
    for za, xa, ca in a:
        for zb, xb, cb in b:
            sign = 1. - 2. * (np.bitwise_count(xa & zb) & 1)
            out.append((za ^ zb, xa ^ xb, ca * cb * sign))
    return out


def fermion_product(ops):
    """ZX expansion of a product of ladder operators.

    ``ops``: list of ``(modes, dagger)`` with equal-length mode arrays.
    Returns flat ``(z, x, coef, term_index)``.
    """
    terms = None
    for modes, dagger in ops:
        lt = _ladder_terms(modes, dagger)
        terms = lt if terms is None else _mul_terms(terms, lt)
    z = np.concatenate([t[0] for t in terms])
    x = np.concatenate([t[1] for t in terms])
    c = np.concatenate([t[2] for t in terms])
    n = len(ops[0][0])
    idx = np.tile(np.arange(n), len(terms))
    return z, x, c, idx


def _zx_to_tables(z, x, c, num_qubits, rtol=1e-12):
    """Aggregate ZX-basis terms and convert to Pauli strings with real coefficients."""
    order = np.lexsort((z, x))
    z, x, c = z[order], x[order], c[order]
    first = np.concatenate([[True], (x[1:] != x[:-1]) | (z[1:] != z[:-1])])
    starts = np.nonzero(first)[0]
    csum = np.add.reduceat(c.astype(np.complex128), starts)
    z, x = z[starts], x[starts]
    alpha = csum * (1j ** (np.bitwise_count(x & z) & 3))
    scale = np.max(np.abs(alpha)) if alpha.size else 1.
    keep = np.abs(alpha) > rtol * scale
    z, x, alpha = z[keep], x[keep], alpha[keep]
    if alpha.size and np.max(np.abs(alpha.imag)) > 1e-9 * scale:
        raise AssertionError('Jordan-Wigner Hamiltonian is not Hermitian')
    return TermTables.from_masks(x, z, alpha.real, num_qubits)


def random_integrals(num_spatial, rng):
    """h symmetric N(0,1)/4; g with 8-fold symmetry, entries N(0, 0.05^2) (chemist order (pq|rs))."""
    M = num_spatial
    h = rng.standard_normal((M, M)) / 4.
    h = np.triu(h) + np.triu(h, 1).T
    raw = rng.standard_normal((M, M, M, M)) * 0.05
    p, q, r, s = np.meshgrid(*(np.arange(M),) * 4, indexing='ij')
    a0, a1 = np.minimum(p, q), np.maximum(p, q)
    b0, b1 = np.minimum(r, s), np.maximum(r, s)
    ka = a0 * M + a1
    kb = b0 * M + b1
    swap = ka > kb
    c0 = np.where(swap, b0, a0)
    c1 = np.where(swap, b1, a1)
    c2 = np.where(swap, a0, b0)
    c3 = np.where(swap, a1, b1)
    g = raw[c0, c1, c2, c3]
    return h, g


def jw_hamiltonian(h, g, rtol=1e-12):
    """Qubit Hamiltonian of ``sum h_pq a+_ps a_qs + 1/2 sum g_pqrs a+_ps a+_rt a_st a_qs``.

    Spin-orbital ``(p, s)`` is qubit ``p + s M`` (s = 0 alpha, 1 beta).
    """
    M = h.shape[0]
    n = 2 * M
    zs, xs, cs = [], [], []
    # one-body
    p, q = np.meshgrid(np.arange(M), np.arange(M), indexing='ij')
    p, q = p.reshape(-1), q.reshape(-1)
    hv = h[p, q]
    for sp in (0, 1):
        P = (p + sp * M).astype(np.uint64)
        Q = (q + sp * M).astype(np.uint64)
        z, x, c, idx = fermion_product([(P, True), (Q, False)])
        zs.append(z)
        xs.append(x)
        cs.append(c * hv[idx])
    # two-body
    grid = np.meshgrid(*(np.arange(M),) * 4, indexing='ij')
    p, q, r, s = (a.reshape(-1) for a in grid)
    gv = 0.5 * g[p, q, r, s]
    for sp in (0, 1):
        for st in (0, 1):
            P = (p + sp * M).astype(np.uint64)
            Q = (q + sp * M).astype(np.uint64)
            R = (r + st * M).astype(np.uint64)
            S = (s + st * M).astype(np.uint64)
            ok = (P != R) & (Q != S) & (gv != 0)
            if not ok.any():
                continue
            z, x, c, idx = fermion_product([(P[ok], True), (R[ok], True), (S[ok], False), (Q[ok], False)])
            zs.append(z)
            xs.append(x)
            cs.append(c * gv[ok][idx])
    return _zx_to_tables(np.concatenate(zs), np.concatenate(xs), np.concatenate(cs), n, rtol=rtol)


def weight_strings(num_bits, weight):
    """All ``num_bits``-bit integers of the given popcount, ascending."""
    vals = [sum(1 << b for b in comb) for comb in combinations(range(num_bits), weight)]
    return np.sort(np.array(vals, dtype=np.uint64))


def product_subspace(num_spatial, n_alpha, n_beta, n_elec_alpha, n_elec_beta, rng):
    """Sampled alpha strings x sampled beta strings -> sorted uint64 states ``alpha | beta << M``."""
    A = weight_strings(num_spatial, n_elec_alpha)
    B = weight_strings(num_spatial, n_elec_beta)
    a = np.sort(A[rng.choice(A.shape[0], size=n_alpha, replace=False)])
    b = np.sort(B[rng.choice(B.shape[0], size=n_beta, replace=False)])
    states = (b[:, None] << np.uint64(num_spatial)) | a[None, :]
    return states.reshape(-1)


# ---------------------------------------------------------------- named workloads

@dataclass
class Workload:
    name: str
    tables: TermTables
    states: np.ndarray          # sorted unique uint64
    num_qubits: int
    strings: list | None = None   # Pauli strings (small random configs only)
    alphas: np.ndarray | None = None


def random20_300t():
    """20-qubit, 300-term random Pauli sum (I/X/Y/Z 0.7/0.1/0.1/0.1, N(0,1)); 8192 states."""
    rng = rng_for(0, 0)
    codes = random_pauli_codes(20, 300, P_SPARSE, rng)
    alpha = rng.standard_normal(300)
    xm, zm = codes_to_masks(codes)
    tables = TermTables.from_masks(xm, zm, alpha, 20)
    states = random_subspace(20, 8192, rng_for(0, 1))
    return Workload('random-20q-300t-8192', tables, states, 20, codes_to_strings(codes), alpha)


def jw_model(num_spatial, n_alpha, n_beta, n_elec, stream):
    rng = rng_for(stream, 0)
    h, g = random_integrals(num_spatial, rng)
    tables = jw_hamiltonian(h, g)
    states = product_subspace(num_spatial, n_alpha, n_beta, n_elec, n_elec, rng_for(stream, 1))
    return Workload(f'jw-{2 * num_spatial}so-{n_alpha}x{n_beta}', tables, states, 2 * num_spatial)


def jw32(n_alpha=1000, n_beta=1000):
    """JW random model, 16 spatial orbitals, 8a+8b electrons, 1000 x 1000 determinants."""
    return jw_model(16, n_alpha, n_beta, 8, 1)


def jw40(n_alpha=5000, n_beta=5000):
    """JW random model, 20 spatial orbitals, 10a+10b electrons, 5000 x 5000 determinants."""
    return jw_model(20, n_alpha, n_beta, 10, 4)


def random36_10kt(num_states=1_000_000, num_terms=10_000):
    """36-qubit, 10k-term random Pauli sum (0.6/0.133/0.133/0.134) on 1M random states."""
    rng = rng_for(3, 0)
    codes = random_pauli_codes(36, num_terms, P_DENSE36, rng)
    alpha = rng.standard_normal(num_terms)
    xm, zm = codes_to_masks(codes)
    tables = TermTables.from_masks(xm, zm, alpha, 36)
    states = random_subspace(36, num_states, rng_for(3, 1))
    return Workload(f'random-36q-{num_terms}t-{num_states}', tables, states, 36, None, alpha)


def batch40_problem(index):
    """Problem ``index`` of the 64-problem batch: 40 qubits, T ~ U[500, 5000], N log-U[1e3, 1e6]."""
    rng = np.random.default_rng(2000 + 17 * index)
    num_terms = int(rng.integers(500, 5001))
    num_states = int(round(math.exp(rng.uniform(math.log(1e3), math.log(1e6)))))
    codes = random_pauli_codes(40, num_terms, P_SPARSE, rng)
    alpha = rng.standard_normal(num_terms)
    xm, zm = codes_to_masks(codes)
    tables = TermTables.from_masks(xm, zm, alpha, 40)
    states = random_subspace(40, num_states, np.random.default_rng(2001 + 17 * index))
    return Workload(f'batch-p{index}-40q-{num_terms}t-{num_states}', tables, states, 40, None, alpha)


def fingerprint(w):
    """Sizes and SHA-256 prefixes of a generated workload's tables and states, so that a
    stored result (a reference ground energy) can be matched to the exact instance."""
    import hashlib

    def sha(a):
        return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]

    t = w.tables
    return {
        'n_states': int(w.states.shape[0]),
        'n_groups': int(t.num_groups),
        'n_terms': int(t.num_terms),
        'states_sha256': sha(w.states),
        'xmask_sha256': sha(t.xmask),
        'grp_off_sha256': sha(t.grp_off),
        'zmask_sha256': sha(t.zmask),
        'coef_abs_sum': float(np.sum(np.abs(t.coef))),
    }
