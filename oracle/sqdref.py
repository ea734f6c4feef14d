"""Oracle: NumPy restatement of ``rqutils/sqd.py`` (test infrastructure only).

Every function works on the reference's packed-byte layout (``B`` uint8 per
row, big endian, pad bit first) exactly like the reference, so that the
product's uint64 representation is checked against an independent
representation (the index-parity lemma: XOR commutes with the constant
bit shift, so ``rank(S[i]^x)`` agrees between layouts for real rows).
"""
import numpy as np


def pack_states(states):
    """``sqd.py:255``: prepend the pad bit and pack big endian."""
    states = np.asarray(states).astype(np.uint8)
    return np.packbits(np.pad(states, ((0, 0), (1, 0))), axis=1)


def _lexsort_rows(rows):
    """Stable lexicographic row order (column 0 most significant)."""
    if rows.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    return np.lexsort(rows.T[::-1])


def uniquify_states(states_p, states_size):
    """``sqd.py:445-482``: lexsort, dedupe, fill with 255 up to ``states_size``."""
    states_p = np.asarray(states_p, dtype=np.uint8)
    perm = _lexsort_rows(states_p)
    srt = states_p[perm]
    is_unique = np.any(srt[1:] != srt[:-1], axis=1)
    total_unique = int(np.sum(is_unique)) + 1
    out = np.full((states_size, states_p.shape[1]), 255, dtype=np.uint8)
    keep = np.concatenate([[True], is_unique])
    uniq = srt[keep]
    out[:total_unique] = uniq[:states_size]
    return out


def get_xsource(xsignature, states_p):
    """``sqd.py:485-548`` (sorted input): sort-merge join of ``[S; S^X]``.

    ``A[i]`` is the row index of ``S[i] ^ X`` in ``S`` or -1.
    """
    states_p = np.asarray(states_p, dtype=np.uint8)
    size = states_p.shape[0]
    mapped = np.bitwise_xor(states_p, np.asarray(xsignature, dtype=np.uint8))
    joined = np.concatenate([states_p, mapped], axis=0)
    idx_sorted = _lexsort_rows(joined)          # stable (``:507``)
    joined_sorted = joined[idx_sorted]
    same = np.all(joined_sorted[:-1] == joined_sorted[1:], axis=1)
    source_idx = np.where(same, idx_sorted[1:] - size, -1).astype(np.int64)
    # cumsum(bincount(cumsum(mask))) nonzero trick (``:533-539``)
    csum = np.cumsum(idx_sorted < size)
    counts = np.bincount(csum, minlength=size + 1)[:size]
    tposition = np.cumsum(counts)
    xsource = np.full(size, -1, dtype=np.int64)
    ok = tposition < source_idx.shape[0]
    xsource[ok] = source_idx[tposition[ok]]
    return xsource


def get_diagonal(zsignatures, coeffs, states_p):
    """``sqd.py:603-628``: ``sum_k c_k (1 - 2 (popc(s & z_k) & 1))`` up to the first zero coef."""
    states_p = np.asarray(states_p, dtype=np.uint8)
    out = np.zeros(states_p.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
    for k in range(coeffs.shape[0]):
        if coeffs[k] == 0:
            break
        anded = np.bitwise_and(states_p, zsignatures[k])
        parity = np.sum(np.unpackbits(anded, axis=1), axis=1) & 1
        out = out + coeffs[k] * (1. - 2. * parity)
    return out


def apply_xgrp(xsource, diagonal, vec):
    """``sqd.py:631-640``: gather with fill 0 at -1, times the diagonal."""
    xvec = np.where(xsource >= 0, vec[np.maximum(xsource, 0)], 0)
    return xvec * diagonal


def xsources_all(x, states_u):
    """``sqd.py:347-351``: the J x N source-index cache."""
    return np.stack([get_xsource(x[j], states_u) for j in range(x.shape[0])])


def apply_h_x_cached(vec, xsources, z, c, states_u):
    """``sqd.py:694-709`` with ``out_block = 0`` and ``offset = 0`` (``:746-751``)."""
    out = np.zeros(vec.shape[0], dtype=np.result_type(vec.dtype, c.dtype))
    for j in range(xsources.shape[0]):
        diagonal = get_diagonal(z[j], c[j], states_u)
        out = out + apply_xgrp(xsources[j], diagonal, vec)
    return out


def vinit(x, z, c, states_u):
    """``sqd.py:406-426``: argmin of the group-0 diagonal over real rows, else 0."""
    if np.all(x[0] == 0):
        diagonal = get_diagonal(z[0], c[0], states_u).real
        diagonal = np.where(states_u[:, 0] == 255, np.inf, diagonal)
        return int(np.argmin(diagonal))
    return 0


def subspace_dim(states_u):
    """``sqd.py:440``: number of real (pad bit 0) rows."""
    return int(np.searchsorted(states_u[:, 0] >> 7, 1))


# ---------------------------------------------------------------------------
# uint64 helpers used as independent cross-checks of the packed restatement
# ---------------------------------------------------------------------------

def xsource_searchsorted(keys, xmask):
    """Rank of ``keys[i] ^ xmask`` in sorted unique ``keys`` or -1 (``np.searchsorted``)."""
    keys = np.asarray(keys, dtype=np.uint64)
    t = keys ^ np.uint64(xmask)
    pos = np.searchsorted(keys, t)
    if keys.shape[0] == 0:
        return pos.astype(np.int64)
    pos_c = np.minimum(pos, keys.shape[0] - 1)
    return np.where(keys[pos_c] == t, pos_c, -1).astype(np.int64)


def diagonal_u64(zmasks, coefs, keys):
    keys = np.asarray(keys, dtype=np.uint64)
    out = np.zeros(keys.shape[0], dtype=np.complex128)
    for zk, ck in zip(zmasks, coefs):
        par = np.bitwise_count(keys & np.uint64(zk)) & 1
        out += ck * (1. - 2. * par)
    return out


def matvec_u64(xmasks, groups, keys, vec):
    """``y = sum_j C_j(s) * v[rank(s ^ x_j)]`` on uint64 keys (real rows only)."""
    keys = np.asarray(keys, dtype=np.uint64)
    out = np.zeros(keys.shape[0], dtype=np.complex128)
    for xm, (zs, cs) in zip(xmasks, groups):
        src = xsource_searchsorted(keys, xm)
        diag = diagonal_u64(zs, cs, keys)
        out += apply_xgrp(src, diag, vec)
    return out
