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


def _group_values(zmask, coef, toff, group, states):
    """``C_j(s) = sum_k c_k (1 - 2 (popc(s & z_k) & 1))`` (``sqd.py:603-628``) for
    many (group, state) pairs at once: terms ``toff[j]..toff[j+1]`` of group ``j``."""
    zmask = np.asarray(zmask, dtype=np.uint64)
    coef = np.asarray(coef, dtype=np.complex128)
    toff = np.asarray(toff, dtype=np.int64)
    group = np.asarray(group, dtype=np.int64)
    states = np.asarray(states, dtype=np.uint64)
    out = np.zeros(group.shape[0], dtype=np.complex128)
    cnt = toff[group + 1] - toff[group]
    ends = np.cumsum(cnt)
    p0 = 0
    while p0 < group.shape[0]:
        # pairs [p0, p1) with at most ~4M term evaluations between them
        p1 = max(p0 + 1, int(np.searchsorted(ends, (ends[p0 - 1] if p0 else 0) + (1 << 22), side='right')))
        g, c, st = group[p0:p1], cnt[p0:p1], states[p0:p1]
        pair = np.repeat(np.arange(g.shape[0]), c)
        first = np.repeat(toff[g] - np.concatenate([[0], np.cumsum(c)[:-1]]), c)
        term = np.arange(pair.shape[0]) + first
        par = np.bitwise_count(st[pair] & zmask[term]) & 1
        np.add.at(out, pair + p0, coef[term] * (1. - 2. * par))
        p0 = p1
    return out


def hits_u64(xmask, grp_off, zmask, coef, keys, g0=0, g1=None):
    """Every nonzero of groups ``g0..g1`` of the projected Hamiltonian as triplets.

    The matrix form of ``apply_h_x_cached`` (``sqd.py:694-709``): group ``j``
    puts ``C_j(s_i)`` at ``(i, rank(s_i ^ x_j))`` for each row whose flipped
    state is in the subspace (``get_xsource``, here ``np.searchsorted``).
    Returns ``(rows int64, cols int64, vals complex128)``.
    """
    keys = np.asarray(keys, dtype=np.uint64)
    xmask = np.asarray(xmask, dtype=np.uint64)
    g1 = xmask.shape[0] if g1 is None else g1
    zmask = np.asarray(zmask, dtype=np.uint64)
    coef = np.asarray(coef, dtype=np.complex128)
    rows, cols, vals = [], [], []
    for j in range(g0, g1):
        src = xsource_searchsorted(keys, xmask[j])
        r = np.nonzero(src >= 0)[0]
        s = keys[r]
        v = np.zeros(r.shape[0], dtype=np.complex128)
        for k in range(grp_off[j], grp_off[j + 1]):
            v += coef[k] * (1. - 2. * (np.bitwise_count(s & zmask[k]) & 1))
        rows.append(r)
        cols.append(src[r])
        vals.append(v)
    rows = np.concatenate(rows) if rows else np.zeros(0, np.int64)
    cols = np.concatenate(cols) if cols else np.zeros(0, np.int64)
    vals = np.concatenate(vals) if vals else np.zeros(0, np.complex128)
    return rows.astype(np.int64), cols.astype(np.int64), vals


def csr_u64(xmask, grp_off, zmask, coef, keys, real=None):
    """Projected Hamiltonian ``H_S`` as a SciPy CSR matrix (duplicates summed).

    ``real`` (default: every coefficient real) stores float64 values.
    """
    import scipy.sparse
    n = np.asarray(keys).shape[0]
    rows, cols, vals = hits_u64(xmask, grp_off, zmask, coef, keys)
    if real is None:
        real = bool(np.all(np.asarray(coef).imag == 0)) if np.iscomplexobj(coef) else True
    if real:
        vals = vals.real
    return scipy.sparse.csr_matrix((vals, (rows, cols)), shape=(n, n))


def matvec_rows_u64(xmask, grp_off, zmask, coef, keys, vec, rows):
    """``(H_S v)_i`` for the listed rows only: every (row, group) pair searched with
    ``np.searchsorted``, ``C_j(s_i)`` summed term by term (``sqd.py:603-709``)."""
    keys = np.asarray(keys, dtype=np.uint64)
    xmask = np.asarray(xmask, dtype=np.uint64)
    rows = np.asarray(rows, dtype=np.int64)
    n = keys.shape[0]
    out = np.zeros(rows.shape[0], dtype=np.complex128)
    if n == 0 or rows.size == 0 or xmask.size == 0:
        return out
    s = keys[rows]
    tgt = s[:, None] ^ xmask[None, :]
    pos = np.searchsorted(keys, tgt.reshape(-1)).reshape(tgt.shape)
    pos_c = np.minimum(pos, n - 1)
    hit = keys[pos_c] == tgt
    ri, gj = np.nonzero(hit)
    vals = _group_values(zmask, coef, grp_off, gj, s[ri])
    np.add.at(out, ri, vals * np.asarray(vec)[pos_c[ri, gj]])
    return out
