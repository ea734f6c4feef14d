"""Oracle: Pauli strings -> X-grouped symplectic tables (test infrastructure only).

Restates the tuple-input path of ``PauliSumXZ.from_paulisum``
(``rqutils/paulis/symplectic.py:57-124``) with NumPy, producing the same
packed-byte layout:

* ``x``: ``(J, B)`` uint8, one packed X signature per group, groups in the
  order of ``np.unique(xbits, axis=0)`` (``:98``);
* ``z``: ``(J, K_max, B)`` uint8, zero padded;
* ``c``: ``(J, K_max)`` phased coefficients ``alpha * (-i)^(popc(x & z) & 3)``
  (``:109-111``), real dtype iff every phased coefficient is real (``:113-114``).

String character ``c`` is qubit ``n-1-c`` (Qiskit order).  With
``add_padding`` a zero bit is prepended to every signature (``:116-119``).
"""
import numpy as np

_PHASES = np.array([1., -1.j, -1., 1.j])


def from_paulisum_tuple(paulis, coeffs, force_real=False, add_padding=False):
    """Restatement of ``symplectic.py:57-124`` for the ``(paulis, coeffs)`` input."""
    if len(paulis) != len(coeffs):
        raise ValueError('Lengths of Pauli and coeff lists do not match')
    if len(set(len(p) for p in paulis)) != 1:
        raise ValueError('Pauli strings have non-uniform lengths')
    coeffs = np.array(coeffs)
    paulis = np.array([list(p.upper()) for p in paulis])
    nonzero = np.nonzero(coeffs)
    coeffs = coeffs[nonzero]
    paulis = paulis[nonzero]
    paulis, indices = np.unique(paulis, axis=0, return_inverse=True)
    indices = indices.reshape(-1)
    # symplectic.py:71-72 builds a dense (T_unique, T_in) mask matrix; a
    # scatter-add gives the same sums.
    summed = np.zeros(paulis.shape[0], dtype=np.result_type(coeffs.dtype, np.float64))
    np.add.at(summed, indices, coeffs)
    coeffs = summed
    xbits = np.logical_or(paulis == 'X', paulis == 'Y')
    zbits = np.logical_or(paulis == 'Y', paulis == 'Z')
    num_qubits = paulis.shape[1]

    if force_real:
        coeffs = np.real(coeffs)

    xuniq, indices, counts = np.unique(xbits, axis=0, return_inverse=True, return_counts=True)
    indices = indices.reshape(-1)
    xsignatures = xuniq.astype(np.uint8)
    shape = (xsignatures.shape[0], np.max(counts))
    zsignatures = np.zeros(shape + zbits.shape[-1:], dtype=np.uint8)
    phcoeffs = np.zeros(shape, dtype=np.complex128)
    for isig, xsig in enumerate(xsignatures):
        ipaulis = np.nonzero(indices == isig)[0]
        zsigs = zbits[ipaulis].astype(np.uint8)
        zsignatures[isig, :counts[isig]] = zsigs
        iphases = np.sum(xsig & zsigs, axis=1) & 3
        phcoeffs[isig, :counts[isig]] = coeffs[ipaulis] * _PHASES[iphases]

    if np.all(phcoeffs.imag == 0.):
        phcoeffs = phcoeffs.real

    if add_padding:
        xsignatures = np.pad(xsignatures, ((0, 0), (1, 0)))
        zsignatures = np.pad(zsignatures, ((0, 0), (0, 0), (1, 0)))

    return (np.packbits(xsignatures, axis=-1), np.packbits(zsignatures, axis=-1),
            phcoeffs, num_qubits)


def packed_to_int(packed, num_qubits, padded):
    """Big-endian packed rows -> Python-int-free uint64 values of the n qubit bits.

    Bit ``q`` of the result is qubit ``q``.  Layout (``sqd.py:251-255``):
    ``[pad?][q_{n-1} ... q_0][zero fill]``.
    """
    packed = np.asarray(packed, dtype=np.uint8)
    nbytes = packed.shape[-1]
    bits = np.unpackbits(packed, axis=-1)
    start = 1 if padded else 0
    bits = bits[..., start:start + num_qubits]
    weights = (np.uint64(1) << np.arange(num_qubits - 1, -1, -1, dtype=np.uint64))
    out = np.zeros(bits.shape[:-1], dtype=np.uint64)
    for q in range(num_qubits):
        out |= bits[..., q].astype(np.uint64) * weights[q]
    del nbytes
    return out


def tables_as_u64(x, z, c, num_qubits, padded):
    """Oracle tables -> (xmask[J], list of (zmask[K_j], coef[K_j])) with trailing zero coefs cut.

    ``get_diagonal`` stops at the first zero coefficient of a group
    (``sqd.py:611-613``); the same cut is applied here.
    """
    xm = packed_to_int(x, num_qubits, padded)
    zm = packed_to_int(z, num_qubits, padded)
    groups = []
    for j in range(xm.shape[0]):
        k = 0
        while k < c.shape[1] and c[j, k] != 0:
            k += 1
        groups.append((zm[j, :k].copy(), np.array(c[j, :k])))
    return xm, groups
