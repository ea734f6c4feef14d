"""Oracle: textbook per-qubit action of Pauli strings (test infrastructure only).

An independent pin for everything symplectic: it never forms X/Z masks or
the ``(-i)^{x.z}`` phase.  Each qubit is acted on with the 2x2 rules

    I|b> = |b>,   X|b> = |1-b>,   Y|b> = i(-1)^b |1-b>,   Z|b> = (-1)^b |b>

and string character ``c`` acts on qubit ``n-1-c`` (the Qiskit convention used
by ``rqutils/sqd.py:132-135`` of the examples and ``tests/sqd.ipynb``).
"""
import numpy as np


def apply_pauli_string(pauli, basis):
    """Return (phase, image) with ``P|basis> = phase |image>`` for uint64 basis states."""
    basis = np.asarray(basis, dtype=np.uint64)
    n = len(pauli)
    phase = np.ones(basis.shape, dtype=np.complex128)
    image = basis.copy()
    for c, ch in enumerate(pauli.upper()):
        q = np.uint64(n - 1 - c)
        bit = ((basis >> q) & np.uint64(1)).astype(np.int64)
        if ch == 'I':
            continue
        if ch == 'X':
            image ^= (np.uint64(1) << q)
        elif ch == 'Y':
            image ^= (np.uint64(1) << q)
            phase *= 1j * (1 - 2 * bit)
        elif ch == 'Z':
            phase *= (1 - 2 * bit)
        else:
            raise ValueError(f'bad Pauli character {ch!r}')
    return phase, image


def projected_matrix(paulis, coeffs, states):
    """Dense ``<s_a|H|s_b>`` for sorted unique uint64 ``states`` (small N only)."""
    states = np.asarray(states, dtype=np.uint64)
    n_states = states.shape[0]
    out = np.zeros((n_states, n_states), dtype=np.complex128)
    for pauli, alpha in zip(paulis, coeffs):
        phase, image = apply_pauli_string(pauli, states)
        pos = np.searchsorted(states, image)
        pos_c = np.minimum(pos, n_states - 1)
        found = states[pos_c] == image
        cols = np.nonzero(found)[0]
        # H|s_b> has amplitude alpha*phase on |image_b>, i.e. row pos[b], column b
        np.add.at(out, (pos_c[cols], cols), alpha * phase[cols])
    return out


def full_matrix(paulis, coeffs, num_qubits):
    """Dense ``2^n x 2^n`` matrix of the Pauli sum (n <= 12)."""
    basis = np.arange(1 << num_qubits, dtype=np.uint64)
    out = np.zeros((1 << num_qubits,) * 2, dtype=np.complex128)
    for pauli, alpha in zip(paulis, coeffs):
        phase, image = apply_pauli_string(pauli, basis)
        out[image.astype(np.int64), basis.astype(np.int64)] += alpha * phase
    return out


def projected_apply(paulis, coeffs, states, vec):
    """``H_S @ vec`` without forming the matrix (moderate N)."""
    states = np.asarray(states, dtype=np.uint64)
    n_states = states.shape[0]
    out = np.zeros(n_states, dtype=np.complex128)
    for pauli, alpha in zip(paulis, coeffs):
        phase, image = apply_pauli_string(pauli, states)
        pos = np.searchsorted(states, image)
        pos_c = np.minimum(pos, n_states - 1)
        found = states[pos_c] == image
        cols = np.nonzero(found)[0]
        np.add.at(out, pos_c[cols], alpha * phase[cols] * vec[cols])
    return out
