r"""Symplectic (XZ) representation of Pauli sums, grouped by X signature.

Drop-in for ``rqutils.paulis.symplectic`` (``rqutils/paulis/symplectic.py:41-149``).

A Pauli string :math:`Q` with X signature :math:`x` and Z signature :math:`z`
is :math:`Q = (-i)^{x\cdot z} Z^z X^x` (``symplectic.py:11-19``).  Terms are
grouped by :math:`x`; each group stores its Z signatures and the phased
coefficients :math:`\alpha (-i)^{\mathrm{popcount}(x \wedge z) \bmod 4}`.

Internally the groups are kept as uint64 CSR tables (``TermTables``: bit q =
qubit q), built in O(T log T).  The reference's packed-byte arrays ``x``,
``z`` and ``c`` (zero padded to the largest group) are materialised lazily
on first access, since they are large for big Hamiltonians.
"""
from dataclasses import dataclass
from typing import Any
import warnings
import numpy as np

try:  # optional, as in the reference (symplectic.py:34-38)
    from qiskit.quantum_info import SparsePauliOp
    HAS_QISKIT = True
except ImportError:  # pragma: no cover - qiskit is not installed here
    HAS_QISKIT = False

_PHASE_TABLE = np.array([1., -1.j, -1., 1.j])


def _popcount(a):
    return np.bitwise_count(np.asarray(a, dtype=np.uint64)).astype(np.int64)


@dataclass(frozen=True, eq=False)
class TermTables:
    """X-grouped Pauli terms as uint64 CSR tables.

    ``xmask[J]`` strictly ascending; group ``j`` owns terms
    ``grp_off[j]:grp_off[j+1]`` with ``zmask`` and phased ``coef`` (complex128).
    """
    xmask: np.ndarray
    grp_off: np.ndarray
    zmask: np.ndarray
    coef: np.ndarray
    num_qubits: int

    @property
    def num_groups(self):
        return int(self.xmask.shape[0])

    @property
    def num_terms(self):
        return int(self.zmask.shape[0])

    @property
    def is_real(self):
        return bool(np.all(self.coef.imag == 0.))

    @property
    def has_diagonal(self):
        return self.num_groups > 0 and int(self.xmask[0]) == 0

    @property
    def max_group(self):
        return int(np.max(np.diff(self.grp_off))) if self.num_groups else 0

    def group_of_terms(self):
        return np.repeat(np.arange(self.num_groups), np.diff(self.grp_off))

    def term_xmask(self):
        """X mask of every term (length T)."""
        return np.repeat(self.xmask, np.diff(self.grp_off))

    def alphas(self):
        """Unphased real coefficients: ``coef * i^(popcount(x & z) & 3)``."""
        ph = _popcount(self.term_xmask() & self.zmask) & 3
        return self.coef * (1j ** ph)

    @classmethod
    def from_masks(cls, xmask, zmask, alpha, num_qubits, force_real=False, order=None, return_index=False):
        """Group raw (x, z, alpha) terms; duplicates are summed, zero sums dropped.

        Groups follow ascending x; within a group terms follow ascending z
        (the order ``np.unique`` gives Pauli strings with I < X < Y < Z).
        With ``return_index`` also returns, for every input term, its position
        in the tables (-1 if it was dropped); duplicates share a position.
        """
        if return_index:
            n_in = np.asarray(xmask).reshape(-1).shape[0]
            tagged = cls.from_masks(xmask, zmask, alpha, num_qubits, force_real=force_real, order=order)
            # locate every input (x, z) among the table terms
            xin = np.asarray(xmask, dtype=np.uint64).reshape(-1)
            zin = np.asarray(zmask, dtype=np.uint64).reshape(-1)
            tx, tz = tagged.term_xmask(), tagged.zmask
            keys_t = np.stack([tx, tz], axis=1).view([('x', np.uint64), ('z', np.uint64)]).reshape(-1)
            keys_i = np.stack([xin, zin], axis=1).view([('x', np.uint64), ('z', np.uint64)]).reshape(-1)
            pos = np.searchsorted(keys_t, keys_i) if keys_t.shape[0] else np.zeros(n_in, np.int64)
            pos_c = np.minimum(pos, max(keys_t.shape[0] - 1, 0))
            found = (keys_t.shape[0] > 0) & (keys_t[pos_c] == keys_i) if keys_t.shape[0] else np.zeros(n_in, bool)
            return tagged, np.where(found, pos_c, -1).astype(np.int64)
        xmask = np.asarray(xmask, dtype=np.uint64).reshape(-1)
        zmask = np.asarray(zmask, dtype=np.uint64).reshape(-1)
        alpha = np.asarray(alpha).reshape(-1)
        if not (xmask.shape == zmask.shape == alpha.shape):
            raise ValueError('xmask, zmask and coefficients must have the same length')
        if num_qubits < 1 or num_qubits > 64:
            raise ValueError('num_qubits must be in [1, 64]')
        if force_real:
            if np.any(np.imag(alpha) != 0.):
                warnings.warn('Found nonzero imaginary part when force_real=True')
            alpha = np.real(alpha).astype(np.float64)
        keep = alpha != 0
        xmask, zmask, alpha = xmask[keep], zmask[keep], alpha[keep]
        if order is None:
            order = np.lexsort((zmask, xmask))
        xmask, zmask, alpha = xmask[order], zmask[order], alpha[order]
        if xmask.shape[0]:
            first = np.concatenate([[True], (xmask[1:] != xmask[:-1]) | (zmask[1:] != zmask[:-1])])
            starts = np.nonzero(first)[0]
            summed = np.add.reduceat(alpha.astype(np.complex128), starts)
            xmask, zmask = xmask[starts], zmask[starts]
            nz = summed != 0
            xmask, zmask, summed = xmask[nz], zmask[nz], summed[nz]
        else:
            summed = np.zeros(0, dtype=np.complex128)
        phases = _PHASE_TABLE[_popcount(xmask & zmask) & 3]
        coef = (summed * phases).astype(np.complex128)
        if xmask.shape[0]:
            gstart = np.nonzero(np.concatenate([[True], xmask[1:] != xmask[:-1]]))[0]
            groups_x = xmask[gstart]
            grp_off = np.concatenate([gstart, [xmask.shape[0]]]).astype(np.int64)
        else:
            groups_x = np.zeros(0, dtype=np.uint64)
            grp_off = np.zeros(1, dtype=np.int64)
        return cls(groups_x.astype(np.uint64), grp_off, zmask.astype(np.uint64), coef, int(num_qubits))


def pauli_strings_to_masks(paulis):
    """Pauli strings (char c = qubit n-1-c) -> (xmask, zmask, n) as uint64 arrays."""
    paulis = list(paulis)
    if len(set(len(p) for p in paulis)) > 1:
        raise ValueError('Pauli strings have non-uniform lengths')
    if not paulis:
        raise ValueError('empty Pauli list')
    n = len(paulis[0])
    if n < 1 or n > 64:
        raise ValueError('Pauli strings must have 1 to 64 characters')
    arr = np.frombuffer(''.join(paulis).upper().encode('ascii'), dtype=np.uint8).reshape(len(paulis), n)
    valid = np.isin(arr, np.frombuffer(b'IXYZ', dtype=np.uint8))
    if not np.all(valid):
        raise ValueError('Pauli strings may only contain I, X, Y, Z')
    xb = (arr == ord('X')) | (arr == ord('Y'))
    zb = (arr == ord('Y')) | (arr == ord('Z'))
    weights = np.uint64(1) << np.arange(n - 1, -1, -1, dtype=np.uint64)
    xm = np.bitwise_or.reduce(np.where(xb, weights, np.uint64(0)), axis=1)
    zm = np.bitwise_or.reduce(np.where(zb, weights, np.uint64(0)), axis=1)
    return xm.astype(np.uint64), zm.astype(np.uint64), n


def masks_to_packed(masks, num_qubits, padded):
    """uint64 masks -> big-endian packed bytes ``[pad?][q_{n-1} .. q_0][0 fill]``."""
    masks = np.asarray(masks, dtype=np.uint64)
    width = num_qubits + (1 if padded else 0)
    nbytes = (width + 7) // 8
    shape = masks.shape
    flat = masks.reshape(-1)
    if nbytes <= 8:
        # one big-endian word per row: the row's bits start at the word's bit 8 nbytes - 1
        qmask = np.uint64((1 << num_qubits) - 1)
        words = ((flat & qmask) << np.uint64(8 * nbytes - width)).astype('>u8')
        packed = words.view(np.uint8).reshape(-1, 8)[:, 8 - nbytes:]
        return np.ascontiguousarray(packed).reshape(shape + (nbytes,))
    q = np.arange(num_qubits - 1, -1, -1, dtype=np.uint64)
    bits = ((flat[:, None] >> q[None, :]) & np.uint64(1)).astype(np.uint8)
    if padded:
        bits = np.pad(bits, ((0, 0), (1, 0)))
    packed = np.packbits(bits, axis=1)
    if packed.shape[1] != nbytes:  # pragma: no cover - packbits always gives ceil(width/8)
        raise AssertionError
    return packed.reshape(shape + (nbytes,))


def packed_to_masks(packed, num_qubits, padded):
    """Inverse of ``masks_to_packed`` (pad bits are discarded)."""
    packed = np.asarray(packed, dtype=np.uint8)
    shape = packed.shape[:-1]
    start = 1 if padded else 0
    nbytes = packed.shape[-1]
    if nbytes <= 8 and 8 * nbytes >= start + num_qubits:
        # one big-endian word per row; the key is the num_qubits bits after the pad bit
        rows = packed.reshape(-1, nbytes)
        buf = np.zeros((rows.shape[0], 8), dtype=np.uint8)
        buf[:, 8 - nbytes:] = rows
        words = buf.view('>u8').reshape(-1).astype(np.uint64)
        shift = np.uint64(8 * nbytes - start - num_qubits)
        qmask = np.uint64((1 << num_qubits) - 1)
        return ((words >> shift) & qmask).reshape(shape)
    bits = np.unpackbits(packed.reshape(-1, packed.shape[-1]), axis=1)
    if bits.shape[1] < start + num_qubits:
        raise ValueError('packed rows are too short for num_qubits')
    bits = bits[:, start:start + num_qubits].astype(np.uint64)
    weights = np.uint64(1) << np.arange(num_qubits - 1, -1, -1, dtype=np.uint64)
    out = np.bitwise_or.reduce(np.where(bits != 0, weights, np.uint64(0)), axis=1) if num_qubits else \
        np.zeros(bits.shape[0], dtype=np.uint64)
    return out.astype(np.uint64).reshape(shape)


class PauliSumXZ:
    """Symplectic (XZ) representation of a sum of Pauli strings (``symplectic.py:41-48``).

    ``PauliSumXZ(x, z, c, num_qubits)`` takes the reference's packed arrays;
    ``from_paulisum`` builds it from Pauli strings.  ``padded`` records whether
    the packed arrays carry the SQD padding bit (``add_padding=True``); when a
    caller constructs the object directly it is inferred from the byte width
    where that is unambiguous and otherwise assumed to follow the SQD
    convention (padded), which is what ``sqd``/``run_sqd``/``hproj`` require.
    """

    def __init__(self, x, z, c, num_qubits, *, padded=None, tables=None):
        self.num_qubits = int(num_qubits)
        if tables is not None:
            self._tables = tables
            self._x = x
            self._z = z
            self._c = c
            self.padded = bool(padded)
            return
        x = np.asarray(x, dtype=np.uint8)
        z = np.asarray(z, dtype=np.uint8)
        c = np.asarray(c)
        if padded is None:
            w_pad = (self.num_qubits + 8) // 8
            w_raw = (self.num_qubits + 7) // 8
            padded = not (x.shape[-1] == w_raw and w_raw != w_pad)
        self.padded = bool(padded)
        xm = packed_to_masks(x, self.num_qubits, self.padded)
        zm = packed_to_masks(z, self.num_qubits, self.padded)
        J = x.shape[0]
        xs, zs, cs = [], [], []
        for j in range(J):
            k = 0
            while k < c.shape[1] and c[j, k] != 0:   # get_diagonal stops at the first zero (sqd.py:611-613)
                k += 1
            xs.append(np.full(k, xm[j], dtype=np.uint64))
            zs.append(zm[j, :k])
            cs.append(np.asarray(c[j, :k], dtype=np.complex128))
        xall = np.concatenate(xs) if xs else np.zeros(0, np.uint64)
        zall = np.concatenate(zs) if zs else np.zeros(0, np.uint64)
        call = np.concatenate(cs) if cs else np.zeros(0, np.complex128)
        # keep the caller's order of groups and terms; groups must have distinct x
        order = np.argsort(xall, kind='stable')
        if J and np.unique(xm).shape[0] != J:
            raise ValueError('PauliSumXZ groups must have distinct X signatures')
        xall, zall, call = xall[order], zall[order], call[order]
        gstart = np.nonzero(np.concatenate([[True], xall[1:] != xall[:-1]]))[0] if xall.size else np.zeros(0, int)
        grp_off = np.concatenate([gstart, [xall.shape[0]]]).astype(np.int64)
        self._tables = TermTables(xall[gstart] if xall.size else np.zeros(0, np.uint64), grp_off,
                                  zall, call, self.num_qubits)
        self._x, self._z, self._c = x, z, c

    # -- constructors ----------------------------------------------------------------
    @classmethod
    def from_paulisum(cls, paulisum: Any, force_real: bool = False, add_padding: bool = False) -> 'PauliSumXZ':
        """``symplectic.py:50-124``: (paulis, coeffs) tuple or SparsePauliOp -> PauliSumXZ."""
        if isinstance(paulisum, tuple):
            paulis, coeffs = paulisum
            if len(paulis) != len(coeffs):
                raise ValueError('Lengths of Pauli and coeff lists do not match')
            if len(set(len(p) for p in paulis)) != 1:
                raise ValueError('Pauli strings have non-uniform lengths')
            xm, zm, n = pauli_strings_to_masks(paulis)
            alpha = np.array(coeffs)
            if alpha.dtype.kind not in 'fc':
                alpha = alpha.astype(np.float64)
        elif HAS_QISKIT and isinstance(paulisum, SparsePauliOp):  # pragma: no cover
            if not np.allclose(paulisum.coeffs.imag, 0.):
                raise ValueError('Coefficients of Paulis must be real for the Hamiltonian to be'
                                 ' Hermitian.')
            paulisum = paulisum.simplify()
            n = paulisum.num_qubits
            weights = np.uint64(1) << np.arange(n, dtype=np.uint64)
            xm = np.bitwise_or.reduce(np.where(paulisum.paulis.x, weights, np.uint64(0)), axis=1)
            zm = np.bitwise_or.reduce(np.where(paulisum.paulis.z, weights, np.uint64(0)), axis=1)
            alpha = np.asarray(paulisum.coeffs)
        else:
            raise ValueError('Unsupported input type')
        tables = TermTables.from_masks(xm, zm, alpha, n, force_real=force_real)
        return cls.from_tables(tables, add_padding=add_padding)

    @classmethod
    def from_tables(cls, tables: TermTables, add_padding: bool = False) -> 'PauliSumXZ':
        return cls(None, None, None, tables.num_qubits, padded=add_padding, tables=tables)

    # -- reference fields ------------------------------------------------------------
    @property
    def tables(self) -> TermTables:
        return self._tables

    @property
    def x(self):
        if self._x is None:
            object.__setattr__(self, '_x', masks_to_packed(self._tables.xmask, self.num_qubits, self.padded))
        return self._x

    @property
    def z(self):
        if self._z is None:
            t = self._tables
            J, K = t.num_groups, max(t.max_group, 1) if t.num_groups else 0
            zpad = np.zeros((J, K), dtype=np.uint64)
            if t.num_terms:
                g = t.group_of_terms()
                pos = np.arange(t.num_terms) - t.grp_off[g]
                zpad[g, pos] = t.zmask
            self._z = masks_to_packed(zpad, self.num_qubits, self.padded)
        return self._z

    @property
    def c(self):
        if self._c is None:
            t = self._tables
            J, K = t.num_groups, max(t.max_group, 1) if t.num_groups else 0
            cpad = np.zeros((J, K), dtype=np.complex128)
            if t.num_terms:
                g = t.group_of_terms()
                pos = np.arange(t.num_terms) - t.grp_off[g]
                cpad[g, pos] = t.coef
            self._c = cpad.real.copy() if np.all(cpad.imag == 0.) else cpad
        return self._c

    @property
    def arrays(self):
        """``symplectic.py:126-128``."""
        return self.x, self.z, self.c

    def matmul(self, rhs):
        """Full-space action ``H @ rhs`` on ``2^n`` rows (``symplectic.py:130-149``).

        ``rhs`` has shape ``(2^n,)`` or ``(2^n, ...)``; every column is applied
        in one launch of the direct gfx950 kernel (``rqx_full_apply``: the gather
        ``rhs[b ^ x_j]`` times the group's signed coefficient sum at ``b``, no
        subspace index).  NumPy input gives a NumPy result, a device tensor a
        device tensor.
        """
        from rqutils_amd.plan import Plan
        import torch
        dim = 1 << self.num_qubits
        if rhs.shape[0] != dim:
            raise ValueError(f'RHS axis 0 size {rhs.shape[0]} is incompatible with'
                             f' num_qubits={self.num_qubits}')
        as_numpy = not isinstance(rhs, torch.Tensor)
        t = torch.as_tensor(np.asarray(rhs) if as_numpy else rhs)
        device = t.device if t.is_cuda else torch.device('cuda', torch.cuda.current_device())
        plan = getattr(self, '_terms_plan', None)
        if plan is None or plan.device != device:
            plan = Plan(self._tables, np.zeros(0, dtype=np.uint64), device=device)
            self._terms_plan = plan
        dtype = torch.complex128 if not self._tables.is_real or t.is_complex() else torch.float64
        out = plan.full_apply(t.to(device=device, dtype=dtype))
        return out.cpu().numpy() if as_numpy else out
