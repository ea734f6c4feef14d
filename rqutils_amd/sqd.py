r"""Sample-based quantum diagonalization of Pauli-sum Hamiltonians on one MI355X.

Drop-in for ``rqutils.sqd`` (``rqutils/sqd.py:183-442``): ``SQDOptions``,
``sqd``, ``hproj`` and ``run_sqd`` keep their signatures and results.

Algorithm (``sqd.py:33-86``): terms are grouped by X signature, so for the
sorted subspace :math:`S` and a vector :math:`v`

.. math::  (H v)_i = \sum_j C_j(s_i)\, v[\mathrm{rank}(s_i \oplus x_j)],\qquad
           C_j(s) = \sum_k \alpha_{jk} (-i)^{x_j z_{jk}} (-1)^{z_{jk} \cdot s}

with the term 0 when :math:`s_i \oplus x_j \notin S`.  States are uint64
bitstrings (bit q = qubit q); the gfx950 kernel (``librqx``) searches the
flipped states in a two-level sorted index instead of building the reference's
``J x N`` source-index cache.  ``cache_strategy`` (1, *) (the reference default)
additionally stores the hits once, source row and coefficient sum, in a
compacted device cache when it fits (``Plan.build_cache``); every strategy
gives the same numbers.  Filler slots (``states_size`` beyond the unique
count) are the reference's 0xFF rows; they stay decoupled and hold 0.
"""
from collections.abc import Sequence
from dataclasses import dataclass
import logging
import time
from numbers import Number
from typing import Any, Optional

import numpy as np
import torch
from scipy.sparse import csr_array

from rqutils_amd.paulis.symplectic import PauliSumXZ, masks_to_packed, packed_to_masks
from rqutils_amd._capi import RqxError
from rqutils_amd.plan import Plan, pack_to_u64, sort_unique_u64
from rqutils_amd.ground_locg import ground_locg
from rqutils_amd.locg_batch import batch_ground_locg

LOG = logging.getLogger(__name__)

HamiltonianInput = PauliSumXZ | tuple[Sequence[str], Sequence[Number]]


@dataclass(frozen=True)
class SQDOptions:
    """``sqd.py:183-196``."""
    return_eigvec: bool = True
    """Whether to return the eigenvector (coefficients and unique state bitstrings)."""
    cache_strategy: tuple[int, int] = (1, 0)
    """(1, *): cache the hits (source rows and coefficient sums) on the device when they fit
    in its free memory; (0, *): matrix-free.  The results are identical."""
    shard_bundle: int = 1
    """Accepted for compatibility (single-GPU plan)."""
    locg_maxiter: int = 1000
    """LOCG maximum iterations."""
    locg_tol: float | None = None
    """LOCG convergence tolerance."""
    locg_precision: str = 'fp64'
    """'fp64': the reference LOCG (``ground_locg``, FP64 throughout).  'mixed' (an
    extension): the graph-captured LOCG step with an FP32 bulk matvec and FP64
    vectors / Rayleigh-Ritz, stopped at relative residual ``locg_tol`` (default
    1e-8); with ``cache_strategy`` (1, *) both cache precisions are built."""

    def __post_init__(self):
        cs = tuple(self.cache_strategy)
        if len(cs) != 2 or cs[0] not in (0, 1) or cs[1] not in (0, 1, 2):
            raise ValueError('cache_strategy must be (0|1, 0|1|2)')
        if self.locg_precision not in ('fp64', 'mixed'):
            raise ValueError("locg_precision must be 'fp64' or 'mixed'")
        if self.shard_bundle < 0:
            raise ValueError('bundle must be positive and divide the mesh size evenly')


def _as_options(options):
    if options is None:
        return SQDOptions()
    if isinstance(options, dict):
        return SQDOptions(**options)
    return options


def states_to_u64(states) -> np.ndarray:
    """Binary ``(N, n)`` rows in string order (column c = qubit n-1-c) -> uint64."""
    states = np.asarray(states)
    if states.ndim != 2:
        raise ValueError('states must be a 2-d (num_states, num_qubits) array')
    n = states.shape[1]
    if n < 1 or n > 64:
        raise ValueError('states must have 1 to 64 columns')
    return packed_to_masks(np.packbits(states.astype(bool).astype(np.uint8), axis=1), n, padded=False)


def u64_to_states(keys, num_qubits) -> np.ndarray:
    """Inverse of ``states_to_u64``: ``(N, n)`` uint8 rows."""
    keys = np.asarray(keys, dtype=np.uint64)
    q = np.arange(num_qubits - 1, -1, -1, dtype=np.uint64)
    return ((keys[:, None] >> q[None, :]) & np.uint64(1)).astype(np.uint8)


def pack_states_u64(keys, num_qubits, states_size=None) -> np.ndarray:
    """uint64 keys -> the reference's packed rows with pad bit, 0xFF filler rows (``sqd.py:255,473-478``)."""
    packed = masks_to_packed(np.asarray(keys, dtype=np.uint64), num_qubits, padded=True)
    if states_size is not None and states_size > packed.shape[0]:
        fill = np.full((states_size - packed.shape[0], packed.shape[1]), 255, dtype=np.uint8)
        packed = np.concatenate([packed, fill], axis=0)
    return packed


def _hamiltonian(hamiltonian, force_real):
    if isinstance(hamiltonian, PauliSumXZ):
        return hamiltonian
    return PauliSumXZ.from_paulisum(hamiltonian, force_real=force_real, add_padding=True)


def sqd(
    hamiltonian: HamiltonianInput,
    states,
    states_size: Optional[int] = None,
    options: Optional[SQDOptions | dict[str, Any]] = None
):
    r"""Sample-based quantum diagonalization (``sqd.py:199-266``).

    Args:
        hamiltonian: ``(paulis, coeffs)`` tuple, Qiskit ``SparsePauliOp`` or ``PauliSumXZ``.
        states: Binary ``(num_states, num_qubits)`` array (column c = qubit n-1-c).
        states_size: Fixed vector length (>= number of rows); extra slots are fillers.
        options: ``SQDOptions`` or a dict of its fields.

    Returns:
        The ground energy, or ``(energy, eigvec, basis_states)`` when
        ``options.return_eigvec``; ``eigvec`` is a NumPy vector over the sorted
        unique states ``basis_states`` (``(dim, num_qubits)`` uint8).
    """
    options = _as_options(options)
    states = np.asarray(states)
    if states_size is None:
        states_size = states.shape[0]
    if states_size < states.shape[0]:
        raise ValueError('states_size smaller than the states array length')
    hamiltonian = _hamiltonian(hamiltonian, force_real=True)
    if states.shape[1] != hamiltonian.num_qubits:
        raise ValueError('states and Hamiltonian have different numbers of qubits')
    keys = states_to_u64(states)
    LOG.debug('Starting SQD with array size %s', states_size)
    start = time.time()
    result = _run(hamiltonian, keys, int(states_size), options)
    LOG.info('Found ground eigenpair in %f seconds.', time.time() - start)
    eigval = float(result[0])
    if options.return_eigvec:
        eigvec, keys_u, subspace_dim = result[1], result[2], result[3]
        basis_states = u64_to_states(keys_u[:subspace_dim], states.shape[1])
        return eigval, eigvec[:subspace_dim].cpu().numpy(), basis_states
    return eigval


def run_sqd(
    hamiltonian: PauliSumXZ,
    states_p,
    states_size: int,
    options: SQDOptions | None = None,
    log_level: int = logging.INFO
):
    """``sqd.py:317-442``: packed states (pad bit first) -> ``(E,)`` or
    ``(E, eigvec[states_size], states_u[states_size, B], subspace_dim)``.

    ``eigvec`` is a device tensor; ``states_u`` the packed unique rows with
    0xFF filler rows, as in the reference.
    """
    options = _as_options(options)
    hamiltonian = _hamiltonian(hamiltonian, force_real=True)
    if states_size < len(states_p):
        raise ValueError('states_size smaller than the states array length')
    keys = packed_keys_device(states_p, hamiltonian.num_qubits)
    result = _run(hamiltonian, keys, int(states_size), options, log_level=log_level)
    if options.return_eigvec:
        eigval, eigvec, keys_u, dim = result
        return eigval, eigvec, pack_states_u64(keys_u, hamiltonian.num_qubits, states_size), dim
    return result


def packed_keys_device(states_p, num_qubits, device=None) -> torch.Tensor:
    """Packed rows (pad bit first, ``sqd.py:255``; numpy array or device tensor)
    -> int64 device tensor of the real rows' uint64 states, unpacked on the GPU."""
    dev = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
    if isinstance(states_p, torch.Tensor):
        rows = states_p.to(dev)
    else:
        rows = torch.from_numpy(np.ascontiguousarray(np.asarray(states_p, dtype=np.uint8))).to(dev)
    if rows.dim() != 2 or rows.shape[1] != (num_qubits + 8) // 8:
        raise ValueError('packed rows must have ceil((num_qubits + 1) / 8) bytes')
    return pack_to_u64(rows, num_qubits)


def uniquify_states(keys, device=None) -> np.ndarray:
    """Sorted unique uint64 states computed on the GPU (``sqd.py:445-482``)."""
    keys = np.ascontiguousarray(np.asarray(keys, dtype=np.uint64))
    if keys.shape[0] == 0:
        return keys
    dev = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
    t = torch.from_numpy(keys.view(np.int64)).to(dev)
    u = sort_unique_u64(t)
    return u.cpu().numpy().view(np.uint64)


def make_plan(hamiltonian: PauliSumXZ, keys, states_size=None, device=None, split_bits=-1) -> Plan:
    """Uniquify ``keys`` (uint64 numpy array or 64-bit device tensor) on the GPU
    and build the device plan from the device keys (no host round trip)."""
    if isinstance(keys, torch.Tensor):
        t = keys.reshape(-1)
        if device is not None:
            t = t.to(device)
    else:
        dev = torch.device('cuda', torch.cuda.current_device()) if device is None else torch.device(device)
        t = torch.from_numpy(np.ascontiguousarray(np.asarray(keys, dtype=np.uint64)).view(np.int64)).to(dev)
    keys_u = sort_unique_u64(t) if t.numel() else t
    if states_size is not None and keys_u.shape[0] > states_size:
        raise ValueError('states_size smaller than the number of unique states')
    return Plan(hamiltonian.tables, keys_u, n_total=states_size, device=keys_u.device, split_bits=split_bits)


def _run(hamiltonian, keys, states_size, options, log_level=logging.INFO):
    plan = make_plan(hamiltonian, keys, states_size)
    tables = hamiltonian.tables
    mixed = options.locg_precision == 'mixed'
    if options.cache_strategy[0] == 1 and plan.n_states > 0:
        # sqd.py:343-351 caches the source index of every (group, row) pair; the
        # hit cache holds the hits with their coefficient sums (strategy (1, 2))
        try:
            info = plan.build_cache(0 if mixed else 64)
            LOG.debug('Hit cache: %d hits, %d bytes', info['nnz'], info['bytes'])
        except RqxError as exc:
            LOG.info('Hit cache does not fit (%s); using the matrix-free kernel', exc)
    dtype = torch.float64 if tables.is_real else torch.complex128
    # vinit (sqd.py:406-426): argmin of the x = 0 group's diagonal over real rows
    if tables.has_diagonal and plan.n_states > 0:
        diag = plan.diagonal().real[:plan.n_states]
        vinit = torch.argmin(diag)
    else:
        vinit = 0
    if mixed and plan.n_states > 0:
        x0 = torch.zeros(states_size, dtype=dtype, device=plan.device)
        x0[vinit] = 1
        tol = 1e-8 if options.locg_tol is None else options.locg_tol
        theta, eigvec, niter, info = batch_ground_locg(plan, x0, tol=tol, maxiter=options.locg_maxiter,
                                                       check_every=5)
        LOG.log(log_level, 'Mixed LOCG: %d iterations, relative residual %g', niter, float(info['rel_residual'][0]))
        if not bool(info['converged'][0]):
            LOG.warning('Mixed LOCG stopped at maxiter=%d before reaching tol=%g', options.locg_maxiter, tol)
        eigval = theta[0]
    else:
        eigval, eigvec, _ = ground_locg(plan.matvec, vinit, maxiter=options.locg_maxiter, tol=options.locg_tol,
                                        vspace=(states_size, dtype, plan.device), log_level=log_level)
    if options.return_eigvec:
        return eigval, eigvec, plan.states, plan.n_states
    return (eigval,)


def hproj(hamiltonian: HamiltonianInput, states, unique_states: bool = False) -> csr_array:
    """Projected Hamiltonian as a sparse CSR matrix (``sqd.py:269-314``).

    Rows/columns follow the sorted unique states.  (The reference pads the
    Hamiltonian but not the states here, ``sqd.py:295,298``; both are aligned.)
    """
    hamiltonian = _hamiltonian(hamiltonian, force_real=False)
    keys = states_to_u64(np.asarray(states))
    if not unique_states:
        keys = np.unique(keys)
    plan = Plan(hamiltonian.tables, keys)
    n = plan.n_states
    # one triplet per nonzero (the reference scans J x N columns and elements, sqd.py:300-314)
    rows, cols, data = plan.hits()
    if hamiltonian.tables.is_real:
        data = data.real
    indptr = torch.zeros(n + 1, dtype=torch.int64, device=rows.device)
    indptr[1:] = torch.cumsum(torch.bincount(rows, minlength=n), 0)
    return csr_array((data.cpu().numpy(), cols.cpu().numpy(), indptr.cpu().numpy()), shape=(n, n))
