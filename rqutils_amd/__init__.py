"""rqutils_amd -- SQD subspace Pauli-sum action and single-vector LOCG on AMD MI355X (gfx950).

Drop-in for the hot path of ``rqutils`` (``rqutils/sqd.py``,
``rqutils/paulis/symplectic.py``, ``rqutils/ground_locg.py``): host code in
Python on PyTorch-ROCm, kernels in ``librqx.so`` (HIP, gfx950) behind a C ABI
(``include/rqx.h``) called through ctypes.

Submodules mirror the reference layout: ``rqutils_amd.sqd`` (``sqd``,
``run_sqd``, ``hproj``, ``SQDOptions``), ``rqutils_amd.ground_locg``
(``ground_locg``), ``rqutils_amd.paulis.symplectic`` (``PauliSumXZ``).
"""
__version__ = '0.1.0'

from rqutils_amd.paulis.symplectic import PauliSumXZ, TermTables  # noqa: F401
