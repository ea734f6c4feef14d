"""Pauli-sum representations used by the SQD path."""
from rqutils_amd.paulis.symplectic import PauliSumXZ, TermTables

__all__ = ['PauliSumXZ', 'TermTables']
