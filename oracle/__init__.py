"""CPU oracle for the SQD Pauli-sum matvec and the single-vector LOCG solver.

TEST INFRASTRUCTURE ONLY.  This package is the checker, never the thing
measured or shipped: only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import it.  The product package
``rqutils_amd`` never imports anything from here, and it fails loudly when
its HIP library is missing instead of falling back to this code.

Modules
-------
``terms``    restatement of ``rqutils/paulis/symplectic.py:50-124`` (Pauli
             strings -> X-grouped packed x/z signatures, phased coefficients).
``sqdref``   restatement of ``rqutils/sqd.py`` (state packing ``:251-255``,
             ``uniquify_states`` ``:445-482``, ``get_xsource`` ``:485-548``,
             ``get_diagonal`` ``:603-628``, ``apply_xgrp`` ``:631-640``,
             ``apply_h_x_cached`` ``:694-709``, vinit ``:406-426``).
``locgref``  restatement of ``rqutils/ground_locg.py:128-514`` in NumPy FP64.
``dense``    an independent pin: textbook per-qubit action of Pauli strings on
             computational basis states (no symplectic algebra), used to build
             projected matrices for small inputs.
``cport``    ctypes binding of ``oracle/rqx_oracle.c``, a C/OpenMP port of the
             reference's default cache strategy (1, 0), timed as the CPU
             baseline.

Pinning (how each function is checked, see ``tests/test_oracle_*.py``):

* ``terms``: dense Pauli matrices from ``dense`` (Kronecker-free per-qubit
  action) must equal the sum of ``c * Z^z X^x`` built from the oracle tables.
* ``sqdref.get_xsource``: bit-exact against ``numpy.searchsorted`` on the
  sorted uint64 keys, and against brute-force dict lookup.
* ``sqdref`` matvec: against the dense projected matrix from ``dense``.
* ``locgref.eigenpair_2x2/3x3``: against ``numpy.linalg.eigh`` including
  diagonal inputs.
* ``locgref.ground_locg``: against ``scipy.sparse.linalg.eigsh`` / ``eigh``
  and the closed-form ground energy of the periodic transverse-field Ising
  chain.
"""
