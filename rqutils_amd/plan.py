"""Device-resident SQD plan: term tables + subspace index behind ``librqx``.

A ``Plan`` replaces the setup half of ``run_sqd`` (``rqutils/sqd.py:341-356``):
the uniquified states and the X-grouped term tables are uploaded once, and
``Plan.matvec`` is the matrix-free action ``y = H_S x`` that
``make_x_cached_matvec`` returns on one device (``sqd.py:746-751``).  No
``J x N`` source-index cache is built: sources are searched on the fly.
"""
import ctypes
import numpy as np
import torch

from rqutils_amd import _capi
from rqutils_amd.paulis.symplectic import TermTables

_DTYPE_CODES = {
    torch.float64: _capi.RQX_F64,
    torch.complex128: _capi.RQX_C128,
    torch.float32: _capi.RQX_F32,
    torch.complex64: _capi.RQX_C64,
}


def _require_cuda():
    if not torch.cuda.is_available():
        raise RuntimeError('rqutils_amd needs a ROCm GPU (torch.cuda.is_available() is False)')


class Plan:
    """Term tables + sorted uint64 subspace index resident on one GPU.

    Args:
        tables: X-grouped terms (``TermTables``).
        sorted_states: sorted, unique uint64 states: a numpy array or CPU tensor,
            or an int64 / uint64 tensor on the plan device (e.g. from
            ``sort_unique_u64``), whose index passes then run on the GPU
            (``rqx_plan_create_device``; the plan is the same).
        n_total: vector length (>= number of states); rows past the states are
            filler slots that always receive 0 (``sqd.py:473-478``).
        device: CUDA device (default: current).
        split_bits: segment split L (default: chosen by the library's cost model).
    """

    def __init__(self, tables: TermTables, sorted_states, n_total=None, device=None, split_bits=-1):
        _require_cuda()
        lib = _capi.lib()
        self.tables = tables
        on_device = isinstance(sorted_states, torch.Tensor) and sorted_states.is_cuda
        if on_device:
            if sorted_states.dtype not in (torch.int64, torch.uint64):
                raise ValueError('device states must be a 64-bit integer tensor')
            states = sorted_states.reshape(-1).contiguous()
            if device is None:
                device = states.device
            elif torch.device(device).index not in (None, states.device.index):
                raise ValueError('device states must be on the plan device')
        else:
            states = np.ascontiguousarray(np.asarray(sorted_states, dtype=np.uint64).reshape(-1))
        self.n_states = int(states.shape[0])
        self.n_total = int(self.n_states if n_total is None else n_total)
        if self.n_total < self.n_states:
            raise ValueError('n_total smaller than the number of states')
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        xmask = np.ascontiguousarray(tables.xmask, dtype=np.uint64)
        grp_off = np.ascontiguousarray(tables.grp_off, dtype=np.int64)
        zmask = np.ascontiguousarray(tables.zmask, dtype=np.uint64)
        coef = np.ascontiguousarray(tables.coef, dtype=np.complex128).view(np.float64)
        handle = ctypes.c_void_p()
        create = lib.rqx_plan_create_device if on_device else lib.rqx_plan_create
        states_ptr = (states.data_ptr() if self.n_states else None) if on_device else _capi.ptr(states)
        with torch.cuda.device(self.device):
            stream = _capi.current_stream(self.device)
            rc = create(ctypes.byref(handle), int(tables.num_qubits),
                        _capi.ptr(xmask), _capi.ptr(grp_off), int(tables.num_groups),
                        _capi.ptr(zmask), _capi.ptr(coef), int(tables.num_terms),
                        states_ptr, self.n_states, self.n_total, int(split_bits), stream)
        _capi.check(rc)
        self._handle = handle
        self._states = states
        self._load_info()

    def _load_info(self):
        lib = _capi.lib()
        info = _capi.PlanInfo()
        _capi.check(lib.rqx_plan_get_info(self._handle, ctypes.byref(info)))
        self.info = info.as_dict()
        self.is_real = bool(self.info['is_real'])
        self.n_problems = int(self.info['n_problems'])
        self.problems = []
        for k in range(self.info['n_problems']):
            pi = _capi.ProblemInfo()
            _capi.check(lib.rqx_plan_get_problem(self._handle, k, ctypes.byref(pi)))
            self.problems.append(pi.as_dict())

    def set_launch(self, persistent, blocks_per_cu=0):
        """Tile kernel launch shape: persistent work queue or one workgroup per tile."""
        _capi.check(_capi.lib().rqx_plan_set_launch(self._handle, int(bool(persistent)), int(blocks_per_cu)))
        self._load_info()

    @classmethod
    def full_space(cls, tables: TermTables, device=None):
        """Plan over the complete ``2^n`` computational basis as a subspace (n <= 30)."""
        n = tables.num_qubits
        if n > 30:
            raise ValueError('a full-space subspace plan needs num_qubits <= 30 (use full_apply)')
        return cls(tables, np.arange(1 << n, dtype=np.uint64), device=device)

    def full_apply(self, rhs, out=None):
        """``H @ rhs`` over the whole ``2^n`` space with no subspace index (``symplectic.py:130-149``).

        ``rhs``: device tensor of shape ``(2^n,)`` or ``(2^n, k)`` (float32/64,
        complex64/128); every column is applied in one launch.  Uses only the
        plan's term tables, so a plan built over an empty subspace serves.
        """
        if not isinstance(rhs, torch.Tensor) or rhs.device != self.device:
            raise ValueError('rhs must be a tensor on the plan device')
        n = self.info['n_qubits']
        dim = 1 << n
        if rhs.shape[0] != dim:
            raise ValueError(f'RHS axis 0 size {rhs.shape[0]} is incompatible with num_qubits={n}')
        if rhs.dtype not in _DTYPE_CODES:
            raise ValueError(f'unsupported dtype {rhs.dtype}')
        cols = rhs.reshape(dim, -1).contiguous()
        if out is None:
            out = torch.empty_like(cols)
        elif out.shape != cols.shape or out.dtype != cols.dtype or not out.is_contiguous():
            raise ValueError('out must be a contiguous tensor like rhs')
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_full_apply(self._handle, cols.data_ptr(), out.data_ptr(), cols.shape[1],
                                               _DTYPE_CODES[cols.dtype], stream))
        return out.reshape(rhs.shape)

    def digest(self):
        """Checksums of the device index arrays (``rqx_plan_digest``): a dict of
        array name -> 64-bit checksum; equal for equal indexes."""
        names = ('keys', 'lo', 'seg_hi', 'seg_off', 'cls_hi', 'cls_goff', 'tile_row0', 'tile_seg',
                 'tile_prob', 'prob', 'h_off', 'h_log', 'h_tab', 'st_off', 'st_list', 'diag64')
        out = (ctypes.c_uint64 * _capi.RQX_DIGEST_ARRAYS)()
        _capi.check(_capi.lib().rqx_plan_digest(self._handle, out, _capi.current_stream(self.device)))
        return dict(zip(names, (int(v) for v in out)))

    # ------------------------------------------------------------------ lifetime
    def close(self):
        handle = getattr(self, '_handle', None)
        if handle is not None and handle.value:
            lib = _capi.lib()
            with torch.cuda.device(self.device):
                torch.cuda.current_stream(self.device).synchronize()
                rc = lib.rqx_plan_destroy(handle)
            self._handle = None
            _capi.check(rc)

    def __del__(self):
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass

    @property
    def states(self):
        """The sorted unique states as a uint64 numpy array (copied from the device
        once for a plan built from device states)."""
        if isinstance(self._states, torch.Tensor):
            self._states = self._states.cpu().numpy().view(np.uint64)
        return self._states

    # ------------------------------------------------------------------ kernels
    def _vec(self, x):
        if not isinstance(x, torch.Tensor):
            raise TypeError('Plan.matvec expects a torch tensor on the plan device')
        if x.device != self.device:
            raise ValueError(f'vector is on {x.device}, plan on {self.device}')
        if x.dim() != 1 or x.shape[0] != self.n_total:
            raise ValueError(f'vector must have shape ({self.n_total},), got {tuple(x.shape)}')
        if x.dtype not in _DTYPE_CODES:
            raise ValueError(f'unsupported dtype {x.dtype}')
        return x.contiguous()

    def matvec(self, x, out=None):
        """``y = H_S x`` (device tensor of dtype float64/complex128/float32/complex64)."""
        x = self._vec(x)
        if out is None:
            out = torch.empty_like(x)
        elif out.dtype != x.dtype or out.shape != x.shape or out.device != x.device or not out.is_contiguous():
            raise ValueError('out must be a contiguous tensor like x')
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_matvec(self._handle, x.data_ptr(), out.data_ptr(),
                                           _DTYPE_CODES[x.dtype], stream))
        return out

    __call__ = matvec

    def lookup(self, values=False):
        """``(J, n_total)`` int32 source indices (-1 where absent), ``get_xsource`` for all groups.

        With ``values=True`` also returns the ``(J, n_total)`` complex128 matrix
        elements ``C_j(s_i)`` at the hits (0 elsewhere).
        """
        shape = (self.info['n_groups'], self.n_total)
        out = torch.empty(shape, dtype=torch.int32, device=self.device)
        vals = torch.empty(shape, dtype=torch.complex128, device=self.device) if values else None
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_lookup(self._handle, out.data_ptr() if out.numel() else None,
                                           vals.data_ptr() if (vals is not None and vals.numel()) else None,
                                           stream))
        return (out, vals) if values else out

    def diagonal(self):
        """Complex ``C_0(s)`` of the x == 0 group for every row (0 on filler rows)."""
        out = torch.empty(self.n_total, dtype=torch.complex128, device=self.device)
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_diagonal(self._handle, out.data_ptr() if out.numel() else None, stream))
        return out

    def grad_coef(self, x, gy, out=None):
        """``g_k = Re <gy, P_k x>``: gradient of ``Re <gy, H x>`` w.r.t. the unphased coefficients.

        Terms follow ``self.tables`` order; returns a float64 device tensor.
        """
        x = self._vec(x)
        gy = self._vec(gy)
        if gy.dtype != x.dtype:
            raise ValueError('x and gy must have the same dtype')
        T = self.info['n_terms']
        if out is None:
            out = torch.empty(T, dtype=torch.float64, device=self.device)
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_grad_coef(self._handle, x.data_ptr(), gy.data_ptr(),
                                              out.data_ptr() if T else None, _DTYPE_CODES[x.dtype], stream))
        return out

    def expect_terms(self, psi):
        """``Re <psi|P_k|psi> / <psi|psi>`` per term: Hellmann-Feynman ``dE/dalpha_k``."""
        psi = self._vec(psi)
        T = self.info['n_terms']
        out = torch.empty(T, dtype=torch.float64, device=self.device)
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_expect_terms(self._handle, psi.data_ptr(), out.data_ptr() if T else None,
                                                 _DTYPE_CODES[psi.dtype], stream))
        nrm2 = torch.linalg.vector_norm(psi).to(torch.float64) ** 2
        return out / nrm2

    def term_phases(self):
        """``(-i)^{popcount(x_k & z_k)}`` per term (complex128 device tensor)."""
        if getattr(self, '_phases', None) is None:
            tabs = self.tables if isinstance(self.tables, (list, tuple)) else [self.tables]
            xm = np.concatenate([t.term_xmask() for t in tabs])
            zm = np.concatenate([t.zmask for t in tabs])
            ph = np.bitwise_count(xm & zm).astype(np.int64) & 3
            table = np.array([1., -1j, -1., 1j])
            self._phases = torch.from_numpy(table[ph]).to(self.device)
        return self._phases

    def set_alphas(self, alpha):
        """Use ``alpha`` (real, one per term, tables order) as the unphased coefficients."""
        alpha = torch.as_tensor(alpha, device=self.device).to(torch.float64).reshape(-1)
        if alpha.shape[0] != self.info['n_terms']:
            raise ValueError('alpha must have one entry per term')
        coef = (alpha.to(torch.complex128) * self.term_phases()).contiguous()
        is_real = bool(torch.all(self.term_phases().imag == 0).item()) if coef.numel() else True
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_plan_update_coef(self._handle, torch.view_as_real(coef).data_ptr()
                                                     if coef.numel() else None, int(is_real), stream))
        self.is_real = is_real

    def count(self):
        """Structural nonzeros, segment-level probes and (tile, target segment) pairs of one matvec."""
        out = torch.zeros(3, dtype=torch.int64, device=self.device)
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_count(self._handle, out.data_ptr(), stream))
        hits, probes, lookups = (int(v) for v in out.cpu())
        return {'hits': hits, 'probes': probes, 'class_lookups': lookups}

    def hits(self):
        """Every nonzero as device tensors ``(rows int64, cols int64, vals complex128)``, sorted by (row, col)."""
        nnz = self.count()['hits']
        rows = torch.empty(max(nnz, 1), dtype=torch.int64, device=self.device)
        cols = torch.empty(max(nnz, 1), dtype=torch.int32, device=self.device)
        vals = torch.empty((max(nnz, 1), 2), dtype=torch.float64, device=self.device)
        count = torch.zeros(1, dtype=torch.int64, device=self.device)
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_plan_hits(self._handle, rows.data_ptr(), cols.data_ptr(), vals.data_ptr(),
                                              nnz, count.data_ptr(), stream))
        got = int(count.item())
        if got != nnz:
            raise _capi.RqxError(_capi.RQX_ERR_HIP, f'hit pass wrote {got} entries, count pass {nnz}')
        rows, cols, vals = rows[:nnz], cols[:nnz].to(torch.int64), torch.view_as_complex(vals[:nnz])
        order = torch.argsort(rows * max(self.n_total, 1) + cols)
        return rows[order], cols[order], vals[order]

    def algorithmic_bytes(self, dtype, hits=None):
        """Per-matvec algorithmic bytes of the HBM model.

        ``8 N + 8 P + s_v H + s_v N + T (8 + s_c) + 12 J`` with ``P = N J'``
        (every (row, non-diagonal group) pair is resolved through the two-level
        index) -- or ``P = 0`` when every group of the problem comes from the
        product-subspace lists (``list_mode`` 2: no pair is searched) --, ``H``
        structural nonzeros incl. the diagonal, ``s_v`` the vector element bytes
        and ``s_c`` the coefficient bytes read per term.
        """
        s_v = torch.empty(0, dtype=dtype).element_size()
        s_c = 16 if dtype in (torch.float64, torch.complex128) else 8
        if hits is None:
            hits = self.count()['hits']
        total = s_v * hits
        for p in self.problems:
            n, J, T = p['n_states'], p['n_groups'], p['n_terms']
            jp = 0 if p['list_mode'] == 2 else J - (1 if p['has_diagonal'] else 0)
            total += 8 * n + 8 * n * jp + s_v * n + T * (8 + s_c) + 12 * J
        return total

    # ------------------------------------------------------------------ hit cache
    def build_cache(self, precision=64, max_bytes=0):
        """Store every hit (source row, ``C_j(s_i)``) once: ``cache_strategy=(1, 2)``.

        The reference caches source indices and coefficient sums for all ``J x N``
        (group, row) pairs (``sqd.py:343-407``); this cache keeps only the hits,
        as sliced ELL (64 rows per slice, hits of groups without low-bit flips in
        a section of their own so that their gathers coalesce).  ``precision`` 64 serves float64 and
        complex128 matvecs, 32 float32 and complex64, 0 builds both in one pass
        (one column array).  Later ``matvec`` calls of that precision read the
        cache and return the same values as the matrix-free kernel.  Raises ``RqxError`` when the cache needs more than
        ``max_bytes`` (0: 90% of the free device memory).
        """
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_plan_build_cache(self._handle, int(precision), int(max_bytes), stream))
        return self.cache_info(precision if precision else 64)

    def drop_cache(self, precision=0):
        """Release the cache of one precision (32 or 64) or both (0)."""
        _capi.check(_capi.lib().rqx_plan_drop_cache(self._handle, int(precision)))

    def cache_info(self, precision=64):
        """``built``, stored hits, stored values, device bytes, slices, complex values,
        stored columns, first-section levels, second-section values (incl. padding),
        first-section values of the levels without a single value, second-section
        hits, and the padding entries among the stored first-section values and
        columns and the first-section levels of padding only."""
        out = np.zeros(14, dtype=np.int64)
        _capi.check(_capi.lib().rqx_plan_cache_info(self._handle, int(precision), _capi.ptr(out)))
        keys = ('built', 'nnz', 'entries', 'bytes', 'slices', 'complex', 'col_entries', 'a_levels',
                'b_entries', 'a_entries', 'b_hits', 'a_pad_values', 'a_pad_columns', 'a_pad_levels')
        info = {k: int(v) for k, v in zip(keys, out)}
        info['built'] = bool(info['built'])
        info['complex'] = bool(info['complex'])
        return info

    def use_cache(self, enable=True):
        """Which path ``matvec`` takes once a cache is built: ``True`` the cache always,
        ``False`` the matrix-free kernels, ``'auto'`` (a new plan's default) the faster
        of the two -- ``build_cache`` times both for each complex dtype the cache serves
        whose matrix-free matvec runs the product kernel (real dtypes keep the cache).
        Both paths give the same bits."""
        mode = 2 if enable == 'auto' else int(bool(enable))
        _capi.check(_capi.lib().rqx_plan_use_cache(self._handle, mode))

    def matvec_path(self, dtype):
        """The kernel ``matvec`` runs for ``dtype`` now: ``'cache'``, ``'product'`` (matrix-free,
        every hit from the product-subspace lists) or ``'tiles'`` (matrix-free tile kernels)."""
        out = np.zeros(1, dtype=np.int32)
        _capi.check(_capi.lib().rqx_plan_matvec_path(self._handle, _DTYPE_CODES[dtype], _capi.ptr(out)))
        return ('tiles', 'product', 'cache')[int(out[0])]

    def cached_bytes(self, dtype):
        """Bytes one cached matvec must move: the cache, the diagonal, ``x`` once and ``y``.

        ``entries s_val + 4 col_entries + 8 a_levels + 116 slices + N (s_d + 2 s_v)``
        with ``s_val`` the stored value bytes, ``col_entries`` the stored
        columns, ``a_levels`` the first-section level codes (offset and
        single-value flag), 116 bytes per slice (four offsets and the section-A
        width: 36, the second section's row order: 64, its chunk widths: 16),
        ``s_d`` the diagonal entry bytes (complex) and ``s_v`` the vector element
        bytes; gathers of ``x`` past its first read are counted as cache hits.
        """
        precision = 64 if dtype in (torch.float64, torch.complex128) else 32
        info = self.cache_info(precision)
        if not info['built']:
            raise ValueError(f'no {precision}-bit cache is built')
        s_v = torch.empty(0, dtype=dtype).element_size()
        s_val = (8 if precision == 64 else 4) * (2 if info['complex'] else 1)
        s_d = 16 if precision == 64 else 8
        return (info['entries'] * s_val + 4 * info['col_entries'] + 8 * info['a_levels'] + 116 * info['slices'] +
                self.n_states * (s_d + 2 * s_v))

    def cached_padding_bytes(self, dtype):
        """Bytes of ``cached_bytes`` that are padding: second-section entries beyond each
        row's hits (value and column), padding lanes of the stored first-section values
        and columns, and first-section levels of padding only (codes and value)."""
        precision = 64 if dtype in (torch.float64, torch.complex128) else 32
        info = self.cache_info(precision)
        if not info['built']:
            raise ValueError(f'no {precision}-bit cache is built')
        s_val = (8 if precision == 64 else 4) * (2 if info['complex'] else 1)
        b_pad = info['b_entries'] - info['b_hits']
        return (b_pad * (s_val + 4) + info['a_pad_values'] * s_val + info['a_pad_columns'] * 4 +
                info['a_pad_levels'] * (8 + s_val))

    # ------------------------------------------------------------------ per-problem vector ops
    def dots(self, pairs):
        """Per-problem ``sum conj(a) * b`` for each ``(a, b)`` pair -> (K, len(pairs)) complex128.

        One fused pass over all pairs; a single-problem plan gives ``K = 1``.
        """
        n = len(pairs)
        a_ptrs = (ctypes.c_void_p * n)(*[a.data_ptr() for a, _ in pairs])
        b_ptrs = (ctypes.c_void_p * n)(*[b.data_ptr() for _, b in pairs])
        out = torch.empty((self.n_problems, n, 2), dtype=torch.float64, device=self.device)
        dt = pairs[0][0].dtype
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_batch_dots(self._handle, n, a_ptrs, b_ptrs, out.data_ptr(),
                                               _DTYPE_CODES[dt], stream))
        return torch.view_as_complex(out)

    def combine(self, vectors, coef, out):
        """``out = sum_q coef[:, q] * vectors[q]`` per problem (coef: (K, n) complex128, device)."""
        n = len(vectors)
        v_ptrs = (ctypes.c_void_p * n)(*[v.data_ptr() for v in vectors])
        c = torch.view_as_real(coef.to(torch.complex128).contiguous()).contiguous()
        stream = _capi.current_stream(self.device)
        _capi.check(_capi.lib().rqx_batch_combine(self._handle, n, v_ptrs, c.data_ptr(), out.data_ptr(),
                                                  _DTYPE_CODES[out.dtype], stream))
        return out


class BatchPlan(Plan):
    """Many (Hamiltonian, subspace) problems resident together; every kernel covers all of them.

    Vectors are the concatenation of the problems' rows (``row_offsets``); term
    outputs (``grad_coef``) the concatenation of their terms.  The matvec runs
    persistent workgroups on a tile queue ordered largest tile first.

    Args:
        problems: sequence of ``(TermTables, sorted uint64 states)``.
    """

    def __init__(self, problems, device=None, split_bits=-1):
        _require_cuda()
        lib = _capi.lib()
        self.tables = [t for t, _ in problems]
        K = len(problems)
        if K < 1:
            raise ValueError('a batch needs at least one problem')
        self.device = torch.device('cuda', torch.cuda.current_device() if device is None else
                                   torch.device(device).index or 0)
        keep = []

        def arr(a, dtype):
            a = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1))
            keep.append(a)
            return a.ctypes.data if a.size else None

        ptrs = {k: (ctypes.c_void_p * K)() for k in ('x', 'g', 'z', 'c', 's')}
        nq = (ctypes.c_int32 * K)()
        ng = (ctypes.c_int32 * K)()
        nt = (ctypes.c_int64 * K)()
        ns = (ctypes.c_int64 * K)()
        self._states = []
        for k, (t, states) in enumerate(problems):
            st = np.ascontiguousarray(np.asarray(states, dtype=np.uint64).reshape(-1))
            self._states.append(st)
            ptrs['x'][k] = arr(t.xmask, np.uint64)
            ptrs['g'][k] = arr(t.grp_off, np.int64)
            ptrs['z'][k] = arr(t.zmask, np.uint64)
            ptrs['c'][k] = arr(np.ascontiguousarray(t.coef, dtype=np.complex128).view(np.float64), np.float64)
            ptrs['s'][k] = st.ctypes.data if st.size else None
            nq[k], ng[k], nt[k], ns[k] = t.num_qubits, t.num_groups, t.num_terms, st.shape[0]
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            stream = _capi.current_stream(self.device)
            rc = lib.rqx_plan_create_batch(ctypes.byref(handle), K, nq, ptrs['x'], ptrs['g'], ng, ptrs['z'],
                                           ptrs['c'], nt, ptrs['s'], ns, int(split_bits), stream)
        _capi.check(rc)
        self._handle = handle
        self._load_info()
        self.n_states = int(self.info['n_states'])
        self.n_total = int(self.info['n_total'])
        offs = [p['row_offset'] for p in self.problems] + [self.n_total]
        self.row_offsets = torch.tensor(offs, dtype=torch.int64, device=self.device)
        self.row_problem = torch.repeat_interleave(
            torch.arange(K, device=self.device), self.row_offsets[1:] - self.row_offsets[:-1])

    @property
    def states(self):
        return np.concatenate(self._states)

    def lookup(self, values=False):
        raise NotImplementedError('lookup tables are per problem: build a Plan for one problem')

def eigenpair_batch(mats: torch.Tensor):
    """Lowest eigenpairs of ``(K, n, n)`` Hermitian device matrices -> (theta (K,), kappa (K, n))."""
    K, n = mats.shape[0], mats.shape[1]
    m = mats.to(torch.complex128).contiguous()
    out = torch.empty((K, 2 + 2 * n), dtype=torch.float64, device=mats.device)
    stream = _capi.current_stream(mats.device)
    _capi.check(_capi.lib().rqx_eigenpair_batch(m.data_ptr(), K, n, out.data_ptr(), stream))
    return out[:, 0], torch.view_as_complex(out[:, 2:].reshape(K, n, 2).contiguous())


def sort_unique_u64(keys: torch.Tensor):
    """Sorted unique copy of a uint64 device tensor (``uniquify_states``, ``sqd.py:445-482``)."""
    _require_cuda()
    if keys.dtype not in (torch.uint64, torch.int64):
        raise ValueError('keys must be a 64-bit integer tensor')
    keys = keys.contiguous()
    out = torch.empty_like(keys)
    n_unique = ctypes.c_int64(0)
    stream = _capi.current_stream(keys.device)
    with torch.cuda.device(keys.device):
        _capi.check(_capi.lib().rqx_sort_unique_u64(keys.data_ptr() if keys.numel() else None, keys.numel(),
                                                    out.data_ptr() if out.numel() else None,
                                                    ctypes.byref(n_unique), stream))
    return out[:n_unique.value]


def pack_to_u64(rows: torch.Tensor, num_qubits: int):
    """Packed state rows on the device (``uint8 (N, ceil((n + 1) / 8))``, pad bit
    first, ``sqd.py:255``) -> int64 device tensor of the real rows' uint64 states
    in input order (rows with the pad bit set are fillers and dropped)."""
    _require_cuda()
    if rows.dtype != torch.uint8 or rows.dim() != 2 or not rows.is_cuda:
        raise ValueError('rows must be a 2-d uint8 device tensor')
    rows = rows.contiguous()
    out = torch.empty(rows.shape[0], dtype=torch.int64, device=rows.device)
    n_real = ctypes.c_int64(0)
    stream = _capi.current_stream(rows.device)
    with torch.cuda.device(rows.device):
        _capi.check(_capi.lib().rqx_pack_to_u64(rows.data_ptr() if rows.numel() else None, rows.shape[0],
                                                rows.shape[1], int(num_qubits),
                                                out.data_ptr() if out.numel() else None,
                                                ctypes.byref(n_real), stream))
    return out[:n_real.value]


def eigenpair_small(mat: torch.Tensor):
    """Lowest eigenpair of a 2x2 / 3x3 Hermitian matrix -> (theta, kappa) tensors on its device.

    A device matrix runs the one-thread HIP kernel on the current stream; a CPU
    matrix runs the same routine on the calling thread.
    """
    n = mat.shape[0]
    m = mat.to(torch.complex128).contiguous()
    out = torch.empty(2 + 2 * n, dtype=torch.float64, device=mat.device)
    if mat.is_cuda:
        stream = _capi.current_stream(mat.device)
        _capi.check(_capi.lib().rqx_eigenpair_small(m.data_ptr(), n, out.data_ptr(), stream))
    else:
        _capi.check(_capi.lib().rqx_eigenpair_small_host(m.data_ptr(), n, out.data_ptr()))
    theta = out[0]
    kappa = torch.view_as_complex(out[2:].reshape(n, 2))
    return theta, kappa
