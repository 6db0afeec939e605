"""MI355X-native block-sparse tensor hot path of QuantiT (btensor contraction,
block SVD/eigh with truncation, two-site DMRG), Python side.

The arithmetic runs in ``libquantit_amd.so`` (host C++ sector bookkeeping +
gfx950 HIP kernels) through its C ABI (``include/quantit_amd.h``); this module
only marshals plain arrays.  There is no CPU fallback: importing works without
a GPU, but creating a :class:`Context` needs the built library and a device.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libquantit_amd.so"

PRECISION_FP64 = 0
PRECISION_SPLIT_FP32 = 1

# The Python types pybind11 gives the reference's exceptions in its own
# binding (python_binding/*.cpp): invalid_argument and domain_error become
# ValueError, out_of_range IndexError, logic_error (no translator of its
# own) RuntimeError, and torch's translator maps c10::Error to RuntimeError.
_STATUS_EXC = {
    1: ValueError,        # std::invalid_argument
    2: IndexError,        # std::out_of_range
    3: ValueError,        # std::domain_error
    4: RuntimeError,      # std::logic_error (NaN in eig2x2)
    5: RuntimeError,      # c10::Error (TORCH_CHECK)
    6: MemoryError,
    7: RuntimeError,
    8: RuntimeError,
}


class QttError(RuntimeError):
    pass


class CheckError(QttError):
    """Raised where the reference raises c10::Error through TORCH_CHECK."""


_lib = None


def library():
    """Loads libquantit_amd.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: run `python -m quantit_amd.build`")
        _lib = C.CDLL(str(LIB_PATH))
        _declare(_lib)
    return _lib


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
I64P = C.POINTER(C.c_int64)
I32P = C.POINTER(C.c_int32)
I16P = C.POINTER(C.c_int16)
U16P = C.POINTER(C.c_uint16)
DP = C.POINTER(C.c_double)


class DmrgOptionsC(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("convergence_criterion", C.c_double), ("maximum_bond", C.c_uint64),
                ("minimum_bond", C.c_uint64), ("maximum_iterations", C.c_uint64)]


LOGGER_INIT_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(DmrgOptionsC))
LOGGER_LOG_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, C.c_double, C.c_int, C.POINTER(C.c_void_p))


class DmrgLoggerC(C.Structure):
    _fields_ = [("user", C.c_void_p), ("init", LOGGER_INIT_FN), ("it_log_all", LOGGER_LOG_FN),
                ("end_log_all", LOGGER_LOG_FN)]


# every symbol of include/quantit_amd.h with its argument types
SIGNATURES = {
    "qtt_last_error": (C.c_char_p, []),
    "qtt_version": (C.c_int, []),
    "qtt_context_create": (C.c_int, [C.c_int, C.c_size_t, PP]),
    "qtt_context_destroy": (C.c_int, [P]),
    "qtt_context_set_precision": (C.c_int, [P, C.c_int]),
    "qtt_context_set_per_block": (C.c_int, [P, C.c_int]),
    "qtt_context_set_overlap": (C.c_int, [P, C.c_int]),
    "qtt_context_set_profiling": (C.c_int, [P, C.c_int]),
    "qtt_context_sync": (C.c_int, [P]),
    "qtt_context_memory": (C.c_int, [P, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
    "qtt_context_kernel_times": (C.c_int, [P, C.c_char_p, C.c_size_t, C.c_int]),
    "qtt_context_svd_stats": (C.c_int, [P, I64P, I64P, I64P]),
    "qtt_context_svd_flops": (C.c_int, [P, DP, DP]),
    "qtt_context_set_lanczos": (C.c_int, [P, C.c_int, C.c_double]),
    "qtt_context_set_graphs": (C.c_int, [P, C.c_int]),
    "qtt_context_lanczos_stats": (C.c_int, [P, I64P, I64P, I64P]),
    "qtt_btensor_create": (C.c_int, [P, C.c_int, I64P, I64P, C.c_int, U16P, I16P, I16P, C.c_int64, I32P, DP, PP]),
    "qtt_btensor_destroy": (C.c_int, [P]),
    "qtt_btensor_info": (C.c_int, [P, C.POINTER(C.c_int), C.POINTER(C.c_int), I64P, I64P, I64P]),
    "qtt_btensor_structure": (C.c_int, [P, I64P, I64P, I16P, I16P, U16P]),
    "qtt_btensor_blocks": (C.c_int, [P, I32P, DP]),
    "qtt_permute": (C.c_int, [P, I64P, PP]),
    "qtt_reshape": (C.c_int, [P, C.c_int, I64P, PP]),
    "qtt_conj": (C.c_int, [P, PP]),
    "qtt_tensordot": (C.c_int, [P, P, C.c_int, I64P, I64P, PP]),
    "qtt_last_tensordot_flops": (C.c_double, []),
    "qtt_tensordot_into": (C.c_int, [P, P, C.c_int, I64P, I64P, P]),
    "qtt_gemm_tuning_save": (C.c_int, [C.c_char_p]),
    "qtt_gemm_tuning_load": (C.c_int, [C.c_char_p]),
    "qtt_gemm_tuning_size": (C.c_int, [I64P]),
    "qtt_mul": (C.c_int, [P, P, PP]),
    "qtt_add": (C.c_int, [P, P, C.c_double, PP]),
    "qtt_full_dot": (C.c_int, [P, P, DP]),
    "qtt_svd": (C.c_int, [P, C.c_size_t, PP, PP, PP]),
    "qtt_svd_truncated": (C.c_int, [P, C.c_size_t, C.c_double, C.c_size_t, C.c_size_t, C.c_double, PP, PP, PP]),
    "qtt_eigh": (C.c_int, [P, C.c_size_t, PP, PP]),
    "qtt_truncate": (C.c_int, [P, P, P, C.c_size_t, C.c_size_t, C.c_double, C.c_double, PP, PP, PP]),
    "qtt_svd_matrix": (C.c_int, [P, C.c_int, C.c_int, PP, PP, PP]),
    "qtt_eigh_truncated": (C.c_int, [P, C.c_size_t, C.c_double, C.c_size_t, C.c_size_t, C.c_double, PP, PP]),
    "qtt_truncate_eigh": (C.c_int, [P, P, C.c_size_t, C.c_size_t, C.c_double, C.c_double, PP, PP]),
    "qtt_dmrg": (C.c_int, [C.c_int, PP, PP, C.POINTER(DmrgOptionsC), DP, C.c_int, DP, DP, I64P, C.POINTER(C.c_int),
                           DP, DP]),
    "qtt_hamil2site_times_state": (C.c_int, [P, P, P, P, PP]),
    "qtt_one_step_lanczos": (C.c_int, [P, P, P, P, PP, DP, DP, DP]),
    "qtt_eig2x2": (C.c_int, [P, C.c_double, C.c_double, C.c_double, DP, DP, DP]),
    "qtt_dmrg_logged": (C.c_int, [C.c_int, PP, PP, C.c_int, C.POINTER(DmrgOptionsC), C.POINTER(DmrgLoggerC), DP]),
    "qtt_dmrg_from_center": (C.c_int, [C.c_int, PP, PP, C.c_int, C.POINTER(DmrgOptionsC), DP, C.c_int, DP, DP, I64P,
                                       C.POINTER(C.c_int), DP, DP]),
    "qtt_mps_move_oc": (C.c_int, [C.c_int, PP, C.c_int, C.c_int]),
    "qtt_mps_prepare": (C.c_int, [C.c_int, PP]),
    "qtt_contract_norm": (C.c_int, [C.c_int, PP, PP, DP]),
    "qtt_contract_energy": (C.c_int, [C.c_int, PP, PP, PP, DP]),
}


def _declare(lib):
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args


def _check(status):
    if status != 0:
        msg = library().qtt_last_error().decode(errors="replace")
        if status == 5:
            raise CheckError(msg)
        raise _STATUS_EXC.get(status, QttError)(msg)


def _arr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


class Context:
    """A device, its HIP stream and its HBM arena (all tensor data lives there)."""

    def __init__(self, device=0, arena_bytes=4 << 30, precision=PRECISION_FP64):
        lib = library()
        h = C.c_void_p()
        _check(lib.qtt_context_create(device, arena_bytes, C.byref(h)))
        self.h = h
        self.set_precision(precision)

    def close(self):
        if self.h:
            library().qtt_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, p):
        _check(library().qtt_context_set_precision(self.h, int(p)))

    def set_per_block(self, enable):
        _check(library().qtt_context_set_per_block(self.h, int(bool(enable))))

    def set_overlap(self, enable):
        """Allow (default) or forbid a second stream for independent work of
        one call; processes running several contexts at once forbid it."""
        _check(library().qtt_context_set_overlap(self.h, int(bool(enable))))

    def set_profiling(self, enable):
        _check(library().qtt_context_set_profiling(self.h, int(bool(enable))))

    def sync(self):
        _check(library().qtt_context_sync(self.h))

    def memory(self):
        u, p, c = C.c_size_t(), C.c_size_t(), C.c_size_t()
        _check(library().qtt_context_memory(self.h, C.byref(u), C.byref(p), C.byref(c)))
        return {"used": u.value, "peak": p.value, "capacity": c.value}

    def svd_flops(self):
        """Block-SVD model FLOPs since creation: (all sectors, wide sectors)."""
        a, w = C.c_double(), C.c_double()
        _check(library().qtt_context_svd_flops(self.h, C.byref(a), C.byref(w)))
        return a.value, w.value

    def set_lanczos(self, max_krylov, tol=1e-12):
        """DMRG eigensolver: 0 = the reference's one-step Lanczos (default);
        2..64 = k-step Lanczos with full reorthogonalisation, stopping at
        residual <= tol |E|."""
        _check(library().qtt_context_set_lanczos(self.h, int(max_krylov), float(tol)))

    def set_graphs(self, enable):
        """k-step mode: replay each update's matvec as a captured HIP graph
        (default) or run it eagerly."""
        _check(library().qtt_context_set_graphs(self.h, int(bool(enable))))

    def lanczos_stats(self):
        u, v, g = C.c_int64(), C.c_int64(), C.c_int64()
        _check(library().qtt_context_lanczos_stats(self.h, C.byref(u), C.byref(v), C.byref(g)))
        return {"updates": u.value, "vectors": v.value, "graph_replays": g.value}

    def svd_stats(self):
        """Block SVD/eigh health: calls, largest sweep count of one call, and
        sectors that reached the Jacobi sweep cap still rotating."""
        c, m, u = C.c_int64(), C.c_int64(), C.c_int64()
        _check(library().qtt_context_svd_stats(self.h, C.byref(c), C.byref(m), C.byref(u)))
        return {"calls": c.value, "max_sweeps": m.value, "unconverged": u.value}

    def kernel_times(self, reset=True):
        buf = C.create_string_buffer(1 << 16)
        _check(library().qtt_context_kernel_times(self.h, buf, len(buf), int(reset)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, ms, n = line.split()
            out[name] = (float(ms), int(n))
        return out


class BTensor:
    """Handle on a device block-sparse tensor (mirrors quantit::btensor)."""

    __slots__ = ("ctx", "h")

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def __del__(self):
        try:
            if self.h:
                library().qtt_btensor_destroy(self.h)
                self.h = None
        except Exception:
            pass

    # -- construction / readback -------------------------------------------------
    @staticmethod
    def from_parts(ctx, groups, secs, cvals, sel, blocks):
        """groups: moduli (0 = Z); secs/cvals: per-dimension section sizes and
        charge tuples; sel: charge tuple; blocks: {index tuple: ndarray}."""
        rank = len(secs)
        ng = len(groups)
        nsec = np.array([len(s) for s in secs], dtype=np.int64)
        sizes = np.array([x for s in secs for x in s], dtype=np.int64)
        cv = np.array([c for d in cvals for ch in d for c in ch], dtype=np.int16).reshape(-1)
        if cv.size == 0:
            cv = np.zeros(1, dtype=np.int16)
        if sizes.size == 0:
            sizes = np.zeros(1, dtype=np.int64)
        if nsec.size == 0:
            nsec = np.zeros(1, dtype=np.int64)
        gm = np.array(groups, dtype=np.uint16)
        sl = np.array(sel, dtype=np.int16)
        keys = sorted(blocks)
        idx = np.array([list(k) for k in keys], dtype=np.int32).reshape(-1)
        if idx.size == 0:
            idx = np.zeros(1, dtype=np.int32)
        data = [np.ascontiguousarray(blocks[k], dtype=np.float64).reshape(-1) for k in keys]
        flat = np.concatenate(data) if data else np.zeros(1)
        if flat.size == 0:
            flat = np.zeros(1)
        h = C.c_void_p()
        _check(library().qtt_btensor_create(ctx.h, rank, _arr(nsec, C.c_int64), _arr(sizes, C.c_int64), ng,
                                            _arr(gm, C.c_uint16), _arr(cv, C.c_int16), _arr(sl, C.c_int16),
                                            len(keys), _arr(idx, C.c_int32), _arr(flat, C.c_double), C.byref(h)))
        return BTensor(ctx, h)

    def info(self):
        r, ng, ts, nb, ne = C.c_int(), C.c_int(), C.c_int64(), C.c_int64(), C.c_int64()
        _check(library().qtt_btensor_info(self.h, C.byref(r), C.byref(ng), C.byref(ts), C.byref(nb), C.byref(ne)))
        return r.value, ng.value, ts.value, nb.value, ne.value

    @property
    def rank(self):
        return self.info()[0]

    def structure(self):
        """(groups, secs, cvals, sel)."""
        r, ng, ts, nb, ne = self.info()
        nsec = np.zeros(max(r, 1), dtype=np.int64)
        sizes = np.zeros(max(ts, 1), dtype=np.int64)
        cv = np.zeros(max(ts * ng, 1), dtype=np.int16)
        sl = np.zeros(ng, dtype=np.int16)
        gm = np.zeros(ng, dtype=np.uint16)
        _check(library().qtt_btensor_structure(self.h, _arr(nsec, C.c_int64), _arr(sizes, C.c_int64),
                                               _arr(cv, C.c_int16), _arr(sl, C.c_int16), _arr(gm, C.c_uint16)))
        secs, cvals = [], []
        pos = 0
        for d in range(r):
            n = int(nsec[d])
            secs.append([int(x) for x in sizes[pos:pos + n]])
            cvals.append([tuple(int(c) for c in cv[(pos + i) * ng:(pos + i + 1) * ng]) for i in range(n)])
            pos += n
        return tuple(int(x) for x in gm), secs, cvals, tuple(int(x) for x in sl)

    def block_list(self):
        """Block indices in the tensor's list order."""
        r, ng, ts, nb, ne = self.info()
        idx = np.zeros(max(nb * r, 1), dtype=np.int32)
        _check(library().qtt_btensor_blocks(self.h, _arr(idx, C.c_int32), None))
        return [tuple(int(x) for x in idx[b * r:(b + 1) * r]) for b in range(nb)]

    def to_parts(self):
        """(groups, secs, cvals, sel, blocks dict) copied to the host."""
        groups, secs, cvals, sel = self.structure()
        r, ng, ts, nb, ne = self.info()
        idx = np.zeros(max(nb * r, 1), dtype=np.int32)
        data = np.zeros(max(ne, 1), dtype=np.float64)
        _check(library().qtt_btensor_blocks(self.h, _arr(idx, C.c_int32), _arr(data, C.c_double)))
        blocks = {}
        pos = 0
        for b in range(nb):
            key = tuple(int(x) for x in idx[b * r:(b + 1) * r])
            shp = tuple(secs[d][key[d]] for d in range(r))
            n = int(np.prod(shp)) if shp else 1
            blocks[key] = data[pos:pos + n].reshape(shp).copy()
            pos += n
        return groups, secs, cvals, sel, blocks

    # -- operations ----------------------------------------------------------------
    def _new(self, fn, *args):
        h = C.c_void_p()
        _check(fn(*args, C.byref(h)))
        return BTensor(self.ctx, h)

    def tensordot(self, other, dims_self, dims_other):
        da = np.array(list(dims_self) or [0], dtype=np.int64)
        db = np.array(list(dims_other) or [0], dtype=np.int64)
        return self._new(library().qtt_tensordot, self.h, other.h, len(list(dims_self)), _arr(da, C.c_int64),
                         _arr(db, C.c_int64))

    def tensordot_into(self, other, dims_self, dims_other, out):
        """The contraction written into ``out`` (an existing tensor with the
        result's structure) with no device allocation: temporaries live in
        the library's static per-GPU scratch.  Returns ``out``."""
        da = np.array(list(dims_self) or [0], dtype=np.int64)
        db = np.array(list(dims_other) or [0], dtype=np.int64)
        _check(library().qtt_tensordot_into(self.h, other.h, len(list(dims_self)), _arr(da, C.c_int64),
                                            _arr(db, C.c_int64), out.h))
        return out

    def permute(self, perm):
        p = np.array(list(perm), dtype=np.int64)
        return self._new(library().qtt_permute, self.h, _arr(p, C.c_int64))

    def reshape(self, index_groups):
        g = np.array(list(index_groups) or [0], dtype=np.int64)
        return self._new(library().qtt_reshape, self.h, len(list(index_groups)), _arr(g, C.c_int64))

    def conj(self):
        return self._new(library().qtt_conj, self.h)

    def mul(self, other):
        return self._new(library().qtt_mul, self.h, other.h)

    def add(self, other, alpha=1.0):
        return self._new(library().qtt_add, self.h, other.h, float(alpha))

    def full_dot(self, other):
        out = C.c_double()
        _check(library().qtt_full_dot(self.h, other.h, C.byref(out)))
        return out.value


def gemm_tuning_save(path):
    """Writes the process-wide GEMM tile-variant table as JSON."""
    _check(library().qtt_gemm_tuning_save(str(path).encode()))


def gemm_tuning_load(path):
    """Merges a table written by gemm_tuning_save into the process's."""
    _check(library().qtt_gemm_tuning_load(str(path).encode()))


def gemm_tuning_size():
    n = C.c_int64()
    _check(library().qtt_gemm_tuning_size(C.byref(n)))
    return n.value


def last_tensordot_flops():
    return library().qtt_last_tensordot_flops()


def tensordot(a, b, da, db):
    """Free ``tensordot(l, r, dl, dr)`` (include/blockTensor/btensor.h:1230-1234)."""
    return a.tensordot(b, da, db)


def _three(ctx, fn, *args):
    hs = [C.c_void_p() for _ in range(3)]
    _check(fn(*args, *[C.byref(h) for h in hs]))
    return tuple(BTensor(ctx, h) for h in hs)


def svd(t, split, tol=None, min_size=1, max_size=2 ** 64 - 1, pow=2.0):
    """``svd(btensor, split)`` and, with ``tol``, the truncating
    ``svd(A, split, tol, min_size, max_size, pow)``
    (include/blockTensor/LinearAlgebra.h:87, :115-116)."""
    if tol is None:
        return _three(t.ctx, library().qtt_svd, t.h, int(split))
    return _three(t.ctx, library().qtt_svd_truncated, t.h, int(split), float(tol), int(min_size), int(max_size),
                  float(pow))


def svd_matrix(t, some=True, compute_uv=True):
    """``svd(btensor, some, compute_uv)`` of a rank-2 tensor
    (include/blockTensor/LinearAlgebra.h:74).  ``some=False`` and
    ``compute_uv=False`` accept square sectors only (ValueError otherwise);
    ``compute_uv=False`` returns zero-filled U and V."""
    return _three(t.ctx, library().qtt_svd_matrix, t.h, int(bool(some)), int(bool(compute_uv)))


def eigh(t, split, tol=None, min_size=1, max_size=2 ** 64 - 1, pow=1.0):
    """``eigh(btensor, split)`` (include/blockTensor/LinearAlgebra.h:167):
    (eigenvalues ascending per sector, U); with ``tol`` the truncating
    ``eigh(A, split, tol, min_size, max_size, pow=1)`` (:180-181)."""
    hs = [C.c_void_p(), C.c_void_p()]
    if tol is None:
        _check(library().qtt_eigh(t.h, int(split), C.byref(hs[0]), C.byref(hs[1])))
    else:
        _check(library().qtt_eigh_truncated(t.h, int(split), float(tol), int(min_size), int(max_size), float(pow),
                                            C.byref(hs[0]), C.byref(hs[1])))
    return BTensor(t.ctx, hs[0]), BTensor(t.ctx, hs[1])


def truncate_eigh(e, S, max_size, min_size, tol, pow=1.0):
    """``truncate(btensor&& e, btensor&& S, max, min, tol, pow)``
    (include/blockTensor/LinearAlgebra.h:242-243): (e, S) truncated together."""
    hs = [C.c_void_p(), C.c_void_p()]
    _check(library().qtt_truncate_eigh(e.h, S.h, int(max_size), int(min_size), float(tol), float(pow),
                                       C.byref(hs[0]), C.byref(hs[1])))
    return BTensor(e.ctx, hs[0]), BTensor(e.ctx, hs[1])


def truncate(U, d, V, max_size, min_size, tol, pow=2.0):
    """``truncate(std::tuple<U, d, V>&&, max, min, tol, pow)``
    (include/blockTensor/LinearAlgebra.h:246-247)."""
    return _three(U.ctx, library().qtt_truncate, U.h, d.h, V.h, int(max_size), int(min_size), float(tol),
                  float(pow))


class DmrgOptions:
    """``dmrg_options`` (include/dmrg_options.h:15-57) with its defaults."""

    def __init__(self, cutoff=1e-6, convergence_criterion=1e-5, maximum_bond=2 ** 64 - 1, minimum_bond=4,
                 maximum_iterations=1000):
        self.cutoff = cutoff
        self.convergence_criterion = convergence_criterion
        self.maximum_bond = maximum_bond
        self.minimum_bond = minimum_bond
        self.maximum_iterations = maximum_iterations


class DmrgLogger:
    """``dmrg_logger`` (include/dmrg_logger.h:27-108): dmrg calls init(options)
    before the first sweep, it_log_all(iteration, energy, state) after every
    sweep and end_log_all(iteration, energy, state) at the end; by default
    both go to log_all, which calls log_step, log_energy and log_bond_dims.
    ``state`` is the list of the MPS tensors, lent for the call only (the
    handles are invalidated when it returns)."""

    def init(self, options):
        pass

    def log_step(self, iteration):
        pass

    def log_energy(self, energy):
        pass

    def log_bond_dims(self, state):
        pass

    def log_all(self, iteration, energy, state):
        self.log_step(iteration)
        self.log_energy(energy)
        self.log_bond_dims(state)

    def it_log_all(self, iteration, energy, state):
        self.log_all(iteration, energy, state)

    def end_log_all(self, iteration, energy, state):
        self.log_all(iteration, energy, state)


def bond_dims(state):
    """Right bond dimension of every MPS tensor (sum of its last leg's sections)."""
    return [int(sum(t.structure()[1][-1])) for t in state]


def dmrg(mpo, mps, options, oc=0, seed=0, logger=None):
    """``dmrg(bMPO&, bMPS&, options, logger)`` (include/dmrg.h:39-40) on a
    state whose orthogonality centre is at site ``oc`` (moved back there at
    the end, as the reference does).  Returns (energy, new mps list, log
    dict).

    With a charge (an int, or a tuple of ints per charge group) in place of
    ``mps``: ``dmrg(bMPO&, qnum, options, logger)`` (sources/dmrg.cpp:109-118),
    which starts from ``random_bMPS(options.minimum_bond, hamil, qnum)``: a
    random state in that sector on the MPO's physical legs
    (``models.random_bmps_general`` seeded with ``seed``), optimised from
    site 0 without a canonical preparation, as the reference does."""
    if isinstance(mps, (int, np.integer)) or (isinstance(mps, tuple) and all(
            isinstance(c, (int, np.integer)) for c in mps)):
        from . import models
        phys, groups = models.mpo_phys_legs(mpo)
        q = (int(mps),) if isinstance(mps, (int, np.integer)) else tuple(int(c) for c in mps)
        if len(q) != len(groups):
            raise ValueError(f"charge {q} does not match the MPO's {len(groups)} charge groups")
        parts = models.random_bmps_general(int(options.minimum_bond), phys, q, seed, groups)
        mps = [BTensor.from_parts(mpo[0].ctx, **p) for p in parts]
        oc = 0
    L = len(mpo)
    mpo_h = (C.c_void_p * L)(*[t.h for t in mpo])
    mps_h = (C.c_void_p * L)(*[t.h for t in mps])
    # the library destroys the handles it replaces
    for t in mps:
        t.h = None
    o = DmrgOptionsC(options.cutoff, options.convergence_criterion, min(options.maximum_bond, 2 ** 64 - 1),
                     options.minimum_bond, options.maximum_iterations)
    if logger is not None:
        return _dmrg_logged(mpo, mpo_h, mps_h, L, int(oc), o, options, logger)
    nmax = int(min(options.maximum_iterations, 100000))
    E = C.c_double()
    es = np.zeros(max(nmax, 1))
    ms = np.zeros(max(nmax, 1))
    mb = np.zeros(max(nmax, 1), dtype=np.int64)
    ns = C.c_int()
    cf, sf = C.c_double(), C.c_double()
    ctx = mpo[0].ctx
    st = library().qtt_dmrg_from_center(L, C.cast(mpo_h, PP), C.cast(mps_h, PP), int(oc), C.byref(o), C.byref(E),
                                        nmax, _arr(es, C.c_double), _arr(ms, C.c_double), _arr(mb, C.c_int64),
                                        C.byref(ns), C.byref(cf), C.byref(sf))
    new_mps = [BTensor(ctx, C.c_void_p(mps_h[i])) for i in range(L)]
    _check(st)
    n = ns.value
    log = {"energies": es[:n].tolist(), "sweep_ms": ms[:n].tolist(), "max_bond": mb[:n].tolist(),
           "contract_flops": cf.value, "svd_flops": sf.value}
    return E.value, new_mps, log


def _dmrg_logged(mpo, mpo_h, mps_h, L, oc, o, options, logger):
    """qtt_dmrg_logged with a Python logger object (see DmrgLogger).  An
    exception raised by a hook is kept and raised once dmrg returns."""
    ctx = mpo[0].ctx
    err = []
    log = {"energies": [], "max_bond": []}

    def lent(n, hs):
        return [BTensor(ctx, C.c_void_p(hs[i])) for i in range(n)]

    def call(fn, *a):
        if err:
            return
        try:
            fn(*a)
        except BaseException as e:  # noqa: BLE001 - reported after dmrg returns
            err.append(e)

    def on_init(_user, _opts):
        call(getattr(logger, "init", lambda _o: None), options)

    def on_it(_user, it, E, n, hs):
        state = lent(n, hs)
        log["energies"].append(E)
        log["max_bond"].append(max(bond_dims(state)))
        call(getattr(logger, "it_log_all", lambda *_: None), int(it), float(E), state)
        for t in state:  # the handles are lent: the library releases them
            t.h = None

    def on_end(_user, it, E, n, hs):
        state = lent(n, hs)
        call(getattr(logger, "end_log_all", lambda *_: None), int(it), float(E), state)
        for t in state:
            t.h = None

    cb = DmrgLoggerC(None, LOGGER_INIT_FN(on_init), LOGGER_LOG_FN(on_it), LOGGER_LOG_FN(on_end))
    E = C.c_double()
    st = library().qtt_dmrg_logged(L, C.cast(mpo_h, PP), C.cast(mps_h, PP), oc, C.byref(o), C.byref(cb), C.byref(E))
    new_mps = [BTensor(ctx, C.c_void_p(mps_h[i])) for i in range(L)]
    _check(st)
    if err:
        raise err[0]
    return E.value, new_mps, log


def hamil2site_times_state(theta, W12, L, R):
    """``details::hamil2site_times_state`` (sources/dmrg.cpp:520-531)."""
    h = C.c_void_p()
    _check(library().qtt_hamil2site_times_state(theta.h, W12.h, L.h, R.h, C.byref(h)))
    return BTensor(theta.ctx, h)


def one_step_lanczos(theta, W12, L, R):
    """``details::one_step_lanczos`` (sources/dmrg.cpp:584-608): returns
    (Hpsi, a0, a1, b) with Hpsi the normalised Lanczos vector."""
    h = C.c_void_p()
    a0, a1, b = C.c_double(), C.c_double(), C.c_double()
    _check(library().qtt_one_step_lanczos(theta.h, W12.h, L.h, R.h, C.byref(h), C.byref(a0), C.byref(a1),
                                          C.byref(b)))
    return BTensor(theta.ctx, h), a0.value, a1.value, b.value


def eig2x2(ctx, a0, a1, b):
    """``details::eig2x2Mat`` (sources/dmrg.cpp:543-572) on the device:
    (E, o, n); a NaN raises RuntimeError (the reference's logic_error)."""
    E, o, n = C.c_double(), C.c_double(), C.c_double()
    _check(library().qtt_eig2x2(ctx.h, float(a0), float(a1), float(b), C.byref(E), C.byref(o), C.byref(n)))
    return E.value, o.value, n.value


def _replace_mps(fn, mps, *args):
    """Calls fn(L, handles, *args); the library replaces the handles in place
    on success (the old ones are destroyed) and leaves them on failure."""
    L = len(mps)
    ctx = mps[0].ctx
    hs = (C.c_void_p * L)(*[t.h for t in mps])
    st = fn(L, C.cast(hs, PP), *args)
    _check(st)
    for t in mps:
        t.h = None
    return [BTensor(ctx, C.c_void_p(hs[i])) for i in range(L)]


def mps_move_oc(mps, oc, target):
    """``bMPS::move_oc`` (sources/MPT.cpp:74-111) on device; returns the new list."""
    return _replace_mps(library().qtt_mps_move_oc, mps, int(oc), int(target))


def mps_prepare(mps):
    """Normalise a bMPS and leave it right-canonical with the orthogonality
    centre at 0, as the reference's drivers do before ``dmrg``
    (tests/DMRG_heisenberg_test.h:34-37); returns the new list."""
    return _replace_mps(library().qtt_mps_prepare, mps)


def contract_norm(a, b):
    L = len(a)
    ah = (C.c_void_p * L)(*[t.h for t in a])
    bh = (C.c_void_p * L)(*[t.h for t in b])
    out = C.c_double()
    _check(library().qtt_contract_norm(L, C.cast(ah, PP), C.cast(bh, PP), C.byref(out)))
    return out.value


def contract_energy(a, b, mpo):
    L = len(a)
    ah = (C.c_void_p * L)(*[t.h for t in a])
    bh = (C.c_void_p * L)(*[t.h for t in b])
    mh = (C.c_void_p * L)(*[t.h for t in mpo])
    out = C.c_double()
    _check(library().qtt_contract_energy(L, C.cast(ah, PP), C.cast(bh, PP), C.cast(mh, PP), C.byref(out)))
    return out.value
