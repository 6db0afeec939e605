"""libtorch-CPU restatement of QuantiT's two-site update: the reference's own
per-block calls on the host.

TEST INFRASTRUCTURE / CPU BASELINE: imported only by ``tests/`` and the
``cpu_baseline`` leg of ``bench.py``, never by the product.

The reference runs its hot path on libtorch's CPU kernels (MKL): per output
block of a contraction one ``torch::mm`` and one ``addmm_`` per further
matching pair (sources/btensor.cpp:2095, :2102) on blocks permuted and
reshaped as ``permute_bl`` does (:1843-1894); per SVD sector one
``torch::svd`` (sources/btensor_linalg.cpp:483) of the sector gathered by
``compact_dense`` (:80-150); the truncation on the host (:657-755).  This
module makes the same calls through PyTorch's CPU backend in the same loop
order, with the block structure planned by the numpy oracle
(``btensor_oracle.tensordot_plan``, ``reorder_by_cvals``).  It is what the
benchmark times as the reference's CPU path.

Differences from the reference that do not change the arithmetic being
timed: blocks live in the oracle's ``BTensor`` container (a dict), and the
truncation drops the U/V blocks of an erased sector outright instead of
through ``remove_unit_blocks`` (whose list-order quirk only decides whether
some empty blocks survive).  ``tests/test_torch_cpu.py`` pins it against the
numpy oracle.
"""
from __future__ import annotations

import numpy as np
import torch

from . import btensor_oracle as bo
from . import dmrg_oracle as dor
from .btensor_oracle import BTensor


def from_oracle(t):
    out = t.copy_shape()
    out.blocks = {k: torch.from_numpy(np.ascontiguousarray(v, dtype=np.float64)) for k, v in t.blocks.items()}
    return out


def to_oracle(t):
    out = t.copy_shape()
    out.blocks = {k: v.contiguous().numpy().copy() for k, v in t.blocks.items()}
    return out


def tensordot(a, b, da, db, flops=None):
    """``btensor::tensordot`` (sources/btensor.cpp:1971-2119) with libtorch
    ``mm`` / ``addmm_``; flops[0] accumulates sum 2 M N K."""
    out, plan, p1, p2 = bo.tensordot_plan(a, b, da, db)
    dl = len(da)
    split_a = a.rank - dl
    for oidx, pairs in plan:
        acc = None
        for ia, ib in pairs:
            A = a.blocks[ia].permute(p1)
            M = int(np.prod(A.shape[:split_a]))
            A = A.reshape(M, -1)
            B = b.blocks[ib].permute(p2)
            K = int(np.prod(B.shape[:dl]))
            B = B.reshape(K, -1)
            if acc is None:
                acc = torch.mm(A, B)
            else:
                acc.addmm_(A, B)
            if flops is not None:
                flops[0] += 2.0 * A.shape[0] * A.shape[1] * B.shape[1]
        out.blocks[oidx] = acc.view(out.block_shape(oidx))
    return out


def conj(t):
    out = t.copy_shape()
    out.blocks = dict(t.blocks)
    return out.inverse_cvals_()


def permute(t, perm):
    out = BTensor(t.groups, [t.secs[p] for p in perm], [t.cvals[p] for p in perm], t.sel)
    out.blocks = {tuple(k[p] for p in perm): v.permute(perm) for k, v in t.blocks.items()}
    return out


def dot(x, y):
    """Rank-0 ``tensordot(x, y.conj(), all, all)`` as a number."""
    acc = 0.0
    for k, v in x.blocks.items():
        w = y.blocks.get(k)
        if w is not None:
            acc += float(torch.dot(v.reshape(-1), w.reshape(-1)))
    return acc


def axpy(x, y, alpha):
    """``add(other, alpha)``: flat_map merge (sources/btensor.cpp:2666-2683)."""
    out = x.copy_shape()
    out.blocks = dict(x.blocks)
    for k, v in y.blocks.items():
        out.blocks[k] = out.blocks[k] + alpha * v if k in out.blocks else alpha * v
    return out


def scal(x, alpha):
    out = x.copy_shape()
    out.blocks = {k: v * alpha for k, v in x.blocks.items()}
    return out


def mul_last(t, d):
    """Broadcast multiply by a rank-1 tensor along the last leg (d's section =
    the last index), as ``v.mul_(d)`` / ``u.mul_(d)`` (sources/dmrg.cpp:190-199)."""
    out = t.copy_shape()
    out.blocks = {k: v * d.blocks[(k[-1],)] for k, v in t.blocks.items() if (k[-1],) in d.blocks}
    return out


def svd_truncated(t, split, tol, min_size, max_size, pow_=2.0, flops=None):
    """Truncating ``svd(A, split, tol, min, max, pow)``: sectors gathered as
    ``compact_dense``, one ``torch.linalg.svd`` per sector, the cut of
    ``truncate_impl`` (s_last (1 - 2 eps), strict >)."""
    rt = t.reshape([split])
    g = rt.groups
    n = bo.q_neutral(g)
    host = rt.copy_shape()   # numpy views of the blocks for the sector gather
    host.blocks = {k: v.contiguous().numpy() for k, v in rt.blocks.items()}
    sectors = bo.compact_dense(host)
    rd_c = [rt.cvals[1][cols[0][0]] for (_, _, _, cols) in sectors]
    dsz = [min(dense.shape) for (dense, _, _, _) in sectors]
    d = BTensor(g, [dsz], [[n] * len(dsz)], n)
    U = BTensor(g, [list(rt.secs[0]), dsz], [list(rt.cvals[0]), rd_c], rt.sel)
    V = BTensor(g, [list(rt.secs[1]), dsz], [list(rt.cvals[1]), [bo.q_inv(g, c) for c in rd_c]], n)
    for b_i, (dense, _, rows, cols) in enumerate(sectors):
        u, s, vh = torch.linalg.svd(torch.from_numpy(dense), full_matrices=False)
        v = vh.T
        if flops is not None:
            m_, n_ = max(dense.shape), min(dense.shape)
            flops[0] += 14.0 * m_ * n_ * n_ + 8.0 * n_ ** 3
        for rw, a, b in rows:
            U.blocks[(rw, b_i)] = u[a:b, :]
        for c, a, b in cols:
            V.blocks[(c, b_i)] = v[a:b, :]
        d.blocks[(b_i,)] = s
    V.inverse_cvals_()
    # truncation (sources/btensor_linalg.cpp:657-755)
    vd = -np.sort(-np.concatenate([d.blocks[k].numpy() for k in sorted(d.blocks)]))
    last = bo.compute_last_index(vd, tol, pow_, min_size, max_size)
    s_last = float(vd[last])
    thr = s_last - 2.0 * s_last * bo.EPS64
    for (kd,), db in sorted(d.blocks.items(), reverse=True):
        k = int((db > thr).sum())
        if k == 0:
            del d.blocks[(kd,)]
            for T in (U, V):
                T.blocks = {ix: v for ix, v in T.blocks.items() if ix[-1] != kd}
        else:
            d.blocks[(kd,)] = db[:k]
            d.secs[0][kd] = k
            for T in (U, V):
                T.secs[-1][kd] = k
                for ix in [ix for ix in T.blocks if ix[-1] == kd]:
                    T.blocks[ix] = T.blocks[ix][:, :k]
    U_shape = t.shape_from([-1 if i < split else 0 for i in range(t.rank)])
    V_shape = t.shape_from([0 if i < split else -1 for i in range(t.rank)])
    U = U.reshape_as(bo.shape_product(U_shape, U.shape_from([0, -1])))
    V = V.reshape_as(bo.shape_product(V_shape.inverse_cvals(), V.shape_from([0, -1])))
    return U, d, V


def left_env(W, Y, L, flops=None):
    out = tensordot(L, Y, [0], [0], flops)
    out = tensordot(out, W, [0, 2], [0, 3], flops)
    return tensordot(out, conj(Y), [0, 2], [0, 1], flops)


def right_env(W, Y, R, flops=None):
    out = tensordot(R, Y, [0], [2], flops)
    out = tensordot(out, W, [0, 3], [2, 3], flops)
    return tensordot(out, conj(Y), [3, 0], [1, 2], flops)


def matvec(theta, W12, L, R, flops=None):
    out = tensordot(L, theta, [0], [0], flops)
    out = tensordot(out, W12, [0, 2, 3], [0, 4, 5], flops)
    return tensordot(out, R, [1, 4], [0, 1], flops)


def environments(mpo, mps, oc):
    """Left environments of sites < oc and right ones of sites > oc + 1
    (generate_env_impl, sources/dmrg.cpp:370-409), with the edges."""
    L = len(mpo)
    Ls = dor.edge_shape_prep(to_oracle_shape(mps[0]), 0)
    LH = dor.edge_shape_prep(to_oracle_shape(mpo[0]), 0)
    env = {-1: from_oracle(bo.ones_like(bo.shape_product(Ls, LH, Ls.inverse_cvals())))}
    Rs = dor.edge_shape_prep(to_oracle_shape(mps[-1]), 2)
    RH = dor.edge_shape_prep(to_oracle_shape(mpo[-1]), 2)
    env[L] = from_oracle(bo.ones_like(bo.shape_product(Rs, RH, Rs.inverse_cvals())))
    for i in range(oc):
        env[i] = left_env(mpo[i], mps[i], env[i - 1])
    for i in range(L - 1, oc + 1, -1):
        env[i] = right_env(mpo[i], mps[i], env[i + 1])
    return env


def to_oracle_shape(t):
    return t.copy_shape()


def two_site_update(mpo, mps, env, oc, step, cutoff, min_bond, max_bond, flops=None):
    """``dmrg_2sites_update::operator()`` (sources/dmrg.cpp:163-206) with the
    one-step Lanczos (:584-608) and eig2x2Mat (:543-572); mps and env are
    updated in place; returns the energy."""
    W12 = permute(tensordot(mpo[oc], mpo[oc + 1], [2], [0], flops), [0, 1, 3, 4, 2, 5])
    theta = tensordot(mps[oc], mps[oc + 1], [2], [0], flops)
    Le, Re = env[oc - 1], env[oc + 2]
    psi = matvec(theta, W12, Le, Re, flops)
    a0 = dot(psi, theta)
    psi = axpy(psi, theta, -a0)
    b = float(np.sqrt(dot(psi, psi)))
    if abs(b) >= 1e-15:
        psi = scal(psi, 1.0 / b)
    a1 = dot(psi, matvec(psi, W12, Le, Re, flops))
    E, o, n = dor.eig2x2(a0, a1, b)
    theta2 = axpy(scal(theta, o), psi, n)
    U, d, V = svd_truncated(theta2, 2, cutoff, min_bond, max_bond, flops=flops)
    nrm = float(np.sqrt(sum(float((v * v).sum()) for v in d.blocks.values())))
    d = scal(d, 1.0 / nrm)
    if step == 1:
        mps[oc] = U
        mps[oc + 1] = permute(conj(mul_last(V, d)), [2, 0, 1])
        env[oc] = left_env(mpo[oc], mps[oc], env[oc - 1], flops)
    else:
        mps[oc] = mul_last(U, d)
        mps[oc + 1] = permute(conj(V), [2, 0, 1])
        env[oc + 1] = right_env(mpo[oc + 1], mps[oc + 1], env[oc + 2], flops)
    return E
