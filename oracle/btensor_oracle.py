"""numpy restatement of QuantiT's block-sparse tensor (``btensor``) hot path.

TEST INFRASTRUCTURE: this module is the oracle.  It is imported only by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py``; the product library never calls it.

Pinning (each function lists how it is checked in ``tests/test_oracle_*.py``):
  * charges            -- group axioms on Z (int16 wrap) and C<N>; tuple order.
  * tensordot          -- dense ``np.tensordot`` of ``to_dense`` on seeded inputs
                          (mirrors include/blockTensor/btensor.h:1581-1608),
                          trace invariance over all 24 index orders (:1551-1580),
                          conservation of every output block.
  * reshape/reshape_as -- round trip, exact block sizes/charges (:1682-1725).
  * svd                -- U.diag(d).V^T reconstruction, orthonormality, singular
                          values equal to dense SVD of ``to_dense`` (library
                          routine), sector order against a hand-built case
                          (include/blockTensor/LinearAlgebra.h:275-465).
  * truncate           -- closed-form rule on synthetic spectra incl. exact ties
                          and the min/max precedence (sources/LinearAlgebra.cpp:57-75,
                          sources/btensor_linalg.cpp:657-755).
  * eigh               -- A.U = U.Lambda and dense ``eigvalsh`` equality.

Charges are tuples of ints; ``groups`` is a tuple of moduli, 0 for the
integer group Z (int16 with wrap-around, include/Conserved/quantity.h:175-239)
and N for the cyclic group C<N> (uint16 mod N, :62-165).  Charge order is the
lexicographic tuple order (include/Conserved/Composite/quantity_impl.h:326).

Reference configuration followed: FP64 blocks with the global torch default
dtype set to double (as tests/2Dheisenberg.cpp:287 does), so the dense sector
matrices of ``compact_dense_single`` and the truncation threshold ``tol**pow``
are FP64.
"""
from __future__ import annotations

import itertools
import numpy as np

EPS64 = float(np.finfo(np.float64).eps)


# ---------------------------------------------------------------------------
# conserved charges  (include/Conserved/quantity.h, Composite/quantity_impl.h)
# ---------------------------------------------------------------------------
def _wrap16(x: int) -> int:
    x &= 0xFFFF
    return x - 0x10000 if x >= 0x8000 else x


def q_neutral(groups):
    return tuple(0 for _ in groups)


def q_add(groups, a, b):
    """Group operation (``op``): Z adds with int16 wrap (quantity.h:215-219),
    C<N> adds mod N (quantity.h:96-101)."""
    return tuple(((x + y) % n) if n else _wrap16(x + y) for n, x, y in zip(groups, a, b))


def q_inv(groups, a):
    """Group inverse: Z negates (quantity.h:220-224); C<N> maps v -> (N-v) if v
    else 0 (quantity.h:107-111)."""
    return tuple(((n - x) if x else 0) if n else _wrap16(-x) for n, x in zip(groups, a))


def q_make(groups, values):
    """Construct a charge like ``C(int16_t)`` (quantity.h:80) / ``Z(int16_t)``."""
    out = []
    for n, v in zip(groups, values):
        if n:
            v = int(v)
            out.append((v != -n) * (v < 0) * n + int(np.fmod(v, n)))
        else:
            out.append(_wrap16(int(v)))
    return tuple(out)


def q_sum(groups, charges):
    acc = q_neutral(groups)
    for c in charges:
        acc = q_add(groups, acc, c)
    return acc


# ---------------------------------------------------------------------------
# btensor  (include/blockTensor/btensor.h:105-978, sources/btensor.cpp)
# ---------------------------------------------------------------------------
class BTensor:
    """Block-sparse tensor: per-dimension sections with sizes and charges, a
    selection rule, and a lexicographically ordered map block-index -> block
    (btensor.h:820-837; flat_map.h:31-33)."""

    __slots__ = ("groups", "secs", "cvals", "sel", "blocks")

    def __init__(self, groups, secs, cvals, sel, blocks=None):
        self.groups = tuple(groups)
        self.secs = [list(map(int, s)) for s in secs]
        self.cvals = [[tuple(c) for c in cv] for cv in cvals]
        self.sel = tuple(sel)
        self.blocks = {} if blocks is None else dict(blocks)

    # -- construction helpers --------------------------------------------
    @staticmethod
    def from_sections(groups, sections, sel):
        """``btensor(vec_list_t, selection_rule)`` (btensor.cpp:341-357)."""
        secs = [[int(s) for s, _ in d] for d in sections]
        cvals = [[tuple(c) for _, c in d] for d in sections]
        return BTensor(groups, secs, cvals, sel)

    def copy_shape(self):
        return BTensor(self.groups, [list(s) for s in self.secs], [list(c) for c in self.cvals], self.sel)

    def copy(self):
        out = self.copy_shape()
        out.blocks = {k: v.copy() for k, v in self.blocks.items()}
        return out

    # -- structure ----------------------------------------------------------
    @property
    def rank(self):
        return len(self.secs)

    def sizes(self):
        return [sum(s) for s in self.secs]

    def block_shape(self, idx):
        return tuple(self.secs[d][i] for d, i in enumerate(idx))

    def block_charge_sum(self, idx):
        return q_sum(self.groups, (self.cvals[d][i] for d, i in enumerate(idx)))

    def allowed(self, idx):
        """``block_conservation_rule_test`` (btensor.cpp:363-372)."""
        return self.block_charge_sum(idx) == self.sel

    def sorted_items(self):
        return sorted(self.blocks.items(), key=lambda kv: kv[0])

    def all_allowed_indices(self):
        for idx in itertools.product(*[range(len(s)) for s in self.secs]):
            if self.allowed(idx):
                yield idx

    def check(self):
        """Invariants of ``check_tensor`` (btensor.cpp:405-467); returns a message
        (empty when the tensor is valid)."""
        msg = []
        for idx, b in self.blocks.items():
            if len(idx) != self.rank:
                msg.append(f"block {idx} has wrong index length")
                continue
            if any(i >= len(self.secs[d]) for d, i in enumerate(idx)):
                msg.append(f"block {idx} index out of range")
                continue
            if not self.allowed(idx):
                msg.append(f"block {idx} violates selection rule")
            if tuple(b.shape) != self.block_shape(idx):
                msg.append(f"block {idx} shape {b.shape} != {self.block_shape(idx)}")
        return "\n".join(msg)

    # -- dense conversion ---------------------------------------------------
    def offsets(self, d):
        return np.concatenate([[0], np.cumsum(self.secs[d])]).astype(int)

    def to_dense(self):
        """``to_dense`` (btensor.cpp:2491-2503)."""
        out = np.zeros(self.sizes(), dtype=np.float64)
        offs = [self.offsets(d) for d in range(self.rank)]
        for idx, b in self.blocks.items():
            sl = tuple(slice(offs[d][i], offs[d][i] + self.secs[d][i]) for d, i in enumerate(idx))
            out[sl] = b
        return out

    def fill_from_dense(self, dense, cutoff=0.0):
        """``from_basic_tensor_like``: keep every allowed block whose slice holds
        an element with |x| > cutoff (btensor.cpp:2505-2529)."""
        offs = [self.offsets(d) for d in range(self.rank)]
        out = self.copy_shape()
        for idx in itertools.product(*[range(len(s)) for s in self.secs]):
            sl = tuple(slice(offs[d][i], offs[d][i] + self.secs[d][i]) for d, i in enumerate(idx))
            blk = np.asarray(dense[sl], dtype=np.float64)
            if blk.size and np.any(np.abs(blk) > cutoff):
                if not self.allowed(idx):
                    raise ValueError(f"dense tensor has weight in forbidden block {idx}")
                out.blocks[idx] = blk.copy()
        return out

    # -- shape transforms -----------------------------------------------------
    def shape_from(self, dims):
        """Member ``shape_from`` (btensor.cpp:492-583): keep dims marked -1, fix
        the others at an element index and absorb the inverse charge of the
        section holding that element into the selection rule."""
        if len(dims) != self.rank:
            raise ValueError("The argument lenght must match the tensor's rank")
        sel = self.sel
        secs, cvals = [], []
        for d, a in enumerate(dims):
            if a < 0:
                secs.append(list(self.secs[d]))
                cvals.append(list(self.cvals[d]))
            else:
                e = a
                s = 0
                while s < len(self.secs[d]) and e >= self.secs[d][s]:
                    e -= self.secs[d][s]
                    s += 1
                if s >= len(self.secs[d]):
                    raise ValueError("element index too large")
                sel = q_add(self.groups, sel, q_inv(self.groups, self.cvals[d][s]))
        return BTensor(self.groups, secs, cvals, sel)

    def neutral_shape_(self):
        """``neutral_shape_`` (btensor.cpp:769-780)."""
        if self.blocks:
            raise RuntimeError("Neutral shape can only function correctly on an empty tensor.")
        n = q_neutral(self.groups)
        self.sel = n
        self.cvals = [[n for _ in cv] for cv in self.cvals]
        return self

    def set_selection_rule_(self, sel):
        if self.blocks:
            raise RuntimeError("This transformation can only be applied to empty btensors")
        self.sel = tuple(sel)
        return self

    def shift_selection_rule_(self, shift):
        if self.blocks:
            raise RuntimeError("This transformation can only be applied to empty btensors")
        self.sel = q_add(self.groups, self.sel, shift)
        return self

    def inverse_cvals_(self):
        """``inverse_cvals_`` (btensor.cpp:2170-2178)."""
        g = self.groups
        self.cvals = [[q_inv(g, c) for c in cv] for cv in self.cvals]
        self.sel = q_inv(g, self.sel)
        return self

    def inverse_cvals(self):
        out = self.copy()
        return out.inverse_cvals_()

    def conj(self):
        """``conj`` = complex conjugate + inverse charges (btensor.cpp:2156-2161).
        Real FP64 blocks are unchanged."""
        out = self.copy()
        return out.inverse_cvals_()

    def permute(self, perm):
        """``permute`` (btensor.cpp:1754-1802)."""
        r = self.rank
        perm = [p + r if p < 0 else p for p in perm]
        out = BTensor(self.groups, [self.secs[p] for p in perm], [self.cvals[p] for p in perm], self.sel)
        for idx, b in self.blocks.items():
            out.blocks[tuple(idx[p] for p in perm)] = np.ascontiguousarray(np.transpose(b, perm))
        return out

    def reshape(self, index_groups):
        """``reshape`` (btensor.cpp:2986-3024): legs [g_k, g_{k+1}) fuse into one
        leg whose sections enumerate the section combinations row-major (last
        index fastest, ``increment_index_right``), sizes multiply, charges add.
        No fusion by charge."""
        bounds = [0] + list(index_groups) + [self.rank]
        g = self.groups
        secs, cvals = [], []
        for a, b in zip(bounds[:-1], bounds[1:]):
            s_out, c_out = [], []
            for combo in itertools.product(*[range(len(self.secs[d])) for d in range(a, b)]):
                size = 1
                ch = q_neutral(g)
                for d, i in zip(range(a, b), combo):
                    size *= self.secs[d][i]
                    ch = q_add(g, ch, self.cvals[d][i])
                s_out.append(size)
                c_out.append(ch)
            secs.append(s_out)
            cvals.append(c_out)
        out = BTensor(g, secs, cvals, self.sel)
        for idx, blk in self.blocks.items():
            new_idx = []
            for a, b in zip(bounds[:-1], bounds[1:]):
                flat = 0
                for d in range(a, b):
                    flat = flat * len(self.secs[d]) + idx[d]
                new_idx.append(flat)
            new_idx = tuple(new_idx)
            out.blocks[new_idx] = blk.reshape(out.block_shape(new_idx))
        return out

    def reshape_as(self, other, overwrite_c_vals=False):
        """``reshape_as`` (btensor.cpp:3026-3083): the flat (row-major) block
        index maps between the two section layouts; every possible block must
        carry the same flux (dims_only mode) and element count."""
        n_self = int(np.prod([len(s) for s in self.secs])) if self.rank else 1
        n_other = int(np.prod([len(s) for s in other.secs])) if other.rank else 1
        if n_self != n_other:
            raise ValueError("incompatible sections layouts")
        g = self.groups
        it_self = itertools.product(*[range(len(s)) for s in self.secs])
        it_other = itertools.product(*[range(len(s)) for s in other.secs])
        for a, b in zip(it_self, it_other):
            if not overwrite_c_vals and self.block_charge_sum(a) != other.block_charge_sum(b):
                raise ValueError("incompatible conserved quantities")
            if int(np.prod(self.block_shape(a))) != int(np.prod(other.block_shape(b))):
                raise ValueError("incompatible block dimensions")
        sel = other.sel if overwrite_c_vals else self.sel
        out = BTensor(g, other.secs, other.cvals, sel)
        for idx, blk in self.blocks.items():
            flat = 0
            for d, i in enumerate(idx):
                flat = flat * len(self.secs[d]) + i
            new_idx = []
            for d in reversed(range(other.rank)):
                n = len(other.secs[d])
                new_idx.append(flat % n)
                flat //= n
            new_idx = tuple(reversed(new_idx))
            if overwrite_c_vals and not other.allowed(new_idx):
                raise ValueError("block not allowed by the new selection rule")
            out.blocks[new_idx] = blk.reshape(other.block_shape(new_idx))
        return out


def shape_product(*tensors):
    """Free ``shape_from(a, b, ...)``: concatenated legs, selection rules added
    (btensor.h:1001-1093)."""
    g = tensors[0].groups
    secs, cvals = [], []
    sel = q_neutral(g)
    for t in tensors:
        secs += [list(s) for s in t.secs]
        cvals += [list(c) for c in t.cvals]
        sel = q_add(g, sel, t.sel)
    return BTensor(g, secs, cvals, sel)


def rand_like(shape, rng, dist="uniform"):
    """``rand_like``: every allowed block filled (btensor.cpp:2399-2404).  The
    values come from the caller's seeded generator (not torch's)."""
    out = shape.copy_shape()
    for idx in shape.all_allowed_indices():
        bs = out.block_shape(idx)
        if dist == "normal":
            out.blocks[idx] = rng.standard_normal(bs)
        else:
            out.blocks[idx] = rng.random(bs)
    return out


def ones_like(shape):
    out = shape.copy_shape()
    for idx in shape.all_allowed_indices():
        out.blocks[idx] = np.ones(out.block_shape(idx))
    return out


def eye_like(shape):
    """``eye_like`` for a rank-2 shape: identity on diagonal blocks."""
    out = shape.copy_shape()
    for idx in shape.all_allowed_indices():
        bs = out.block_shape(idx)
        if idx[0] == idx[1] and bs[0] == bs[1]:
            out.blocks[idx] = np.eye(bs[0])
    return out


# ---------------------------------------------------------------------------
# contraction  (sources/btensor.cpp:782-883, 1843-1946, 1971-2119)
# ---------------------------------------------------------------------------
class ContractionError(RuntimeError):
    """Plays the role of ``c10::Error`` from TORCH_CHECK."""


def check_product_compat(a, b, da, db, throws=True):
    """``check_product_compat`` (btensor.cpp:782-825): contracted dims must have
    equal section counts and pairwise opposite charges."""
    if len(da) != len(db):
        if throws:
            raise ContractionError("both dimension lists should have the same length.")
        return False
    g = a.groups
    n = q_neutral(g)
    for x, y in zip(da, db):
        if len(a.secs[x]) != len(b.secs[y]):
            if throws:
                raise ContractionError(
                    f"contracted dimensions need to match, but first has {len(a.secs[x])} sections along dim {x} "
                    f"and second has {len(b.secs[y])} sections along dim {y}")
            return False
        for c1, c2 in zip(a.cvals[x], b.cvals[y]):
            if q_add(g, c1, c2) != n:
                if throws:
                    raise ContractionError(
                        f"contracted conserved numbers need to sum to zero, but there is a violation when "
                        f"contracting dim {x} of the left tensor with dim {y} of the right tensor")
                return False
    return True


def tensordot_plan(a, b, da, db):
    """Structure of ``btensor::tensordot`` (btensor.cpp:1971-2119) without the
    arithmetic: returns (out shape, plan) where plan is the ordered list of
    (out index, [(a index, b index), ...]) in the merge-join order, the pair
    list being the accumulation order (first pair -> ``mm``, others ->
    ``addmm_``)."""
    da = [d + a.rank if d < 0 else d for d in da]
    db = [d + b.rank if d < 0 else d for d in db]
    if a.groups != b.groups:
        raise ContractionError("the two tensors have different type of conserved quantities")
    check_product_compat(a, b, da, db, throws=True)
    dl = len(da)
    free_a = [i for i in range(a.rank) if i not in da]
    free_b = [i for i in range(b.rank) if i not in db]
    p1 = free_a + list(da)
    p2prime = free_b + list(db)
    out = BTensor(a.groups, [a.secs[i] for i in free_a] + [b.secs[i] for i in free_b],
                  [a.cvals[i] for i in free_a] + [b.cvals[i] for i in free_b],
                  q_add(a.groups, a.sel, b.sel))
    t1 = sorted((tuple(idx[p] for p in p1), idx) for idx in a.blocks)
    t2 = sorted((tuple(idx[p] for p in p2prime), idx) for idx in b.blocks)
    na = len(free_a)
    nb = len(free_b)

    def groups_of(t, nfree):
        res = []
        for key, grp in itertools.groupby(t, key=lambda kv: kv[0][:nfree]):
            res.append((key, list(grp)))
        return res

    g1 = groups_of(t1, na)
    g2 = groups_of(t2, nb)
    plan = []
    for pre1, grp1 in g1:
        for pre2, grp2 in g2:
            i = j = 0
            pairs = []
            while i < len(grp1) and j < len(grp2):
                s1 = grp1[i][0][na:]
                s2 = grp2[j][0][nb:]
                if s1 < s2:
                    i += 1
                elif s2 < s1:
                    j += 1
                else:
                    pairs.append((grp1[i][1], grp2[j][1]))
                    i += 1
            if pairs:
                plan.append((pre1 + pre2, pairs))
    return out, plan, p1, [*db, *free_b]


def tensordot(a, b, da, db):
    """``btensor::tensordot`` (btensor.cpp:1971-2119): per output block one
    ``mm`` plus one ``addmm_`` per further matching block pair, in merge-join
    order; blocks are matricised as ``permute_bl`` does (btensor.cpp:1843-1894):
    A block -> (free_a..., contracted...) reshaped (M,K), B block ->
    (contracted..., free_b...) reshaped (K,N)."""
    out, plan, p1, p2 = tensordot_plan(a, b, da, db)
    dl = len(da)
    split_a = a.rank - dl
    for oidx, pairs in plan:
        acc = None
        for ia, ib in pairs:
            A = np.transpose(a.blocks[ia], p1)
            M = int(np.prod(A.shape[:split_a]))
            A = A.reshape(M, -1)
            B = np.transpose(b.blocks[ib], p2)
            K = int(np.prod(B.shape[:dl]))
            B = B.reshape(K, -1)
            if acc is None:
                acc = A @ B
            else:
                acc += A @ B
        out.blocks[oidx] = acc.reshape(out.block_shape(oidx))
    return out


def tensordot_flops(a, b, da, db):
    """Host-counted algorithmic FLOPs of a contraction: sum over matched block
    pairs of 2*M*N*K."""
    _, plan, p1, p2 = tensordot_plan(a, b, da, db)
    dl = len(da)
    total = 0
    for oidx, pairs in plan:
        for ia, ib in pairs:
            sa = a.block_shape(ia)
            sb = b.block_shape(ib)
            k = int(np.prod([sa[d] for d in p1[a.rank - dl:]]))
            m = int(np.prod(sa)) // max(k, 1) if k else 0
            n = int(np.prod(sb)) // max(k, 1) if k else 0
            total += 2 * m * n * k
    return total


# ---------------------------------------------------------------------------
# elementwise / broadcast  (sources/btensor.cpp:977-1297, 2666-2752)
# ---------------------------------------------------------------------------
def _shape_compute(small, large):
    """``mul_helpers::shape_compute`` (btensor.cpp:980-1083)."""
    sr, lr = small.rank, large.rank
    g = large.groups
    comp = [0] * sr
    new_secs = [list(s) for s in large.secs]
    new_cvals = [list(c) for c in large.cvals]
    for n in range(1, sr + 1):
        s_sec = small.secs[sr - n]
        l_sec = large.secs[lr - n]
        m = int(len(s_sec) == 1 and s_sec[0] == 1) + 2 * int(len(l_sec) == 1 and l_sec[0] == 1)
        comp[sr - n] = m
        if not m and len(s_sec) != len(l_sec):
            raise ValueError("section mismatch at non-singleton dimension")
        if len(l_sec) > len(s_sec):
            src_secs, src_c, short_c = l_sec, large.cvals[lr - n], small.cvals[sr - n]
        else:
            src_secs, src_c, short_c = s_sec, small.cvals[sr - n], large.cvals[lr - n]
        cv = []
        k = len(short_c) - 1
        for j in reversed(range(len(src_c))):
            cv.append(q_add(g, src_c[j], short_c[k]))
            if k > 0:
                k -= 1
        new_secs[lr - n] = list(src_secs)
        new_cvals[lr - n] = list(reversed(cv))
    sel = q_add(g, large.sel, small.sel)
    return comp, new_secs, new_cvals, sel


def mul(x, y):
    """Broadcast ``mul`` (btensor.cpp:1750-1753 -> broadcast_operation
    :1631-1689): trailing dimensions align; a block pair contributes when the
    trailing indices match (or the dimension is a broadcast singleton); only
    matched pairs produce output blocks."""
    if x.rank >= y.rank:
        large, small, x_large = x, y, True
    else:
        large, small, x_large = y, x, False
    comp, secs, cvals, sel = _shape_compute(small, large)
    out = BTensor(large.groups, secs, cvals, sel)
    sr = small.rank
    other = y
    for other_idx, ob in other.blocks.items():
        for this_idx, tb in x.blocks.items():
            li, si = (this_idx, other_idx) if x_large else (other_idx, this_idx)
            ok = all(comp[k] or li[large.rank - sr + k] == si[k] for k in range(sr))
            if not ok:
                continue
            oidx = list(li)
            for k in range(sr):
                oidx[large.rank - sr + k] = max(li[large.rank - sr + k], si[k])
            out.blocks[tuple(oidx)] = tb * ob
    return out


def scal(x, alpha):
    """``mul(Scalar)``: every block scaled."""
    out = x.copy_shape()
    out.blocks = {k: v * alpha for k, v in x.blocks.items()}
    return out


def add(x, y, alpha=1.0):
    """``add(other, alpha)`` (btensor.cpp:2666-2683): ``flat_map::merge`` -- on
    collision a += alpha*b, blocks only in y are inserted as alpha*b."""
    if x.cvals != y.cvals:
        raise ValueError("The conserved quantities of the tensors must be a perfect match")
    if x.sel != y.sel:
        raise ValueError("The selection rules of the tensors must be the same")
    if x.secs != y.secs:
        raise ValueError("The blocks of the tensors must have the same dimensions")
    out = x.copy()
    for k, v in y.blocks.items():
        if k in out.blocks:
            out.blocks[k] = out.blocks[k] + alpha * v
        else:
            out.blocks[k] = v * alpha
    return out


def full_dot(x, y):
    """``tensordot(x, y, all dims, all dims)`` with y conjugated (real data):
    rank-0 result; the sum runs over blocks present in both, in sorted order
    (btensor.cpp:2048-2117 with an empty output index)."""
    t = tensordot(x, y.conj(), list(range(x.rank)), list(range(y.rank)))
    if not t.blocks:
        return 0.0, t
    return float(next(iter(t.blocks.values()))), t


# ---------------------------------------------------------------------------
# block linear algebra  (sources/btensor_linalg.cpp)
# ---------------------------------------------------------------------------
def reorder_by_cvals(t):
    """``reorder_by_cvals`` (btensor_linalg.cpp:30-67): stable sort of the
    (lexicographically ordered) block list by (row charge, column charge) of the
    last two dimensions."""
    r = t.rank
    items = t.sorted_items()
    return sorted(items, key=lambda kv: (t.cvals[r - 2][kv[0][r - 2]], t.cvals[r - 1][kv[0][r - 1]]))


def compact_dense(t):
    """``compact_dense`` + ``compact_dense_single`` (btensor_linalg.cpp:80-255):
    returns a list of sectors (dense matrix, other indices, row slices, column
    slices).  Rows get slices in order of first appearance of a new row section
    while walking the sector's blocks; columns likewise in encounter order, the
    slice list itself kept sorted by column section index."""
    r = t.rank
    items = reorder_by_cvals(t)

    def key(idx):
        return (idx[:r - 2], t.cvals[r - 2][idx[r - 2]], t.cvals[r - 1][idx[r - 1]])

    sectors = []
    start = 0
    while start < len(items):
        end = start + 1
        while end < len(items) and key(items[end][0]) == key(items[start][0]):
            end += 1
        grp = items[start:end]
        rows = []  # (row section, start, stop) in encounter order
        cols = {}  # col section -> (start, stop)
        row_acc = col_acc = 0
        cur_row = grp[0][0][r - 2]
        rs = grp[0][1].shape[r - 2]
        rows.append((cur_row, 0, rs))
        row_acc = rs
        for idx, blk in grp:
            c = idx[r - 1]
            if c not in cols:
                cs = blk.shape[r - 1]
                cols[c] = (col_acc, col_acc + cs)
                col_acc += cs
            if idx[r - 2] != cur_row:
                cur_row = idx[r - 2]
                rs = blk.shape[r - 2]
                rows.append((cur_row, row_acc, row_acc + rs))
                row_acc += rs
        other_shape = grp[0][1].shape[:r - 2]
        dense = np.zeros(tuple(other_shape) + (row_acc, col_acc))
        rmap = {rw: (a, b) for rw, a, b in rows}
        for idx, blk in grp:
            a, b = rmap[idx[r - 2]]
            c0, c1 = cols[idx[r - 1]]
            dense[..., a:b, c0:c1] = blk
        col_list = sorted((c, a, b) for c, (a, b) in cols.items())
        sectors.append((dense, tuple(grp[0][0][:r - 2]), rows, col_list))
        start = end
    return sectors


def svd_rank2(t):
    """``svd(const btensor&, some=true, uv=true)`` (btensor_linalg.cpp:390-502)
    for a rank-2 btensor.  Returns (U, d, V) with U (rows, k), d (k) carrying
    neutral charges, V (cols, k) with inverted charges."""
    if t.rank != 2:
        raise ValueError("svd_rank2 expects a rank-2 btensor")
    g = t.groups
    n = q_neutral(g)
    sectors = compact_dense(t)
    rd_c = [t.cvals[1][cols[0][0]] for (_, _, _, cols) in sectors]
    ld_c = [q_inv(g, c) for c in rd_c]
    dsz = [min(dense.shape) for (dense, _, _, _) in sectors]
    d = BTensor(g, [dsz], [[n] * len(dsz)], n)
    U = BTensor(g, [list(t.secs[0]), dsz], [list(t.cvals[0]), rd_c], t.sel)
    # V = shape_from(rank0, cols leg, leftD) with neutral rule, then inverse_cvals_
    V = BTensor(g, [list(t.secs[1]), dsz], [list(t.cvals[1]), ld_c], n)
    for b_i, (dense, _, rows, cols) in enumerate(sectors):
        u, s, vh = np.linalg.svd(dense, full_matrices=False)
        v = vh.T
        for rw, a, b in rows:
            U.blocks[(rw, b_i)] = u[a:b, :].copy()
        for c, a, b in cols:
            V.blocks[(c, b_i)] = v[a:b, :].copy()
        d.blocks[(b_i,)] = s.copy()
    V.inverse_cvals_()
    return U, d, V


def svd_split(t, split):
    """``svd(const btensor&, size_t split)`` (btensor_linalg.cpp:503-534)."""
    rt = t.reshape([split])
    rU, d, rV = svd_rank2(rt)
    U_shape = t.shape_from([-1 if i < split else 0 for i in range(t.rank)])
    V_shape = t.shape_from([0 if i < split else -1 for i in range(t.rank)])
    U = rU.reshape_as(shape_product(U_shape, rU.shape_from([0, -1])))
    V = rV.reshape_as(shape_product(V_shape.inverse_cvals(), rV.shape_from([0, -1])))
    return U, d, V


def eigh_rank2(t):
    """``eigh(const btensor&, BOOL)`` (btensor_linalg.cpp:294-389)."""
    g = t.groups
    n = q_neutral(g)
    sectors = compact_dense(t)
    rd_c = [t.cvals[1][cols[0][0]] for (_, _, _, cols) in sectors]
    dsz = [dense.shape[-1] for (dense, _, _, _) in sectors]
    d = BTensor(g, [dsz], [[n] * len(dsz)], n)
    u_sel = q_add(g, t.shape_from([-1, 0]).sel, q_inv(g, t.cvals[1][0]))
    U = BTensor(g, [list(t.secs[0]), dsz], [list(t.cvals[0]), rd_c], u_sel)
    for b_i, (dense, _, rows, cols) in enumerate(sectors):
        w, v = np.linalg.eigh(dense, UPLO="U")
        for rw, a, b in rows:
            idx = (rw, b_i)
            if not U.allowed(idx):
                raise ValueError(f"block index {idx} not allowed by selection rule")
            U.blocks[idx] = v[a:b, :].copy()
        d.blocks[(b_i,)] = w.copy()
    return d, U


def eigh_split(t, split):
    """``eigh(const btensor&, size_t split)`` (btensor_linalg.cpp:370-389)."""
    rt = t.reshape([split])
    d, rU = eigh_rank2(rt)
    U_shape = t.shape_from([-1 if i < split else 0 for i in range(t.rank)])
    U = rU.reshape_as(shape_product(U_shape, rU.shape_from([0, -1])))
    return d, U


def compute_last_index(vd, tol, pow_, min_size, max_size):
    """``compute_last_index`` (sources/LinearAlgebra.cpp:57-75) on a descending
    FP64 vector: walk from the end accumulating |d|^pow until the discarded
    weight would exceed tol^pow (and the index is below max_size), never going
    below min_size-1."""
    L = len(vd)
    toln = float(tol) ** pow_
    last = L - 1
    trunc = abs(float(vd[last])) ** pow_
    while last >= min_size:
        if trunc > toln and last < max_size:
            break
        last -= 1
        trunc += abs(float(vd[last])) ** pow_
    return last


def _remove_unit_blocks(U, kd):
    """``remove_unit_blocks`` lambda of ``truncate_impl``
    (btensor_linalg.cpp:690-709), restated step for step on the block list in
    its current order (the dict's insertion order stands for the flat_map's
    vector).  It advances past one matching block per iteration, so a
    matching block that directly follows another matching block is kept (as
    in the reference), and its swaps can leave the list unsorted: the next
    call (the next erased sector) then walks that unsorted order, as the
    reference's does."""
    lst = list(U.blocks.items())
    n = len(lst)
    src = dest = 0
    while dest != n:
        dest_last = lst[dest][0][-1]
        dest += int(dest_last == kd)
        if dest != src and dest != n:
            lst[dest], lst[src] = lst[src], lst[dest]
        dest += int(dest != n)
        src += 1
    keep = lst[:n - (dest - src)]
    U.blocks = dict(keep)


def truncate_impl(d, unitaries, max_size, min_size, tol, pow_=2.0):
    """``truncate_impl`` (btensor_linalg.cpp:657-755).

    All singular values are sorted descending; the cut value is
    s_last*(1 - 2*eps) (computed as s - 2*s*eps); each d block keeps its
    leading entries strictly greater than it.  A block keeping none is erased
    together with its unitary blocks (section sizes left stale); otherwise d
    and every unitary block of that sector are sliced and the section size of
    the last leg is updated.  d blocks are visited in reverse order."""
    d = d.copy()
    unitaries = [u.copy() for u in unitaries]
    for u in unitaries:   # the flat_map order the reference starts from
        u.blocks = dict(u.sorted_items())
    items = d.sorted_items()
    if not items:
        return d, unitaries, None
    vd = np.concatenate([b for _, b in items]).astype(np.float64)
    vd = -np.sort(-vd, kind="stable")
    last = compute_last_index(vd, tol, pow_, min_size, max_size)
    s = float(vd[last])
    thr = s - 2.0 * s * EPS64
    for kd_idx, db in reversed(items):
        kd = kd_idx[-1]
        k = 0
        while k < db.shape[0] and db[k] > thr:
            k += 1
        if k == 0:
            for u in unitaries:
                _remove_unit_blocks(u, kd)
            del d.blocks[kd_idx]
        else:
            d.blocks[kd_idx] = db[:k].copy()
            d.secs[0][kd] = k
            for u in unitaries:
                u.secs[-1][kd] = k
                for ui in list(u.blocks):
                    if ui[-1] == kd:
                        u.blocks[ui] = u.blocks[ui][..., :k].copy()
    return d, unitaries, thr


def svd_truncated(t, split, tol, min_size=1, max_size=2 ** 63 - 1, pow_=2.0):
    """Truncating ``svd(A, split, tol, min, max, pow)``
    (btensor_linalg.cpp:805-809)."""
    U, d, V = svd_split(t, split)
    d2, (U2, V2), _ = truncate_impl(d, [U, V], max_size, min_size, tol, pow_)
    return U2, d2, V2
