"""Host-side construction of the inputs of a DMRG solve: U(1) spin-1/2 MPOs
and random starting states, uploaded as device btensors.

The reference builds its Hamiltonians as dense MPOs (sources/models.cpp:30-68)
and infers the bond charges with ``guess_btensor`` (tests/2Dheisenberg.cpp:194-280);
its random starting state is ``random_bMPS`` (sources/MPT.cpp:311-494).  This
module does the same work for the solves this library serves:

* ``spin_mpo``: finite-state-machine MPO for
  H = sum_(i,j) in bonds J (Sx_i Sx_j + Sy_i Sy_j + delta Sz_i Sz_j) + sum_i h_i Sz_i
  on any bond list (chains, snake-ordered cylinders), index order
  (left bond, s, right bond, s') like the reference's dense MPOs.
* ``to_bmpo``: charge inference on every bond, channels regrouped by charge
  (a gauge permutation of each internal bond) so that each bond has as few
  sections as it has distinct charges, then one device btensor per site.
* ``random_bmps``: the block structure of ``random_bMPS`` (random physical
  strings in the target charge sector, bond sections from their partial
  charges) with seeded random blocks.

Everything here is host preparation; the solve itself runs on the device
(``quantit_amd.mps_prepare`` and ``quantit_amd.dmrg``).
"""
from __future__ import annotations

import numpy as np

from . import BTensor

# Physical basis: element 0 = spin up (2Sz = +1), element 1 = spin down
# (tests/DMRG_heisenberg_test.h:26 ordering).
UP_FIRST = ((1,), (-1,))
_SP = np.array([[0.0, 1.0], [0.0, 0.0]])   # S+ |down> = |up>
_SM = _SP.T.copy()
_SZ = np.diag([0.5, -0.5])
_ID = np.eye(2)


def read_mpo_text(text):
    """Dense MPO tensor (l, s, r, s') from the text format of the reference's
    cylinder fixtures (tests/2dHeisenberg/mpotensor_*.txt, read by
    tests/2Dheisenberg.cpp:130-175): one non-zero per line, "l s s' r value",
    1-based indices, optionally wrapped in a C++ raw-string literal.  Note
    the order of the indices on a line: the third number is the output
    physical index s' and the fourth the right bond.  Bond dimensions are the
    largest indices that occur."""
    body = text.strip()
    if body.startswith('R"('):
        body = body[3:]
    if body.endswith(')"'):
        body = body[:-2]
    rows = [ln.split() for ln in body.splitlines() if len(ln.split()) >= 5]
    idx = np.array([[int(x) for x in r[:4]] for r in rows], dtype=np.int64).reshape(-1, 4)
    val = np.array([float(r[4]) for r in rows], dtype=np.float64)
    if idx.size and idx.min() < 1:
        raise ValueError("indices are 1-based")
    lb, d, rb = (int(idx[:, 0].max()), max(2, int(idx[:, 1:3].max())), int(idx[:, 3].max())) if idx.size else (1, 2, 1)
    W = np.zeros((lb, d, rb, d))
    W[idx[:, 0] - 1, idx[:, 1] - 1, idx[:, 3] - 1, idx[:, 2] - 1] = val
    return W


def chain_bonds(L):
    return [(i, i + 1) for i in range(L - 1)]


def cylinder_bonds(width, length):
    """Bonds of a cylinder periodic in ``width`` and open in ``length``, sites
    numbered ring by ring in snake order: ring x holds sites x*width .. x*width
    + width-1, running forward on even rings and backward on odd ones."""
    def site(x, y):
        return x * width + (y if x % 2 == 0 else width - 1 - y)

    out = set()
    for x in range(length):
        for y in range(width):
            a = site(x, y)
            if width > 2 or y + 1 < width:   # ring bond (no double bond for width 2)
                b = site(x, (y + 1) % width)
                if a != b:
                    out.add((min(a, b), max(a, b)))
            if x + 1 < length:               # bond to the next ring
                b = site(x + 1, y)
                out.add((min(a, b), max(a, b)))
    return sorted(out)


def spin_mpo(L, bonds, J=1.0, delta=1.0, fields=None):
    """Dense MPO tensors W_i[l, s, r, s'] of the spin-1/2 XXZ model on
    ``bonds`` (pairs i < j) with optional fields h_i on Sz.

    The machine's states on the cut after site c are: 'done' (index 0), one
    triple (S+, S-, Sz pending) for every bond origin a <= c whose partner lies
    beyond c, and 'start' (last index).  The left edge keeps only 'start', the
    right edge only 'done'.  A bond (a, b) enters at a with the bare operator
    and closes at b with J/2 S-+ (for S+-) or J delta Sz (for Sz)."""
    bonds = sorted({(min(a, b), max(a, b)) for a, b in bonds})
    if any(a == b or a < 0 or b >= L for a, b in bonds):
        raise ValueError("bonds must join two distinct sites of the chain")
    fields = np.zeros(L) if fields is None else np.asarray(fields, dtype=np.float64)
    partners = [set() for _ in range(L)]
    for a, b in bonds:
        partners[a].add(b)
    kinds = (("p", _SP, _SM, 0.5 * J), ("m", _SM, _SP, 0.5 * J), ("z", _SZ, _SZ, J * delta))

    def states(cut):
        if cut < 0:
            return ["start"]
        if cut >= L - 1:
            return ["done"]
        pend = sorted(a for a in range(cut + 1) if any(b > cut for b in partners[a]))
        return ["done"] + [(a, k[0]) for a in pend for k in kinds] + ["start"]

    out = []
    for i in range(L):
        left, right = states(i - 1), states(i)
        li = {s: n for n, s in enumerate(left)}
        ri = {s: n for n, s in enumerate(right)}
        W = np.zeros((len(left), 2, len(right), 2))

        def put(src, dst, op):
            if src in li and dst in ri:
                W[li[src], :, ri[dst], :] += op

        put("done", "done", _ID)
        put("start", "start", _ID)
        if fields[i] != 0.0:
            put("start", "done", fields[i] * _SZ)
        for name, open_op, close_op, coef in kinds:
            put("start", (i, name), open_op)
            for a in range(i):
                if (a, name) in li:
                    if (a, name) in ri:
                        put((a, name), (a, name), _ID)
                    if i in partners[a]:
                        put((a, name), "done", coef * close_op)
        out.append(W)
    return out


def _charge_runs(charges):
    secs, cv = [], []
    for c in charges:
        if cv and cv[-1] == c:
            secs[-1] += 1
        else:
            secs.append(1)
            cv.append(c)
    return secs, cv


def to_bmpo(ctx, dense, phys=UP_FIRST, cutoff=1e-14):
    """Device bMPO (one btensor per site) from dense MPO tensors."""
    return [BTensor.from_parts(ctx, **p) for p in bmpo_parts(dense, phys, cutoff)]


def random_bmps(ctx, L, bond, q=0, seed=0, phys=UP_FIRST):
    """Device random bMPS (see ``random_bmps_parts``)."""
    return [BTensor.from_parts(ctx, **p) for p in random_bmps_parts(L, bond, q, seed, phys)]


def mpo_phys_legs(mpo):
    """Physical legs of the states an MPO acts on: the inverse of each
    site's last leg, ``hamil[i].shape_from({0,0,0,-1}).inverse_cvals()``
    (sources/MPT.cpp:480-488), as lists of (section size, charge)."""
    out = []
    for W in mpo:
        groups, secs, cvals, _ = W.structure()
        out.append([(int(n), _q_inv(groups, tuple(c))) for n, c in zip(secs[3], cvals[3])])
    return out, tuple(groups)


def bmpo_parts(dense, phys=UP_FIRST, cutoff=1e-14):
    """Block structure and data of a U(1) bMPO from dense MPO tensors (l, s, r, s').
    Returns one dict (groups, secs, cvals, sel, blocks) per site.

    Every element above ``cutoff`` fixes the charge of its right-bond channel,
    c_r = c_s' - c_l - c_s (the block is then neutral).  Channels with no
    element on either side are dropped; the rest are ordered by charge.
    Raises ValueError for a non-conserving MPO."""
    dense = [np.asarray(W, dtype=np.float64).copy() for W in dense]
    L = len(dense)
    for i in range(L - 1):   # drop channels that are empty on either side
        A, B = dense[i], dense[i + 1]
        keep = [r for r in range(A.shape[2])
                if np.abs(A[:, :, r, :]).max() > cutoff and np.abs(B[r]).max() > cutoff]
        dense[i], dense[i + 1] = A[:, :, keep, :], B[keep]
    ps = [c[0] for c in phys]
    left = [0] * dense[0].shape[0]    # charges (2Sz) of the left-bond channels
    out = []
    for i, W in enumerate(dense):
        rch = [None] * W.shape[2]
        for l, s, r, sp in zip(*np.nonzero(np.abs(W) > cutoff)):
            c = ps[sp] - left[l] - ps[s]
            if rch[r] is None:
                rch[r] = c
            elif rch[r] != c:
                raise ValueError(f"site {i}: the MPO does not conserve Sz on bond channel {r}")
        if any(c is None for c in rch):
            raise ValueError(f"site {i}: empty bond channel")
        order = sorted(range(len(rch)), key=lambda r: (rch[r], r))
        W = W[:, :, order, :]
        if i + 1 < L:
            dense[i + 1] = dense[i + 1][order]
        rc = [rch[r] for r in order]
        lsecs, lcv = _charge_runs(left)
        rsecs, rcv = _charge_runs(rc)
        loff = np.concatenate([[0], np.cumsum(lsecs)])
        roff = np.concatenate([[0], np.cumsum(rsecs)])
        blocks = {}
        for a, ca in enumerate(lcv):
            for s in range(2):
                for b, cb in enumerate(rcv):
                    for sp in range(2):
                        if ca + ps[s] + cb - ps[sp] != 0:
                            continue
                        blk = W[loff[a]:loff[a + 1], s:s + 1, roff[b]:roff[b + 1], sp:sp + 1]
                        if np.abs(blk).max() > 0.0:
                            blocks[(a, s, b, sp)] = blk
        out.append(dict(groups=(0,), secs=[lsecs, [1, 1], rsecs, [1, 1]],
                        cvals=[[(c,) for c in lcv], [(p,) for p in ps], [(c,) for c in rcv], [(-p,) for p in ps]],
                        sel=(0,), blocks=blocks))
        left = [-c for c in rc]
    return out


def _q_add(groups, a, b):
    """Sum of two charges: Z components add, C<N> components add mod N
    (stored in [0, N)); int16 wrap-around as the stored values."""
    out = []
    for g, x, y in zip(groups, a, b):
        v = x + y
        v = v % g if g else ((v + 32768) % 65536) - 32768
        out.append(int(v))
    return tuple(out)


def _q_inv(groups, a):
    return tuple(int((-x) % g) if g else int(((-x + 32768) % 65536) - 32768) for g, x in zip(groups, a))


def _q_dist2(groups, a, b):
    """Squared distance between charges (the repair metric of
    generate_random_string): the difference per component, taken as the
    shorter way round for C<N>."""
    t = 0
    for g, x, y in zip(groups, a, b):
        d = x - y
        if g:
            d %= g
            d = min(d, g - d)
        t += d * d
    return t


def random_bmps_general(bond, phys, q, seed=0, groups=(0,)):
    """Random bMPS structure and blocks of ``random_bMPS_impl``
    (sources/MPT.cpp:423-462) for any charge groups and physical legs.

    phys[i] is site i's physical leg as a list of (section size, charge),
    i.e. the inverse of the MPO's last leg (``random_bMPS(bond, hamil, q)``,
    sources/MPT.cpp:478-493).  ``bond`` strings of physical sections are drawn
    uniformly and repaired in one pass toward total charge q
    (``generate_random_string``, :311-371: a site's section is replaced only
    by one strictly closer to q); the right bond of site i has one section
    per distinct partial charge q - sum_{t<=i} c(s_t), sized by its
    multiplicity (``make_right_side``, :387-421); site 0 carries the selection
    rule q, the others the neutral one; the last right bond keeps one element.
    Blocks are uniform in [0, 1) from numpy's PCG64 seeded with ``seed``, each
    site normalised."""
    L = len(phys)
    q = tuple(q)
    neutral = tuple(0 for _ in groups)
    rng = np.random.default_rng(seed)
    counts = [len(p) for p in phys]
    if len(set(counts)) == 1:
        strings = rng.integers(0, counts[0], size=(bond, L))
    else:
        strings = np.stack([rng.integers(0, n, size=bond) for n in counts], axis=1)
    pc = [[c for _, c in p] for p in phys]
    for j in range(bond):
        total = neutral
        for i in range(L):
            total = _q_add(groups, total, pc[i][strings[j, i]])
        cur = _q_dist2(groups, total, q)
        for i in range(L):
            if cur == 0:
                break
            rest = _q_add(groups, total, _q_inv(groups, pc[i][strings[j, i]]))
            best = int(strings[j, i])
            for s_ in range(counts[i]):
                nd = _q_dist2(groups, _q_add(groups, rest, pc[i][s_]), q)
                if nd < cur:
                    best, cur = s_, nd
            strings[j, i] = best
            total = _q_add(groups, rest, pc[i][best])
        if cur != 0:
            raise ValueError("charge sector unreachable with these physical legs")
    acc = [q] * bond
    lsecs, lcv = [1], [neutral]
    out = []
    for i in range(L):
        acc = [_q_add(groups, a, _q_inv(groups, pc[i][strings[j, i]])) for j, a in enumerate(acc)]
        rsecs, rcv = _charge_runs(sorted(acc))
        if i == L - 1:
            rsecs, rcv = [1], rcv[:1]
        sel = q if i == 0 else neutral
        psz = [n for n, _ in phys[i]]
        blocks = {}
        for a, ca in enumerate(lcv):
            for s_ in range(counts[i]):
                for b, cb in enumerate(rcv):
                    if _q_add(groups, _q_add(groups, ca, pc[i][s_]), cb) == sel:
                        blocks[(a, s_, b)] = rng.random((lsecs[a], psz[s_], rsecs[b]))
        nrm = np.sqrt(sum(float(np.sum(x * x)) for x in blocks.values()))
        blocks = {k: v / nrm for k, v in blocks.items()}
        out.append(dict(groups=tuple(groups), secs=[lsecs, psz, rsecs], cvals=[list(lcv), list(pc[i]), list(rcv)],
                        sel=sel, blocks=blocks))
        lsecs, lcv = rsecs, [_q_inv(groups, c) for c in rcv]
    return out


def random_bmps_parts(L, bond, q=0, seed=0, phys=UP_FIRST):
    """Random U(1) bMPS of L spin-1/2 sites in the sector of total 2Sz = q:
    ``random_bmps_general`` with the physical leg ``phys`` (one state per
    section) on every site."""
    return random_bmps_general(bond, [[(1, c) for c in phys]] * L, (q,), seed, (0,))
