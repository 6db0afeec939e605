"""numpy restatement of QuantiT's two-site DMRG driver and its inputs.

TEST INFRASTRUCTURE: the oracle for the DMRG path.  Imported only by tests,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``.

Pinning (tests/test_oracle_dmrg.py):
  * eig2x2            -- known answer of include/dmrg.h:115-139 (constructed
                         2x2 with a random angle).
  * matvec / <H>      -- <psi|H|psi> via ``contract`` == via W12 == via
                         ``hamil2site_times_state`` (include/dmrg.h:141-290),
                         and == dense <psi|H|psi> from the dense MPO.
  * Heisenberg MPO    -- exact int8 8x8 3-site matrix (include/models.h:62-73).
  * dmrg energies     -- exact diagonalisation for L <= 12 chains and 4x2/4x3
                         cylinders; XX chain free-fermion closed form.
  * cylinder builder  -- <psi|H|psi> equal to the 4x8 fixture MPO
                         (tests/2dHeisenberg/mpotensor_*.txt).
  * lanczos_ground    -- (this build's k-step eigensolver, no reference
                         counterpart) with two vectors it equals the one-step
                         update (eig2x2 of one_step_lanczos); its Ritz value
                         equals the lowest eigenvalue of the dense two-site
                         Hamiltonian built column by column from the matvec;
                         DMRG with it reaches exact-diagonalisation energies.
The random initial MPS follows the structure of ``random_bMPS``
(sources/MPT.cpp:311-462) but draws from numpy's PCG64, not from std::mt19937
and torch's CPU generator, so it is this project's own seeded input.
"""
from __future__ import annotations

import numpy as np

from .btensor_oracle import (BTensor, add, eye_like, full_dot, mul, ones_like, q_add, q_inv, q_neutral,
                             rand_like, scal, shape_product, svd_split, svd_truncated, tensordot)


# ---------------------------------------------------------------------------
# dense MPO builders (inputs)
# ---------------------------------------------------------------------------
def pauli():
    """``pauli()`` (sources/operators.cpp:59-88): Sx, iSy, Sz, lo, id."""
    sx = np.array([[0, 1], [1, 0]], dtype=np.float64)
    isy = np.array([[0, 1], [-1, 0]], dtype=np.float64)
    sz = np.array([[1, 0], [0, -1]], dtype=np.float64)
    lo = np.array([[0, 1], [0, 0]], dtype=np.float64)
    idm = np.eye(2)
    return sx, isy, sz, lo, idm


def heisenberg_impl(Jp, L):
    """``details::Heisenberg_impl`` (sources/models.cpp:38-64): dense MPO with
    index order (left bond, s, right bond, s'), bond 5, row 4 at the left edge
    and column 0 at the right edge."""
    sx, isy, sz, lo, idm = pauli()
    hi = lo.T
    W = np.zeros((5, 2, 5, 2))
    W[0, :, 0, :] = idm
    W[1, :, 0, :] = lo
    W[2, :, 0, :] = hi
    W[3, :, 0, :] = sz
    W[4, :, 1, :] = 2 * Jp * hi
    W[4, :, 2, :] = 2 * Jp * lo
    W[4, :, 3, :] = Jp * sz
    W[4, :, 4, :] = idm
    mpo = [W.copy() for _ in range(L)]
    mpo[0] = mpo[0][4:5]
    mpo[-1] = mpo[-1][:, :, 0:1, :]
    return mpo


def heisenberg_dense(J, L):
    """``Heisenberg(J, L)`` = Heisenberg_impl(-J/4) (sources/models.cpp:65-68).
    J = -1 gives H = sum_i S_i . S_{i+1} (antiferromagnet)."""
    return heisenberg_impl(-J / 4.0, L)


def xxz_dense(delta, L, J=1.0):
    """H = J sum_i (Sx Sx + Sy Sy + delta Sz Sz), same operator layout as
    ``Heisenberg_impl`` (this project's own builder for the XXZ requests)."""
    sx, isy, sz, lo, idm = pauli()
    hi = lo.T
    W = np.zeros((5, 2, 5, 2))
    W[0, :, 0, :] = idm
    W[1, :, 0, :] = lo
    W[2, :, 0, :] = hi
    W[3, :, 0, :] = sz
    W[4, :, 1, :] = 0.5 * J * hi
    W[4, :, 2, :] = 0.5 * J * lo
    W[4, :, 3, :] = 0.25 * J * delta * sz
    W[4, :, 4, :] = idm
    mpo = [W.copy() for _ in range(L)]
    mpo[0] = mpo[0][4:5]
    mpo[-1] = mpo[-1][:, :, 0:1, :]
    return mpo


def xxz_field_dense(delta, fields, J=1.0):
    """H = J sum_i (Sx Sx + Sy Sy + delta Sz Sz) + sum_i h_i Sz_i: the XXZ
    layout above with the on-site field on the start -> done channel.  Random
    fields remove every symmetry but U(1), so Schmidt spectra carry no exact
    multiplets (a test input with tie-free truncations)."""
    L = len(fields)
    mpo = xxz_dense(delta, L, J)
    sz = pauli()[2]
    for i, h in enumerate(fields):
        row = 0 if i == 0 else 4
        mpo[i][row, :, 0, :] = 0.5 * h * sz
    return mpo


def cylinder_bonds(width, length):
    """Nearest-neighbour bonds of a width-periodic, length-open cylinder in
    snake order: site (x, y) -> x*width + y for even x, x*width + width-1-y
    for odd x."""
    def site(x, y):
        return x * width + (y if x % 2 == 0 else width - 1 - y)
    bonds = set()
    for x in range(length):
        for y in range(width):
            a = site(x, y)
            if width > 2 or y + 1 < width:
                b = site(x, (y + 1) % width)
                if a != b:
                    bonds.add((min(a, b), max(a, b)))
            if x + 1 < length:
                b = site(x + 1, y)
                bonds.add((min(a, b), max(a, b)))
    return sorted(bonds)


def longrange_heisenberg_dense(N, bonds, J=1.0, delta=1.0):
    """Finite-state-automaton MPO for H = J sum_(i,j) in bonds
    (Sx_i Sx_j + Sy_i Sy_j + delta Sz_i Sz_j); index order (l, s, r, s'),
    channel 0 = done, last channel = start.  Channels: one (S+, S-, Sz)
    triple per open bond origin that is still waiting for its partner."""
    sx, isy, sz, lo, idm = pauli()
    hi = lo.T
    # Sx Sx + Sy Sy = (S+ S- + S- S+)/2 with S+- = hi/lo (unit entries); Sz = sz/2
    ops = {"p": hi, "m": lo, "z": 0.5 * sz}
    partner = {"p": "m", "m": "p", "z": "z"}
    coef = {"p": 0.5 * J, "m": 0.5 * J, "z": J * delta}
    Ws = []
    # open channels after site i: origins a <= i with a partner j > i
    open_after = []
    for i in range(N):
        open_after.append(sorted({a for (a, b) in bonds if a <= i < b}))
    for i in range(N):
        left = ["done"] + [(a, k) for a in (open_after[i - 1] if i > 0 else []) for k in "pmz"] + ["start"]
        right = ["done"] + [(a, k) for a in (open_after[i] if i < N - 1 else []) for k in "pmz"] + ["start"]
        if i == 0:
            left = ["start"]
        if i == N - 1:
            right = ["done"]
        li = {c: n for n, c in enumerate(left)}
        ri = {c: n for n, c in enumerate(right)}
        W = np.zeros((len(left), 2, len(right), 2))
        if "done" in li and "done" in ri:
            W[li["done"], :, ri["done"], :] = idm
        if "start" in li and "start" in ri:
            W[li["start"], :, ri["start"], :] = idm
        # carry open channels through
        for c in left:
            if isinstance(c, tuple) and c in ri:
                W[li[c], :, ri[c], :] = idm
        # open new channels at i
        for k in "pmz":
            if (i, k) in ri and "start" in li:
                W[li["start"], :, ri[(i, k)], :] = ops[k]
        # close channels whose partner is i
        for c in left:
            if isinstance(c, tuple):
                a, k = c
                if (a, i) in bonds and "done" in ri:
                    W[li[c], :, ri["done"], :] += coef[k] * ops[partner[k]]
        Ws.append(W)
    return Ws


def mpo_to_dense_matrix(mpo):
    """Contract a dense MPO (l, s, r, s') into the d^N x d^N matrix
    <s_1..s_N| H |s'_1..s'_N> (row index = bra string, first site slowest)."""
    T = np.transpose(mpo[0][0], (1, 0, 2))  # (r, s, s')
    for W in mpo[1:]:
        out = np.einsum("rab,rscd->casbd", T, W)  # (r2, A, s, B, s')
        T = out.reshape(W.shape[2], T.shape[1] * W.shape[1], T.shape[2] * W.shape[3])
    return T[0]


def parse_mpo_fixture(text):
    """``string2structure`` + ``make_tensor`` (tests/2Dheisenberg.cpp:130-175):
    lines "l s s' r coeff", 1-based; stored at (l-1, s-1, r-1, s'-1) -- the
    reference's parser writes index ind[2] (the third number) to the last
    axis and ind[3] to the right bond."""
    text = text.strip()
    if text.startswith('R"('):
        text = text[3:]
    if text.endswith(')"'):
        text = text[:-2]
    entries = []
    lb = rb = 0
    for line in text.strip().splitlines():
        parts = line.split()
        if len(parts) < 5:
            continue
        l, s, sp, r = (int(x) for x in parts[:4])
        c = float(parts[4])
        lb = max(lb, l)
        rb = max(rb, r)
        entries.append((l, s, sp, r, c))
    W = np.zeros((lb, 2, rb, 2))
    for l, s, sp, r, c in entries:
        W[l - 1, s - 1, r - 1, sp - 1] = c
    return W


# ---------------------------------------------------------------------------
# charge inference: dense MPO -> block MPO
# ---------------------------------------------------------------------------
def guess_bmpo(dense_mpo, phys_charges, groups=(0,), cutoff=1e-4):
    """Site-by-site charge inference like ``guess_btensor``
    (tests/2Dheisenberg.cpp:194-280) with the driver's wiring (:303-316):
    left bond known, phys leg s with ``phys_charges``, s' with their inverse;
    every non-negligible element fixes the right-bond charge
    c_r = -(c_l + c_s - c_s').  Right-bond sections are maximal runs of
    consecutive bond indices with equal charge.  A bond channel with no
    non-negligible element cannot be assigned a charge and is rejected."""
    n = q_neutral(groups)
    left_secs = [1]
    left_c = [n]
    out = []
    phys = list(phys_charges)
    physdag = [q_inv(groups, c) for c in phys]
    for site, W in enumerate(dense_mpo):
        lb, d, rb, d2 = W.shape
        lcharge_of = []
        for s, c in zip(left_secs, left_c):
            lcharge_of += [c] * s
        rch = [None] * rb
        for l, s, r, sp in zip(*np.nonzero(np.abs(W) > cutoff)):
            c = q_add(groups, q_add(groups, lcharge_of[l], phys[s]), physdag[sp])
            c = q_inv(groups, c)
            if rch[r] is None:
                rch[r] = c
            elif rch[r] != c:
                raise ValueError(f"site {site}: non-conserving MPO at right bond {r}")
        if any(c is None for c in rch):
            raise ValueError(f"site {site}: empty right-bond channel, drop it first")
        secs, cv = [], []
        for c in rch:
            if cv and cv[-1] == c:
                secs[-1] += 1
            else:
                secs.append(1)
                cv.append(c)
        shape = BTensor(groups, [left_secs, [1] * d, secs, [1] * d2], [left_c, phys, cv, physdag], n)
        out.append(shape.fill_from_dense(W, cutoff=0.0))
        left_secs, left_c = secs, [q_inv(groups, c) for c in cv]
    return out


def drop_zero_channels(dense_mpo, tol=0.0):
    """Remove bond channels that carry no element on either side (the all-zero
    right-bond channel 0 of the first Heisenberg site, for example)."""
    mpo = [W.copy() for W in dense_mpo]
    for i in range(len(mpo) - 1):
        A, B = mpo[i], mpo[i + 1]
        keep = [r for r in range(A.shape[2]) if np.any(np.abs(A[:, :, r, :]) > tol) and np.any(np.abs(B[r]) > tol)]
        mpo[i] = A[:, :, keep, :]
        mpo[i + 1] = B[keep]
    return mpo


def u1_spin_half_phys(up_first=True):
    """Physical leg charges in 2Sz units: element 0 = spin up (+1) when
    ``up_first`` (tests/DMRG_heisenberg_test.h:26 order)."""
    return [(1,), (-1,)] if up_first else [(-1,), (1,)]


# ---------------------------------------------------------------------------
# MPS helpers  (sources/MPT.cpp)
# ---------------------------------------------------------------------------
def random_bmps(bond, bmpo, q, seed, dist="uniform"):
    """Structure of ``random_bMPS_impl`` (sources/MPT.cpp:423-494): ``bond``
    random physical-charge strings satisfying the total charge ``q``
    (``generate_random_string`` :311-371, greedy one-pass repair by squared
    charge distance), bond sections = sorted distinct partial charges with
    multiplicities (``make_right_side`` :387-421), all allowed blocks random
    and each site tensor normalised; the last right leg keeps element 0."""
    rng = np.random.default_rng(seed)
    g = bmpo[0].groups
    n = q_neutral(g)
    L = len(bmpo)
    phys = [[q_inv(g, c) for c in W.cvals[3]] for W in bmpo]   # hamil[i].shape_from({0,0,0,-1}).inverse
    phys_secs = [list(W.secs[3]) for W in bmpo]

    def d2(a, b):
        return sum((x - y) ** 2 for x, y in zip(a, b))

    strings = []
    for _ in range(bond):
        st = [int(rng.integers(0, len(phys[i]))) for i in range(L)]
        S = n
        for i in range(L):
            S = q_add(g, S, phys[i][st[i]])
        cur = d2(S, q)
        if cur:
            for i in range(L):
                if cur == 0:
                    break
                S = q_add(g, S, q_inv(g, phys[i][st[i]]))
                best = st[i]
                for j in range(len(phys[i])):
                    nd = d2(q_add(g, S, phys[i][j]), q)
                    if nd < cur:
                        best, cur = j, nd
                st[i] = best
                S = q_add(g, S, phys[i][best])
        if cur:
            raise ValueError("cannot satisfy the sum rule")
        strings.append(st)
    acc = [q for _ in range(bond)]
    left_secs, left_c = [1], [n]
    out = []
    for i in range(L):
        for j in range(bond):
            acc[j] = q_add(g, acc[j], q_inv(g, phys[i][strings[j][i]]))
        srt = sorted(acc)
        secs, cv = [], []
        for c in srt:
            if cv and cv[-1] == c:
                secs[-1] += 1
            else:
                secs.append(1)
                cv.append(c)
        sel = q if i == 0 else n
        shape = BTensor(g, [left_secs, phys_secs[i], secs], [left_c, phys[i], cv], sel)
        t = rand_like(shape, rng, dist)
        nrm = np.sqrt(sum(float(np.sum(b * b)) for b in t.blocks.values()))
        t = scal(t, 1.0 / nrm)
        out.append(t)
        left_secs, left_c = secs, [q_inv(g, c) for c in cv]
    # basic_create_view({-1,-1,0}, true): last right leg restricted to element 0
    last = out[-1]
    r_secs = last.secs[2]
    new = BTensor(g, [last.secs[0], last.secs[1], [1]], [last.cvals[0], last.cvals[1], [last.cvals[2][0]]], last.sel)
    for idx, b in last.blocks.items():
        if idx[2] == 0:
            new.blocks[(idx[0], idx[1], 0)] = b[:, :, :1].copy()
    out[-1] = new
    return out


def contract_norm(a, b):
    """``contract(bMPS, bMPS)`` = <b|a> (sources/MPT.cpp:268-292)."""
    left = eye_like(shape_product(a[0].shape_from([-1, 0, 0]).inverse_cvals_(), b[0].shape_from([-1, 0, 0])))
    right = eye_like(shape_product(a[-1].shape_from([0, 0, -1]).inverse_cvals_(), b[-1].shape_from([0, 0, -1])))
    for x, y in zip(a, b):
        left = tensordot(left, x, [0], [0])
        left = tensordot(left, y.conj(), [0, 1], [0, 1])
    res = tensordot(left, right, [0, 1], [0, 1])
    return float(next(iter(res.blocks.values()))) if res.blocks else 0.0


def edge_shape_prep(t, dim):
    """``details::edge_shape_prep`` (sources/dmrg.cpp:334-342).  The two
    arguments of the final ``shape_from`` are evaluated right to left (g++
    order): the rank-0 shape is taken before ``inverse_cvals_``."""
    vec = [0] * t.rank
    vec[dim] = -1
    S = t.shape_from(vec)
    S.sel = q_neutral(t.groups)
    s0 = S.shape_from([0])
    S.inverse_cvals_()
    return shape_product(S, s0)


def contract_energy(a, b, mpo):
    """``contract(bMPS, bMPS, bMPO)`` = <b|H|a> (sources/MPT.cpp:211-231)."""
    left = eye_like(shape_product(a[0].shape_from([-1, 0, 0]).inverse_cvals_(), b[0].shape_from([-1, 0, 0])))
    right = eye_like(shape_product(a[-1].shape_from([0, 0, -1]).inverse_cvals_(), b[-1].shape_from([0, 0, -1])))
    left = tensordot(left, ones_like(edge_shape_prep(mpo[0], 0)), [], []).permute([0, 2, 1])
    right = tensordot(right, ones_like(edge_shape_prep(mpo[-1], 2)), [], []).permute([0, 2, 1])
    for x, W, y in zip(a, mpo, b):
        left = tensordot(left, x, [0], [0])
        left = tensordot(left, W, [0, 2], [0, 3])
        left = tensordot(left, y.conj(), [0, 2], [0, 1])
    res = tensordot(left, right, [0, 1, 2], [0, 1, 2])
    return float(next(iter(res.blocks.values()))) if res.blocks else 0.0


def move_oc(mps, oc, target):
    """``bMPS::move_oc`` (sources/MPT.cpp:75-111) with untruncated block SVDs."""
    mps = list(mps)
    while target < oc:
        U, d, V = svd_split(mps[oc], 1)
        mps[oc] = V.conj().permute([2, 0, 1])
        ud = mul(U, d)
        mps[oc - 1] = tensordot(mps[oc - 1], ud, [2], [0])
        oc -= 1
    while target > oc:
        U, d, V = svd_split(mps[oc], 2)
        mps[oc] = U
        dv = mul(V, d).conj()
        mps[oc + 1] = tensordot(dv, mps[oc + 1], [0], [0])
        oc += 1
    return mps, oc


def prepare_state(mps):
    """Driver preparation (tests/DMRG_heisenberg_test.h:34-37): normalise,
    move the orthogonality centre to the end and back to 0."""
    nrm = contract_norm(mps, mps)
    mps = list(mps)
    mps[0] = scal(mps[0], 1.0 / np.sqrt(nrm))
    mps, oc = move_oc(mps, 0, len(mps) - 1)
    mps, oc = move_oc(mps, oc, 0)
    return mps


# ---------------------------------------------------------------------------
# DMRG  (sources/dmrg.cpp)
# ---------------------------------------------------------------------------
def compute_left_env(W, Y, Lenv):
    """``compute_left_env_impl`` (dmrg.cpp:424-459)."""
    out = tensordot(Lenv, Y, [0], [0])
    out = tensordot(out, W, [0, 2], [0, 3])
    return tensordot(out, Y.conj(), [0, 2], [0, 1])


def compute_right_env(W, Y, Renv):
    """``compute_right_env_impl`` (dmrg.cpp:468-493)."""
    out = tensordot(Renv, Y, [0], [2])
    out = tensordot(out, W, [0, 3], [2, 3])
    return tensordot(out, Y.conj(), [3, 0], [1, 2])


def two_site_hamil(mpo):
    """``compute_2sitesHamil_impl`` (dmrg.cpp:503-515)."""
    return [tensordot(mpo[i], mpo[i + 1], [2], [0]).permute([0, 1, 3, 4, 2, 5]) for i in range(len(mpo) - 1)]


def matvec(theta, W12, Lenv, Renv):
    """``hamil2site_times_state_impl`` (dmrg.cpp:520-531)."""
    out = tensordot(Lenv, theta, [0], [0])
    out = tensordot(out, W12, [0, 2, 3], [0, 4, 5])
    return tensordot(out, Renv, [1, 4], [0, 1])


def _scalar(t):
    return float(next(iter(t.blocks.values()))) if t.blocks else 0.0


def eig2x2(a0, a1, b):
    """``eig2x2Mat_impl`` (dmrg.cpp:543-572): lowest eigenpair of
    [[a0, b], [b, a1]] in the closed form of arXiv:1908.03795."""
    crit = np.sqrt((a0 - a1) ** 2 + 4 * (b * b))
    E0 = (a0 + a1 - crit) / 2
    delt = E0 - a1
    with np.errstate(invalid="ignore", divide="ignore"):
        o = np.sqrt(delt / (-crit))
        n = (b * o) / delt
    zero_o = (o + E0) == E0 or np.isnan(o)
    if zero_o:
        n = 1.0
        o = 0.0
    if np.isnan(o) or np.isnan(n):
        raise ArithmeticError("nan found in output tensor")
    return float(E0), float(o), float(n)


def one_step_lanczos(theta, W12, Lenv, Renv):
    """``one_step_lanczos_impl`` (dmrg.cpp:584-608)."""
    psi = matvec(theta, W12, Lenv, Renv)
    a0t = tensordot(psi, theta.conj(), [0, 1, 2, 3], [0, 1, 2, 3])
    psi = add(psi, mul(theta, a0t), -1.0)
    bt = tensordot(psi, psi.conj(), [0, 1, 2, 3], [0, 1, 2, 3])
    b = np.sqrt(_scalar(bt))
    if abs(b) >= 1e-15:
        psi = scal(psi, 1.0 / b)
    a1t = tensordot(psi.conj(), matvec(psi, W12, Lenv, Renv), [0, 1, 2, 3], [0, 1, 2, 3])
    return psi, _scalar(a0t), _scalar(a1t), float(b)


def lanczos_ground(theta, W12, Lenv, Renv, max_krylov, tol):
    """k-step Lanczos with full reorthogonalisation on the two-site matvec
    (the production eigensolver of this build; the reference has only the
    one-step form, to which this reduces for max_krylov = 2).

    v0 = theta/|theta|; step j: w = H v_j, two classical Gram-Schmidt passes
    against v_0..v_j (coefficients c, c'), alpha_j = c_j + c'_j,
    beta_j = |w|; the lowest eigenpair (E, y) of the (j+1)-square tridiagonal
    (numpy's dense symmetric eigensolver), y[0] >= 0; stop when
    beta_j |y_j| <= tol |E|, beta_j <= 1e-15 |E| or j+1 = max_krylov,
    else v_{j+1} = w / beta_j.  Returns (E, sum_i y_i v_i, vectors used)."""
    nrm = np.sqrt(full_dot(theta, theta)[0])
    V = [scal(theta, 1.0 / nrm)]
    alpha, beta = [], []
    E, y = None, None
    for j in range(max_krylov):
        w = matvec(V[j], W12, Lenv, Renv)
        c = [full_dot(w, v)[0] for v in V]
        for ci, v in zip(c, V):
            w = add(w, v, -ci)
        c2 = [full_dot(w, v)[0] for v in V]
        for ci, v in zip(c2, V):
            w = add(w, v, -ci)
        alpha.append(c[j] + c2[j])
        b = float(np.sqrt(full_dot(w, w)[0]))
        beta.append(b)
        m = j + 1
        T = np.diag(alpha) + np.diag(beta[:m - 1], 1) + np.diag(beta[:m - 1], -1)
        vals, vecs = np.linalg.eigh(T)
        E, y = float(vals[0]), vecs[:, 0]
        if y[0] < 0:
            y = -y
        if b * abs(y[-1]) <= tol * abs(E) or not (b > 1e-15 * abs(E)) or m == max_krylov:
            break
        V.append(scal(w, 1.0 / b))
    out = scal(V[0], float(y[0]))
    for yi, v in zip(y[1:], V[1:]):
        out = add(out, v, float(yi))
    return E, out, len(y)


def two_sites_update(theta, W12, Lenv, Renv):
    """``two_sites_update_impl`` (dmrg.cpp:623-638)."""
    psi, a0, a1, b = one_step_lanczos(theta, W12, Lenv, Renv)
    E, o, n = eig2x2(a0, a1, b)
    new = add(scal(theta, o), scal(psi, n), 1.0)
    return E, new, (a0, a1, b)


def generate_env(mpo, mps, oc):
    """``generate_env_impl`` (dmrg.cpp:370-409); returns a dict keyed -1..L."""
    L = len(mpo)
    Ls = edge_shape_prep(mps[0], 0)
    LH = edge_shape_prep(mpo[0], 0)
    env = {-1: ones_like(shape_product(Ls, LH, Ls.inverse_cvals()))}
    Rs = edge_shape_prep(mps[-1], 2)
    env[L] = ones_like(shape_product(Rs, edge_shape_prep(mpo[-1], 2), Rs.inverse_cvals()))
    for i in range(oc):
        env[i] = compute_left_env(mpo[i], mps[i], env[i - 1])
    for i in range(L - 1, oc, -1):
        env[i] = compute_right_env(mpo[i], mps[i], env[i + 1])
    return env


class DmrgOptions:
    """``dmrg_options`` (include/dmrg_options.h:15-57) defaults."""

    def __init__(self, cutoff=1e-6, convergence_criterion=1e-5, maximum_bond=2 ** 63 - 1,
                 minimum_bond=4, maximum_iterations=1000, max_krylov=0, lanczos_tol=1e-12):
        self.cutoff = cutoff
        self.convergence_criterion = convergence_criterion
        self.maximum_bond = maximum_bond
        self.minimum_bond = minimum_bond
        self.maximum_iterations = maximum_iterations
        # 0: the reference's one-step Lanczos; >= 2: k-step Lanczos
        self.max_krylov = max_krylov
        self.lanczos_tol = lanczos_tol


def dmrg_update(mpo, W12, mps, env, oc, step, opts, trace=None):
    """``dmrg_2sites_update::operator()`` (dmrg.cpp:163-206)."""
    theta = tensordot(mps[oc], mps[oc + 1], [2], [0])
    if getattr(opts, "max_krylov", 0) >= 2:
        E, theta2, nvec = lanczos_ground(theta, W12[oc], env[oc - 1], env[oc + 2], opts.max_krylov,
                                         opts.lanczos_tol)
        coeffs = (None, None, None)
    else:
        E, theta2, coeffs = two_sites_update(theta, W12[oc], env[oc - 1], env[oc + 2])
    u, d, v = svd_truncated(theta2, 2, opts.cutoff, opts.minimum_bond, opts.maximum_bond)
    nrm = np.sqrt(sum(float(np.sum(b * b)) for b in d.blocks.values()))
    d = scal(d, 1.0 / nrm)
    if step == 1:
        mps[oc] = u
        mps[oc + 1] = mul(v, d).conj().permute([2, 0, 1])
        env[oc] = compute_left_env(mpo[oc], mps[oc], env[oc - 1])
    else:
        mps[oc] = mul(u, d)
        mps[oc + 1] = v.conj().permute([2, 0, 1])
        env[oc + 1] = compute_right_env(mpo[oc + 1], mps[oc + 1], env[oc + 2])
    if trace is not None:
        trace.append(dict(oc=oc, step=step, E=E, a0=coeffs[0], a1=coeffs[1], b=coeffs[2],
                          kept=[int(x.shape[0]) for _, x in d.sorted_items()]))
    return E


def dmrg(mpo, mps, opts, log=None, trace=None, oc=0):
    """``dmrg(bMPO&, bMPS&, options, logger)`` + ``details::dmrg_impl``
    (dmrg.cpp:92-100, 219-273) with the orthogonality centre at ``oc``:
    sweeps start toward the far end (a centre on the last site starts one
    site to the left, :232-236) and the centre is moved back at the end
    (:257-266).  Returns (E, mps, per-sweep energies)."""
    L = len(mpo)
    mps = list(mps)
    env = generate_env(mpo, mps, oc)
    W12 = two_site_hamil(mpo)
    n_step = len(W12) - 1 + (len(W12) == 1)
    init_pos = oc
    step = 1 if oc == 0 else -1
    if len(W12) == 1:
        step = 0
    if oc == L - 1:
        init_pos -= 1
        oc -= 1
    E0 = 100000.0
    energies = []
    for it in range(opts.maximum_iterations):
        E = None
        for _ in range(2 * n_step):
            E = dmrg_update(mpo, W12, mps, env, oc, step, opts, trace)
            oc += step
            if oc == 0 or oc == L - 2:
                step = -step
        energies.append(E)
        if log is not None:
            log(it, E, mps)
        Eold, E0 = E0, E
        if not (abs((E0 - Eold) / E0) > opts.convergence_criterion):
            break
    if oc != init_pos:
        if oc != init_pos - 1 and init_pos != L - 1:
            raise RuntimeError(f"the orthogonality center finished somewhere surprising! final oc: {oc}")
        mps, oc = move_oc(mps, oc, init_pos)
    return E0, mps, energies


def bond_dims(mps):
    return [sum(t.secs[2]) for t in mps[:-1]]
