"""FP64 CPU restatement of subby's per-column sufficient-statistics path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it may be imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and nowhere else.  The product path (``subby_amd``) never calls
it; when the HIP library is missing the product raises instead of falling
back here.

It restates, column-vectorised in NumPy float64, the explicit-loop oracle of
the reference (``subby/oracle/oracle.py``) and the JAX data flow it mirrors
(``subby/jax/{pruning,outside,eigensub,f81_fast,components,vjp}.py``).  The
arithmetic order per column follows the reference: the inside pass rescales
the parent after *each* child is multiplied in, with the 1e-300 clamp, and
log-normalisers are accumulated in float64.

Pinning (what checks each function; see tests/test_oracle_golden.py,
tests/test_oracle_invariants.py):
  * every function below is compared with the six golden fixtures the
    reference ships (tests/golden/*.npz, converted from the reference's
    tests/golden/*.json) at <= 1e-12;
  * closed forms / brute force: HKY85 eigensystem round trip, J(t) closed
    forms, M(0)=I, rows of M sum to 1, cherry log-likelihood, exhaustive
    enumeration of internal states for logLike / root posteriors / counts on
    small trees (counts via numerical quadrature of the I^{ab}_{ij}
    integral), Steiner mask vs brute-force graph definition, F81 fast path
    == general eigensub path, dwell sums == branch-length sums, and the
    gradient vs central finite differences; per-branch counts and node /
    branch posteriors against the same enumeration (and Counts = sum of
    BranchCounts);
  * irreversible models (tests/test_oracle_irrev.py): M(t) = expm(Q t),
    complex J(t) by quadrature, logLike / root / posteriors by enumeration,
    counts of both conventions by quadrature (the reference convention with
    Q^T in the integrand, the exact one with Q), exact convention on a
    reversible Q == the reversible path.
No function here is "parity unpinned".
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np


class Tree(NamedTuple):
    """Rooted binary tree in preorder (reference ``subby/formats.py:15-17``)."""
    parentIndex: np.ndarray       # (R,) int, parentIndex[0] == -1, parentIndex[n] < n
    distanceToParent: np.ndarray  # (R,) float64, distanceToParent[0] == 0


class DiagModel(NamedTuple):
    """Eigensystem of the symmetrised rate matrix (``subby/jax/types.py:8-11``)."""
    eigenvalues: np.ndarray   # (A,)
    eigenvectors: np.ndarray  # (A, A), V[a, k] = component a of eigenvector k
    pi: np.ndarray            # (A,)


class IrrevModel(NamedTuple):
    """Eigensystem of an irreversible rate matrix Q = V diag(mu) V^{-1}
    (``subby/jax/types.py:14-18`` IrrevDiagModel).

    ``exact`` selects the counts convention: False (default) restates the
    reference's formulas (eigensub.py:271-385), whose integrand is that of
    Q^T; True projects with (V^T, V^{-1}) so that the counts are the expected
    counts under Q itself.  The two agree on LogLike, RootProb and posteriors.
    """
    eigenvalues: np.ndarray       # (A,) complex
    eigenvectors: np.ndarray      # (A, A) complex, right eigenvectors V[a, k]
    eigenvectors_inv: np.ndarray  # (A, A) complex, V^{-1}
    pi: np.ndarray                # (A,) real root distribution
    exact: bool = False


def is_irrev(model):
    return isinstance(model, IrrevModel)


# ---------------------------------------------------------------------------
# Tree utilities  (subby/oracle/oracle.py:19-62)
# ---------------------------------------------------------------------------

def children_of(parentIndex):
    """left/right child and sibling tables (oracle.py:19-49).

    The first child seen in index order is the left child.
    """
    parentIndex = np.asarray(parentIndex)
    R = len(parentIndex)
    left = np.full(R, -1, dtype=np.int64)
    right = np.full(R, -1, dtype=np.int64)
    for n in range(1, R):
        p = parentIndex[n]
        if left[p] == -1:
            left[p] = n
        else:
            right[p] = n
    sibling = np.full(R, -1, dtype=np.int64)
    for n in range(1, R):
        p = parentIndex[n]
        sibling[n] = right[p] if left[p] == n else left[p]
    return left, right, sibling


def validate_binary_tree(parentIndex):
    """Every node has 0 or 2 children and parents precede children (oracle.py:52-62)."""
    parentIndex = np.asarray(parentIndex)
    R = len(parentIndex)
    if R < 1 or parentIndex[0] != -1:
        raise ValueError("parentIndex[0] must be -1")
    cnt = np.zeros(R, dtype=np.int64)
    for n in range(1, R):
        p = parentIndex[n]
        if not (0 <= p < n):
            raise ValueError(f"parentIndex[{n}]={p} is not a preorder parent")
        cnt[p] += 1
    bad = np.nonzero((cnt != 0) & (cnt != 2))[0]
    if len(bad):
        raise ValueError(f"Node {bad[0]} has {cnt[bad[0]]} children; expected 0 or 2.")


def token_to_likelihood(alignment, A):
    """(R,C) tokens -> (R,C,A) one-hot / all-ones (oracle.py:69-96).

    0..A-1 observed; A (ungapped-unobserved), A+1 (gap) and -1 (gap) all-ones.
    """
    aln = np.asarray(alignment)
    obs = (aln >= 0) & (aln < A)
    L = np.ones(aln.shape + (A,), dtype=np.float64)
    L[obs] = 0.0
    r, c = np.nonzero(obs)
    L[r, c, aln[r, c]] = 1.0
    return L


# ---------------------------------------------------------------------------
# Models  (oracle.py:103-205, 304-335; subby/jax/models.py:7-123)
# ---------------------------------------------------------------------------

def hky85_diag(kappa, pi):
    """Closed-form HKY85 eigensystem (oracle.py:103-161)."""
    pi = np.asarray(pi, dtype=np.float64)
    pA, pC, pG, pT = pi
    pR, pY = pA + pG, pC + pT
    beta = 1.0 / (2.0 * pR * pY + 2.0 * kappa * (pA * pG + pC * pT))
    ev = np.array([0.0, -beta, -beta * (pR + kappa * pY), -beta * (pY + kappa * pR)])
    s = np.sqrt(pi)
    w0 = s.copy()
    w1 = np.array([s[0] * pY, -s[1] * pR, s[2] * pY, -s[3] * pR]) / np.sqrt(pR * pY)
    w2 = np.array([0.0, s[1] * pT, 0.0, -s[3] * pC]) / np.sqrt(pC * pT * pY)
    w3 = np.array([s[0] * pG, 0.0, -s[2] * pA, 0.0]) / np.sqrt(pA * pG * pR)
    return DiagModel(ev, np.column_stack([w0, w1, w2, w3]), pi)


def hky85_rate_matrix(kappa, pi):
    """HKY85 rate matrix Q_ij = beta * kappa^{ts(i,j)} * pi_j, normalised to rate 1.

    States are ordered A, C, G, T; transitions are A<->G and C<->T.
    """
    pi = np.asarray(pi, dtype=np.float64)
    ts = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=bool)
    Q = np.where(ts, kappa, 1.0) * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    Q[np.diag_indices(4)] = -Q.sum(1)
    return Q / -(pi @ np.diag(Q))


def jukes_cantor_model(A):
    """JC: eigenvalues (0, -A/(A-1), ...) and a QR basis (oracle.py:164-180)."""
    pi = np.ones(A) / A
    ev = np.zeros(A)
    ev[1:] = -float(A) / (A - 1.0)
    v0 = np.ones(A) / np.sqrt(A)
    Q, _ = np.linalg.qr(np.column_stack([v0, np.eye(A)]))
    return DiagModel(ev, Q[:, :A], pi)


def f81_model(pi):
    """F81: eigenvalues (0, -mu, ...), mu = 1/(1-sum pi^2) (oracle.py:183-205)."""
    pi = np.asarray(pi, dtype=np.float64)
    A = len(pi)
    ev = np.zeros(A)
    ev[1:] = -1.0 / (1.0 - np.sum(pi ** 2))
    Q, _ = np.linalg.qr(np.column_stack([np.sqrt(pi), np.eye(A)]))
    return DiagModel(ev, Q[:, :A], pi)


def diagonalize_rate_matrix(subRate, rootProb):
    """eigh of S = Pi^{1/2} Q Pi^{-1/2}, symmetrised (oracle.py:304-335)."""
    Q = np.asarray(subRate, dtype=np.float64)
    pi = np.asarray(rootProb, dtype=np.float64)
    s = np.sqrt(pi)
    S = Q * s[:, None] / s[None, :]
    S = 0.5 * (S + S.T)
    ev, V = np.linalg.eigh(S)
    return DiagModel(ev, V, pi)


def reconstruct_rate_matrix(model):
    """Q_ij = sqrt(pi_j/pi_i) (V diag(mu) V^T)_ij (subby/jax/diagonalize.py:29-47)."""
    V, mu, pi = model.eigenvectors, model.eigenvalues, model.pi
    s = np.sqrt(pi)
    return (V * mu) @ V.T / s[:, None] * s[None, :]


# ---------------------------------------------------------------------------
# Substitution matrices and J  (oracle.py:747-777, 931-959)
# ---------------------------------------------------------------------------

def compute_sub_matrices(model, distances):
    """M_ij(t) = sqrt(pi_j/pi_i) sum_k V_ik e^{mu_k t} V_jk, (R,A,A) (oracle.py:747-777).

    Irreversible models: compute_sub_matrices_irrev."""
    if is_irrev(model):
        return compute_sub_matrices_irrev(model, distances)
    V, mu, pi = model.eigenvectors, model.eigenvalues, model.pi
    t = np.asarray(distances, dtype=np.float64)
    e = np.exp(mu[None, :] * t[:, None])                      # (R, A)
    S = np.einsum('ik,rk,jk->rij', V, e, V)
    s = np.sqrt(pi)
    return S / s[None, :, None] * s[None, None, :]


def compute_J(eigenvalues, distances):
    """J_kl(t) per branch; degenerate when |mu_k-mu_l| < 1e-8 (oracle.py:931-959)."""
    mu = np.asarray(eigenvalues, dtype=np.float64)
    t = np.asarray(distances, dtype=np.float64)[:, None, None]
    ek = np.exp(mu[None, :, None] * t)
    el = np.exp(mu[None, None, :] * t)
    diff = (mu[:, None] - mu[None, :])[None]
    deg = np.abs(diff) < 1e-8
    with np.errstate(divide='ignore', invalid='ignore'):
        nondeg = (ek - el) / np.where(deg, 1.0, diff)
    return np.where(deg, t * ek, nondeg)


# ---------------------------------------------------------------------------
# Inside / outside  (oracle.py:784-924)
# ---------------------------------------------------------------------------

def upward_pass(alignment, tree, subMatrices, rootProb):
    """Inside vectors with per-child max-rescaling (oracle.py:784-838).

    Returns U (R,C,A) rescaled, logNormU (R,C), logLike (C,).
    """
    parentIndex = np.asarray(tree.parentIndex)
    R, C = np.asarray(alignment).shape
    A = subMatrices.shape[-1]
    U = token_to_likelihood(alignment, A)
    lnU = np.zeros((R, C))
    for n in range(R - 1, 0, -1):
        p = parentIndex[n]
        U[p] *= U[n] @ subMatrices[n].T
        m = np.maximum(U[p].max(axis=1), 1e-300)
        U[p] /= m[:, None]
        lnU[p] += lnU[n] + np.log(m)
    with np.errstate(divide='ignore'):
        ll = lnU[0] + np.log(U[0] @ np.asarray(rootProb))
    return U, lnU, ll


def downward_pass(U, logNormU, tree, subMatrices, rootProb, alignment):
    """Outside vectors (oracle.py:845-924).

    D[n]_a = (M_s U_s)_a * obs(p)_a * (pi_a if p == 0 else (D[p] M_p)_a),
    rescaled by its max (clamp 1e-300); D[0] stays zero.
    """
    parentIndex = np.asarray(tree.parentIndex)
    R, C, A = U.shape
    _, _, sib = children_of(parentIndex)
    obs = token_to_likelihood(alignment, A)
    D = np.zeros((R, C, A))
    lnD = np.zeros((R, C))
    pi = np.asarray(rootProb)
    for n in range(1, R):
        p, s = parentIndex[n], sib[n]
        sc = U[s] @ subMatrices[s].T
        if p == 0:
            pc = np.broadcast_to(pi, (C, A))
            prior = 0.0
        else:
            pc = D[p] @ subMatrices[p]
            prior = lnD[p]
        raw = sc * pc * obs[p]
        m = np.maximum(raw.max(axis=1), 1e-300)
        D[n] = raw / m[:, None]
        lnD[n] = logNormU[s] + prior + np.log(m)
    return D, lnD


# ---------------------------------------------------------------------------
# Eigensubstitution accumulation  (oracle.py:966-1137)
# ---------------------------------------------------------------------------

def eigenbasis_project(U, D, model):
    """U~ = (U*sqrt(pi)) V ; D~ = (D/sqrt(pi)) V (oracle.py:966-1003)."""
    s = np.sqrt(model.pi)
    return (U * s) @ model.eigenvectors, (D / s) @ model.eigenvectors


def branch_scale(logNormU, logNormD, logLike):
    """scale[n,c] = exp(lnD + lnU - ll) (oracle.py:1036-1037)."""
    with np.errstate(invalid='ignore', over='ignore'):
        return np.exp(logNormD + logNormU - logLike[None, :])


def accumulate_C(D_tilde, U_tilde, J, logNormU, logNormD, logLike, parentIndex,
                 branch_mask=None):
    """C_kl,c = sum_{n>0} D~_k J_kl U~_l scale (oracle.py:1010-1044)."""
    sc = branch_scale(logNormU, logNormD, logLike)[1:]
    if branch_mask is not None:
        sc = np.where(np.asarray(branch_mask, dtype=bool)[1:], sc, 0.0)
    Dt = D_tilde[1:] * sc[..., None]
    return np.einsum('nck,nkl,ncl->klc', Dt, J[1:], U_tilde[1:])


def rate_matrix_symmetric(model):
    """S = V diag(mu) V^T (oracle.py:1107-1114)."""
    V = model.eigenvectors
    return (V * model.eigenvalues) @ V.T


def back_transform(Ceig, model):
    """counts_ij = VCV_ij (i==j: dwell) or S_ij VCV_ij (oracle.py:1087-1137)."""
    V = model.eigenvectors
    VCV = np.einsum('ik,klc,jl->ijc', V, Ceig, V, optimize=True)
    S = rate_matrix_symmetric(model)
    A = len(model.pi)
    eye = np.eye(A, dtype=bool)[..., None]
    return np.where(eye, VCV, S[..., None] * VCV)


def f81_counts(U, D, logNormU, logNormD, logLike, distances, pi, parentIndex,
               branch_mask=None):
    """Closed-form F81/JC counts, O(CRA^2) (oracle.py:1190-1256; f81_fast.py:7-141)."""
    R, C, A = U.shape
    pi = np.asarray(pi)
    mu = 1.0 / (1.0 - np.sum(pi ** 2))
    t = np.asarray(distances, dtype=np.float64)[1:]
    e = np.exp(-mu * t)
    p = 1.0 - e
    alpha = t * e
    beta = p / mu - t * e
    gamma = t * (1.0 + e) - 2.0 * p / mu
    sc = branch_scale(logNormU, logNormD, logLike)[1:]
    if branch_mask is not None:
        sc = np.where(np.asarray(branch_mask, dtype=bool)[1:], sc, 0.0)
    Ds = D[1:] * sc[..., None]
    Ub = U[1:]
    piU = Ub @ pi                       # (R-1, C)
    Dsum = Ds.sum(-1)                   # (R-1, C)
    t1 = np.einsum('n,nci,ncj->ijc', alpha, Ds, Ub)
    t2a = np.einsum('n,nci,nc->ic', beta, Ds, piU)
    t2b = np.einsum('n,nc,ncj->jc', beta, Dsum, Ub)
    t3 = np.einsum('n,nc,nc->c', gamma, Dsum, piU)
    I = t1 + t2a[:, None, :] + pi[:, None, None] * t2b[None, :, :] + \
        pi[:, None, None] * t3[None, None, :]
    out = mu * pi[None, :, None] * I
    idx = np.arange(A)
    out[idx, idx, :] = I[idx, idx, :]
    return out


# ---------------------------------------------------------------------------
# Steiner branch mask  (oracle.py:1372-1435; NOT the depth-bounded JAX version)
# ---------------------------------------------------------------------------

def compute_branch_mask(alignment, parentIndex, A):
    """Branch n active iff n and parent(n) lie in the Steiner tree of ungapped leaves.

    An ungapped leaf has token 0..A (A = ungapped-unobserved counts as ungapped).
    Sequential over the tree, exact on any depth (oracle.py:1372-1435).
    """
    aln = np.asarray(alignment)
    parentIndex = np.asarray(parentIndex)
    R, C = aln.shape
    cnt = np.zeros(R, dtype=np.int64)
    for n in range(1, R):
        cnt[parentIndex[n]] += 1
    is_leaf = cnt == 0
    ung_leaf = is_leaf[:, None] & (aln >= 0) & (aln <= A)
    has = ung_leaf.copy()
    for n in range(R - 1, 0, -1):
        has[parentIndex[n]] |= has[n]
    nchild = np.zeros((R, C), dtype=np.int64)
    for n in range(1, R):
        nchild[parentIndex[n]] += has[n]
    st = ung_leaf | (nchild >= 2)
    for n in range(1, R):
        st[n] |= st[parentIndex[n]] & has[n]
    mask = np.zeros((R, C), dtype=bool)
    for n in range(1, R):
        mask[n] = st[n] & st[parentIndex[n]]
    return mask


# ---------------------------------------------------------------------------
# Irreversible models  (oracle.py:265-288, 474-740; eigensub.py:234-434)
# ---------------------------------------------------------------------------

def diagonalize_irreversible(subRate, rootProb, exact=False):
    """Q = V diag(mu) V^{-1} by eig, no symmetrisation (oracle.py:265-288)."""
    Q = np.asarray(subRate, dtype=np.float64)
    mu, V = np.linalg.eig(Q)
    return IrrevModel(mu.astype(np.complex128), V.astype(np.complex128),
                      np.linalg.inv(V).astype(np.complex128),
                      np.asarray(rootProb, dtype=np.float64), exact)


def compute_sub_matrices_irrev(model, distances):
    """M_ij(t) = Re sum_k V_ik e^{mu_k t} Vinv_kj, (R,A,A) (oracle.py:474-502)."""
    t = np.asarray(distances, dtype=np.float64)
    e = np.exp(model.eigenvalues[None, :] * t[:, None])            # (R, A) complex
    return np.einsum('ik,rk,kj->rij', model.eigenvectors, e, model.eigenvectors_inv).real


def compute_J_complex(eigenvalues, distances):
    """Complex J_kl(t); degenerate when |mu_k - mu_l| < 1e-8 (oracle.py:509-535)."""
    mu = np.asarray(eigenvalues, dtype=np.complex128)
    t = np.asarray(distances, dtype=np.float64)[:, None, None]
    ek = np.exp(mu[None, :, None] * t)
    el = np.exp(mu[None, None, :] * t)
    diff = (mu[:, None] - mu[None, :])[None]
    deg = np.abs(diff) < 1e-8
    with np.errstate(divide='ignore', invalid='ignore'):
        nondeg = (ek - el) / np.where(deg, 1.0, diff)
    return np.where(deg, t * ek, nondeg)


def eigenbasis_project_irrev(U, D, model):
    """Reference convention (oracle.py:538-571): U~_k = sum_b U_b V_bk,
    D~_k = sum_a D_a Vinv_ka.  Exact convention: U^_l = sum_b Vinv_lb U_b,
    D^_k = sum_a D_a V_ak (the factors of M_ai(s) and M_jb(t-s))."""
    V, Vi = model.eigenvectors, model.eigenvectors_inv
    if model.exact:
        return U @ Vi.T, D @ V
    return U @ V, D @ Vi.T


def accumulate_C_complex(D_tilde, U_tilde, J, logNormU, logNormD, logLike, parentIndex,
                         branch_mask=None):
    """C_kl,c = sum_{n>0} D~_k J_kl U~_l scale, complex (oracle.py:574-611)."""
    sc = branch_scale(logNormU, logNormD, logLike)[1:]
    if branch_mask is not None:
        sc = np.where(np.asarray(branch_mask, dtype=bool)[1:], sc, 0.0)
    return np.einsum('nck,nkl,ncl->klc', D_tilde[1:] * sc[..., None], J[1:], U_tilde[1:])


def accumulate_C_complex_per_branch(D_tilde, U_tilde, J, logNormU, logNormD, logLike,
                                    branch_mask=None):
    """accumulate_C_complex keeping the branch axis, branch 0 zeros (oracle.py:614-648)."""
    sc = _branch_scale_masked(logNormU, logNormD, logLike, branch_mask)
    return np.einsum('nck,nkl,ncl->nklc', D_tilde * sc[..., None], J, U_tilde)


def rate_matrix_irrev(model):
    """R = V diag(mu) V^{-1} (complex; real up to rounding) (oracle.py:668-675)."""
    return (model.eigenvectors * model.eigenvalues) @ model.eigenvectors_inv


def _vcv_irrev(Ceig, model, per_branch):
    V, Vi = model.eigenvectors, model.eigenvectors_inv
    if model.exact:     # VCV_ij = sum_kl Vinv_ki C_kl V_jl
        return np.einsum('ki,...klc,jl->...ijc', Vi, Ceig, V, optimize=True)
    return np.einsum('ik,...klc,lj->...ijc', V, Ceig, Vi, optimize=True)   # V C V^{-1}


def back_transform_irrev(Ceig, model):
    """counts_ii = Re VCV_ii, counts_ij = Re(R_ij VCV_ij) (oracle.py:651-696)."""
    VCV = _vcv_irrev(Ceig, model, False)
    Rm = rate_matrix_irrev(model)
    A = len(model.pi)
    eye = np.eye(A, dtype=bool)[..., None]
    return np.where(eye, VCV, Rm[..., None] * VCV).real


def back_transform_irrev_per_branch(Cn, model):
    """back_transform_irrev keeping the branch axis (oracle.py:699-740)."""
    VCV = _vcv_irrev(Cn, model, True)
    Rm = rate_matrix_irrev(model)
    A = len(model.pi)
    eye = np.eye(A, dtype=bool)[None, ..., None]
    return np.where(eye, VCV, Rm[None, ..., None] * VCV).real


def _eigensub_counts(U, D, lnU, lnD, ll, tree, model, mask):
    if is_irrev(model):
        J = compute_J_complex(model.eigenvalues, tree.distanceToParent)
        Ut, Dt = eigenbasis_project_irrev(U, D, model)
        return back_transform_irrev(
            accumulate_C_complex(Dt, Ut, J, lnU, lnD, ll, tree.parentIndex, branch_mask=mask), model)
    J = compute_J(model.eigenvalues, tree.distanceToParent)
    Ut, Dt = eigenbasis_project(U, D, model)
    return back_transform(accumulate_C(Dt, Ut, J, lnU, lnD, ll, tree.parentIndex, branch_mask=mask),
                          model)


# ---------------------------------------------------------------------------
# Public API  (oracle.py:1454-1595; subby/jax/__init__.py:207-650)
# ---------------------------------------------------------------------------

def LogLike(alignment, tree, model):
    M = compute_sub_matrices(model, tree.distanceToParent)
    return upward_pass(alignment, tree, M, model.pi)[2]


def RootProb(alignment, tree, model):
    """q[a,c] = pi_a U[0,c,a] exp(lnU[0,c] - ll[c]) (oracle.py:1571-1595)."""
    M = compute_sub_matrices(model, tree.distanceToParent)
    U, lnU, ll = upward_pass(alignment, tree, M, model.pi)
    return (model.pi[None, :] * U[0] * np.exp(lnU[0] - ll)[:, None]).T


def _resolve_mask(alignment, tree, model, branch_mask):
    if isinstance(branch_mask, str):
        if branch_mask != "auto":
            raise ValueError(f"unknown branch_mask {branch_mask!r}")
        return compute_branch_mask(alignment, tree.parentIndex, len(model.pi))
    return branch_mask


def inside_outside(alignment, tree, model, root_pi=None):
    """All intermediates of one Counts call, for tests.  ``root_pi``: root
    distribution of both passes when not the model's own (model lists)."""
    pi = model.pi if root_pi is None else root_pi
    M = compute_sub_matrices(model, tree.distanceToParent)
    U, lnU, ll = upward_pass(alignment, tree, M, pi)
    D, lnD = downward_pass(U, lnU, tree, M, pi, alignment)
    return M, U, lnU, ll, D, lnD


def Counts(alignment, tree, model, f81_fast=False, branch_mask="auto"):
    """(A,A,C) dwell (diagonal) and substitution counts (oracle.py:1470-1514)."""
    mask = _resolve_mask(alignment, tree, model, branch_mask)
    M, U, lnU, ll, D, lnD = inside_outside(alignment, tree, model)
    if f81_fast:
        assert not is_irrev(model), "F81 fast path is reversible-only"
        return f81_counts(U, D, lnU, lnD, ll, tree.distanceToParent, model.pi,
                          tree.parentIndex, branch_mask=mask)
    return _eigensub_counts(U, D, lnU, lnD, ll, tree, model, mask)


def counts_loglike_root(alignment, tree, model, f81_fast=False, branch_mask="auto"):
    """Counts, logLike and RootProb of one block from one inside/outside sweep."""
    mask = _resolve_mask(alignment, tree, model, branch_mask)
    M, U, lnU, ll, D, lnD = inside_outside(alignment, tree, model)
    if f81_fast:
        assert not is_irrev(model), "F81 fast path is reversible-only"
        cnt = f81_counts(U, D, lnU, lnD, ll, tree.distanceToParent, model.pi,
                         tree.parentIndex, branch_mask=mask)
    else:
        cnt = _eigensub_counts(U, D, lnU, lnD, ll, tree, model, mask)
    root = (model.pi[None, :] * U[0] * np.exp(lnU[0] - ll)[:, None]).T
    return cnt, ll, root


# ---------------------------------------------------------------------------
# Per-branch counts and posteriors  (oracle.py:1517-1568, 1618-1902;
# subby/jax/__init__.py:450-540, 681-1040; eigensub.py:175-231;
# f81_fast.py:144-...)
# ---------------------------------------------------------------------------

def _branch_scale_masked(lnU, lnD, ll, branch_mask):
    sc = branch_scale(lnU, lnD, ll)
    sc[0] = 0.0
    if branch_mask is not None:
        sc = np.where(np.asarray(branch_mask, dtype=bool), sc, 0.0)
    return sc


def accumulate_C_per_branch(D_tilde, U_tilde, J, logNormU, logNormD, logLike, branch_mask=None):
    """C[n,k,l,c] = D~_k J_kl U~_l scale, no sum over n; branch 0 zeros (eigensub.py:175-203)."""
    sc = _branch_scale_masked(logNormU, logNormD, logLike, branch_mask)
    return np.einsum('nck,nkl,ncl->nklc', D_tilde * sc[..., None], J, U_tilde)


def back_transform_per_branch(Cn, model):
    """back_transform keeping the branch axis (eigensub.py:206-231)."""
    V = model.eigenvectors
    VCV = np.einsum('ik,nklc,jl->nijc', V, Cn, V, optimize=True)
    S = rate_matrix_symmetric(model)
    A = len(model.pi)
    eye = np.eye(A, dtype=bool)[None, ..., None]
    return np.where(eye, VCV, S[None, ..., None] * VCV)


def f81_counts_per_branch(U, D, logNormU, logNormD, logLike, distances, pi, branch_mask=None):
    """f81_counts keeping the branch axis; branch 0 zeros (f81_fast.py:144-...)."""
    R, C, A = U.shape
    pi = np.asarray(pi)
    mu = 1.0 / (1.0 - np.sum(pi ** 2))
    t = np.asarray(distances, dtype=np.float64)
    e = np.exp(-mu * t)
    p = 1.0 - e
    alpha = t * e
    beta = p / mu - t * e
    gamma = t * (1.0 + e) - 2.0 * p / mu
    sc = _branch_scale_masked(logNormU, logNormD, logLike, branch_mask)
    Ds = D * sc[..., None]
    piU = U @ pi                        # (R, C)
    Dsum = Ds.sum(-1)                   # (R, C)
    I = (np.einsum('n,nci,ncj->nijc', alpha, Ds, U)
         + np.einsum('n,nci,nc->nic', beta, Ds, piU)[:, :, None, :]
         + pi[None, :, None, None] * np.einsum('n,nc,ncj->njc', beta, Dsum, U)[:, None, :, :]
         + pi[None, :, None, None] * np.einsum('n,nc,nc->nc', gamma, Dsum, piU)[:, None, None, :])
    out = mu * pi[None, None, :, None] * I
    idx = np.arange(A)
    out[:, idx, idx, :] = I[:, idx, idx, :]
    out[0] = 0.0
    return out


def BranchCounts(alignment, tree, model, f81_fast=False, branch_mask=None):
    """(R,A,A,C) per-branch dwell (diagonal) and substitution counts, branch 0 zeros.

    The JAX entry point applies no mask (subby/jax/__init__.py:450-540); the
    reference oracle's takes branch_mask="auto" by default (oracle.py:1517-1568).
    Here the default is None (the JAX behaviour); "auto" and arrays are honoured.
    """
    mask = _resolve_mask(alignment, tree, model, branch_mask)
    M, U, lnU, ll, D, lnD = inside_outside(alignment, tree, model)
    if f81_fast:
        assert not is_irrev(model), "F81 fast path is reversible-only"
        return f81_counts_per_branch(U, D, lnU, lnD, ll, tree.distanceToParent, model.pi,
                                     branch_mask=mask)
    if is_irrev(model):
        J = compute_J_complex(model.eigenvalues, tree.distanceToParent)
        Ut, Dt = eigenbasis_project_irrev(U, D, model)
        return back_transform_irrev_per_branch(
            accumulate_C_complex_per_branch(Dt, Ut, J, lnU, lnD, ll, branch_mask=mask), model)
    J = compute_J(model.eigenvalues, tree.distanceToParent)
    Ut, Dt = eigenbasis_project(U, D, model)
    return back_transform_per_branch(
        accumulate_C_per_branch(Dt, Ut, J, lnU, lnD, ll, branch_mask=mask), model)


def _with_root_outside(D, lnD, pi):
    """D[0] = pi / max(pi), lnD[0] = log max(pi) (oracle.py:1648-1660)."""
    D = D.copy()
    lnD = lnD.copy()
    mx = max(float(np.max(pi)), 1e-300)
    D[0] = np.asarray(pi)[None, :] / mx
    lnD[0] = np.log(mx)
    return D, lnD


def node_posterior(alignment, tree, model, root_pi=None):
    """(R,A,C) posterior state distribution of every node (oracle.py:1755-1832).

    Root: pi_a U_0a normalised.  Non-root n: (sum_a D[n]_a M_n[a,j]) U_nj normalised,
    D[n] being the outside vector of branch n (indexed by the parent's state).
    ``root_pi``: root distribution of the passes if not the model's own; the
    root row keeps the model's pi (D[0] = model.pi, __init__.py:768-780).
    """
    M, U, lnU, ll, D, lnD = inside_outside(alignment, tree, model, root_pi)
    D, lnD = _with_root_outside(D, lnD, model.pi)
    with np.errstate(invalid='ignore', over='ignore'):
        sc = np.exp(lnU + lnD - ll[None, :])
    Dt = np.einsum('nca,naj->ncj', D, M)
    post = Dt * U * sc[..., None]
    post[0] = D[0] * U[0] * sc[0][:, None]
    post = post / post.sum(-1, keepdims=True)
    return np.moveaxis(post, -1, -2)


def branch_posterior(alignment, tree, model, root_pi=None):
    """(R,A,A,C) joint posterior P(x_parent(n)=i, x_n=j | column), branch 0 zeros
    (oracle.py:1833-1902): D[n]_i M_n[i,j] U[n]_j normalised over (i, j).
    ``root_pi`` as in node_posterior."""
    M, U, lnU, ll, D, lnD = inside_outside(alignment, tree, model, root_pi)
    D, lnD = _with_root_outside(D, lnD, model.pi)
    with np.errstate(invalid='ignore', over='ignore'):
        sc = np.exp(lnU + lnD - ll[None, :])
    joint = D[:, :, :, None] * M[:, None, :, :] * U[:, :, None, :] * sc[:, :, None, None]
    joint = joint / joint.sum(axis=(-2, -1), keepdims=True)
    joint[0] = 0.0
    return np.moveaxis(joint, 1, -1)


# ---------------------------------------------------------------------------
# Gradient  (subby/jax/vjp.py:62-133; rate parameters via HR2002 eqs. 3-4)
# ---------------------------------------------------------------------------

def loglike_branch_grad_columns(alignment, tree, model):
    """Per-column derivatives d logL_c / d t_n, (R, C) (root row 0), and logL
    (vjp.py:62-133): per_branch[n,c] = scale[n,c] sum_k D~_k mu_k e^{mu_k t_n} U~_k.
    Irreversible models project D with V and U with V^{-1} and keep the real
    part (vjp.py:80-88, 113-131)."""
    M, U, lnU, ll, D, lnD = inside_outside(alignment, tree, model)
    if is_irrev(model):
        Dt = D @ model.eigenvectors
        Ut = U @ model.eigenvectors_inv.T
    else:
        Ut, Dt = eigenbasis_project(U, D, model)
    mu = model.eigenvalues
    t = np.asarray(tree.distanceToParent, dtype=np.float64)
    me = mu[None, :] * np.exp(mu[None, :] * t[:, None])          # (R, A)
    sc = branch_scale(lnU, lnD, ll)
    per = (np.einsum('nck,nk,nck->nc', Dt, me, Ut) * sc).real
    per[0] = 0.0
    return per, ll


def loglike_branch_grad(alignment, tree, model, g=None):
    """d(sum_c g_c logL_c)/d t_n, n = 0..R-1 (root entry 0) (vjp.py:62-133)."""
    per, ll = loglike_branch_grad_columns(alignment, tree, model)
    C = per.shape[1]
    g = np.ones(C) if g is None else np.asarray(g, dtype=np.float64)
    return per @ g, ll


def hky85_loglike_kappa_grad(alignment, tree, kappa, pi, g=None):
    """d(sum_c g_c logL_c)/d log(kappa) for HKY85 via the Fisher identity.

    dlogL/dtheta = sum_{i!=j} N_ij dlog Q_ij/dtheta + sum_i T_i dQ_ii/dtheta
    with unmasked expected counts N (off-diagonal) and dwell times T (diagonal)
    (Holmes & Rubin 2002, eqs. 3-4; reference references/Holmes_Rubin_2002.tex:183-195).
    Q_ij = beta(kappa) kappa^{ts(ij)} pi_j with beta normalising the rate to 1.
    """
    pi = np.asarray(pi, dtype=np.float64)
    model = hky85_diag(kappa, pi)
    N = Counts(alignment, tree, model, branch_mask=None)         # (4,4,C)
    C = N.shape[2]
    g = np.ones(C) if g is None else np.asarray(g, dtype=np.float64)
    Q = hky85_rate_matrix(kappa, pi)
    pA, pC, pG, pT = pi
    pR, pY = pA + pG, pC + pT
    denom = 2.0 * pR * pY + 2.0 * kappa * (pA * pG + pC * pT)
    # d log beta / d log kappa
    dlb = -2.0 * kappa * (pA * pG + pC * pT) / denom
    ts = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=float)
    dlogQ = dlb + ts                                              # off-diagonal d log Q_ij
    off = ~np.eye(4, dtype=bool)
    dQ = np.where(off, Q * dlogQ, 0.0)
    dQii = -dQ.sum(1)
    per_col = np.einsum('ij,ijc->c', np.where(off, dlogQ, 0.0), np.where(off[..., None], N, 0.0)) \
        + np.einsum('i,iic->c', dQii, N)
    return float(per_col @ g)


# ---------------------------------------------------------------------------
# Reference data flow for CPU timing (subby/jax/pruning.py:41-56)
# ---------------------------------------------------------------------------

# ---------------------------------------------------------------------------
# Rate categories, mixtures, per-column model lists and per-branch model grids
# (subby/jax/__init__.py:52-204, 207-283, 338-411, 543-632, 653-678;
#  subby/jax/models.py:126-186; subby/jax/mixture.py; oracle.py:208-262, 1334-1370, 1598-1616)
# ---------------------------------------------------------------------------

def gamma_rate_categories(alpha, K):
    """Yang (1994) discretised gamma: medians of K equal-probability bins of
    Gamma(alpha, 1/alpha) by 64 bisection steps on gammainc, normalised to mean
    rate 1; uniform weights (oracle.py:208-241, models.py:126-158)."""
    from scipy.special import gammainc
    p = (np.arange(K) + 0.5) / K
    lo = np.zeros(K)
    hi = np.full(K, max(50.0, 10.0 / alpha))
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        below = gammainc(alpha, mid * alpha) < p
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    r = 0.5 * (lo + hi)
    return r * K / r.sum(), np.full(K, 1.0 / K)


def scale_model(model, rate_multiplier):
    """Eigenvalues times a rate multiplier (oracle.py:244-262, models.py:161-186)."""
    if is_irrev(model):
        return model._replace(eigenvalues=model.eigenvalues * rate_multiplier)
    return DiagModel(model.eigenvalues * rate_multiplier, model.eigenvectors, model.pi)


def mixture_posterior(log_likes, log_weights):
    """softmax_k(log_likes[k, c] + log_weights[k]) (mixture.py:6-30, oracle.py:1334-1365)."""
    lj = np.asarray(log_likes, dtype=np.float64) + np.asarray(log_weights, dtype=np.float64)[:, None]
    e = np.exp(lj - lj.max(0, keepdims=True))
    return e / e.sum(0, keepdims=True)


def MixturePosterior(alignment, tree, models, log_weights):
    """(K, C) posterior over mixture components (__init__.py:653-678, oracle.py:1598-1616)."""
    return mixture_posterior(np.stack([LogLike(alignment, tree, m) for m in models]), log_weights)


# --- per-column model lists: model[c] for column c ---

def LogLike_columns(alignment, tree, models):
    """Per-column models, with the root distribution of models[0] for every
    column, as the reference stacks the models and passes stacked.pi[0] to the
    inside pass (__init__.py:264-274)."""
    aln = np.asarray(alignment)
    pi0 = models[0].pi
    out = np.empty(aln.shape[1])
    for c, m in enumerate(models):
        out[c] = upward_pass(aln[:, c:c + 1], tree, compute_sub_matrices(m, tree.distanceToParent), pi0)[2][0]
    return out


def RootProb_columns(alignment, tree, models):
    """q[:, c] = pi_c U0 exp(lnU0 - ll_c) with ll_c from LogLike_columns
    (root distribution of models[0]) (__init__.py:617-632)."""
    aln = np.asarray(alignment)
    pi0 = models[0].pi
    out = np.empty((len(pi0), aln.shape[1]))
    for c, m in enumerate(models):
        U, lnU, ll = upward_pass(aln[:, c:c + 1], tree, compute_sub_matrices(m, tree.distanceToParent), pi0)
        out[:, c] = m.pi * U[0, 0] * np.exp(lnU[0, 0] - ll[0])
    return out


def node_posterior_columns(alignment, tree, models):
    """InsideOutside(models list).node_posterior() (subby/jax/__init__.py:719-780,
    900-960): column c under models[c]'s matrices with the root distribution of
    models[0] in both passes; the root row with models[c].pi (D_root is the
    stacked per-column pi there)."""
    aln = np.asarray(alignment)
    pi0 = models[0].pi
    return np.concatenate([node_posterior(aln[:, c:c + 1], tree, m, root_pi=pi0)
                           for c, m in enumerate(models)], axis=-1)


def branch_posterior_columns(alignment, tree, models):
    """InsideOutside(models list).branch_posterior() (__init__.py:1008-1040):
    column c under models[c] with the root distribution of models[0]."""
    aln = np.asarray(alignment)
    pi0 = models[0].pi
    return np.concatenate([branch_posterior(aln[:, c:c + 1], tree, m, root_pi=pi0)
                           for c, m in enumerate(models)], axis=-1)


def _per_column_mask(branch_mask, c):
    if branch_mask is None or isinstance(branch_mask, str):
        return branch_mask
    return np.asarray(branch_mask)[:, c:c + 1]


def Counts_columns(alignment, tree, models, f81_fast=False, branch_mask="auto"):
    """Column c is Counts of that column alone under models[c] (__init__.py:395-411)."""
    aln = np.asarray(alignment)
    return np.concatenate([Counts(aln[:, c:c + 1], tree, m, f81_fast, _per_column_mask(branch_mask, c))
                           for c, m in enumerate(models)], axis=-1)


def BranchCounts_columns(alignment, tree, models, f81_fast=False, branch_mask=None):
    """Column c is BranchCounts of that column alone under models[c] (__init__.py:484-500)."""
    aln = np.asarray(alignment)
    return np.concatenate([BranchCounts(aln[:, c:c + 1], tree, m, f81_fast, _per_column_mask(branch_mask, c))
                           for c, m in enumerate(models)], axis=-1)


# --- per-branch models: model grid (1, R), models_row[n] for branch n, models_row[0] the root ---

def sub_matrices_rows(models_row, distances):
    """(R, A, A): M of branch n from its own model (__init__.py:73-99)."""
    t = np.asarray(distances, dtype=np.float64)
    return np.stack([compute_sub_matrices(m, t[n:n + 1])[0] for n, m in enumerate(models_row)])


def LogLike_rows(alignment, tree, models_row):
    """Inside pass with per-branch M and the root distribution of models_row[0]
    (__init__.py:230-243)."""
    return upward_pass(alignment, tree, sub_matrices_rows(models_row, tree.distanceToParent),
                       models_row[0].pi)[2]


def RootProb_rows(alignment, tree, models_row):
    """(__init__.py:562-583)"""
    U, lnU, ll = upward_pass(alignment, tree, sub_matrices_rows(models_row, tree.distanceToParent),
                             models_row[0].pi)
    return (models_row[0].pi[None, :] * U[0] * np.exp(lnU[0] - ll)[:, None]).T


def BranchCounts_rows(alignment, tree, models_row, branch_mask=None):
    """(R, A, A, C) per-branch counts, branch n projected, accumulated with its
    own J and back-transformed in the eigenbasis of models_row[n]
    (_branch_counts_per_row, __init__.py:102-144; eigensub.py:440-650)."""
    M = sub_matrices_rows(models_row, tree.distanceToParent)
    pi0 = models_row[0].pi
    U, lnU, ll = upward_pass(alignment, tree, M, pi0)
    D, lnD = downward_pass(U, lnU, tree, M, pi0, alignment)
    mask = _resolve_mask(alignment, tree, models_row[0], branch_mask)
    sc = _branch_scale_masked(lnU, lnD, ll, mask)
    R, C, A = U.shape
    t = np.asarray(tree.distanceToParent, dtype=np.float64)
    out = np.zeros((R, A, A, C))
    for n in range(1, R):
        m = models_row[n]
        sl = slice(n, n + 1)
        if is_irrev(m):
            J = compute_J_complex(m.eigenvalues, t[sl])
            Ut, Dt = eigenbasis_project_irrev(U[sl], D[sl], m)
            Cn = np.einsum('nck,nkl,ncl->nklc', Dt * sc[sl, :, None], J, Ut)
            out[n] = back_transform_irrev_per_branch(Cn, m)[0]
        else:
            J = compute_J(m.eigenvalues, t[sl])
            Ut, Dt = eigenbasis_project(U[sl], D[sl], m)
            Cn = np.einsum('nck,nkl,ncl->nklc', Dt * sc[sl, :, None], J, Ut)
            out[n] = back_transform_per_branch(Cn, m)[0]
    return out


def Counts_rows(alignment, tree, models_row, branch_mask="auto"):
    """Sum over branches of the masked per-branch counts (__init__.py:362-377)."""
    return BranchCounts_rows(alignment, tree, models_row, branch_mask).sum(0)


# --- model grid (C, R): grid[c][n] ---

def LogLike_grid(alignment, tree, grid):
    """Column c: LogLike_rows of that column with grid[c] (__init__.py:244-263)."""
    aln = np.asarray(alignment)
    return np.concatenate([LogLike_rows(aln[:, c:c + 1], tree, row) for c, row in enumerate(grid)])


def RootProb_grid(alignment, tree, grid):
    """(__init__.py:584-615)"""
    aln = np.asarray(alignment)
    return np.concatenate([RootProb_rows(aln[:, c:c + 1], tree, row) for c, row in enumerate(grid)], axis=1)


def Counts_grid(alignment, tree, grid, branch_mask="auto"):
    """Column c: Counts of that column with the per-branch models grid[c] (__init__.py:378-392)."""
    aln = np.asarray(alignment)
    return np.concatenate([Counts_rows(aln[:, c:c + 1], tree, row, _per_column_mask(branch_mask, c))
                           for c, row in enumerate(grid)], axis=-1)


def counts_chunked(alignment, tree, model, chunk=128, f81_fast=False, branch_mask="auto"):
    """Counts with the inside pass run over 128-column chunks, as subby.jax does.

    Same numbers as ``Counts``; used only to time the CPU baseline.
    """
    aln = np.asarray(alignment)
    R, C = aln.shape
    M = compute_sub_matrices(model, tree.distanceToParent)
    parts = [upward_pass(aln[:, i:i + chunk], tree, M, model.pi) for i in range(0, C, chunk)]
    U = np.concatenate([p[0] for p in parts], axis=1)
    lnU = np.concatenate([p[1] for p in parts], axis=1)
    ll = np.concatenate([p[2] for p in parts])
    D, lnD = downward_pass(U, lnU, tree, M, model.pi, aln)
    mask = _resolve_mask(aln, tree, model, branch_mask)
    if f81_fast:
        return f81_counts(U, D, lnU, lnD, ll, tree.distanceToParent, model.pi,
                          tree.parentIndex, branch_mask=mask)
    J = compute_J(model.eigenvalues, tree.distanceToParent)
    Ut, Dt = eigenbasis_project(U, D, model)
    return back_transform(accumulate_C(Dt, Ut, J, lnU, lnD, ll, tree.parentIndex,
                                       branch_mask=mask), model)
