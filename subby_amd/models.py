"""Host-side model construction in FP64 (the eigendecomposition stays on the host).

Mirrors subby/jax/models.py (hky85_diag :7-73, jukes_cantor_model :76-97,
f81_model :100-123) and subby/jax/diagonalize.py (diagonalize_rate_matrix
:6-26, reconstruct_rate_matrix :29-47) with NumPy float64.
"""
from __future__ import annotations

from typing import NamedTuple

import numpy as np


class Tree(NamedTuple):
    """subby/formats.py:15-17: parentIndex (R,) preorder, distanceToParent (R,)."""
    parentIndex: object
    distanceToParent: object


class DiagModel(NamedTuple):
    """subby/jax/types.py:8-11: eigenvalues (A,), eigenvectors V[a,k], pi (A,)."""
    eigenvalues: object
    eigenvectors: object
    pi: object


class IrrevDiagModel(NamedTuple):
    """subby/jax/types.py:14-18: Q = V diag(mu) V^{-1} with complex eigenvalues (A,),
    right eigenvectors V[a,k] (A,A), V^{-1} (A,A) and a real root distribution pi (A,)."""
    eigenvalues: object
    eigenvectors: object
    eigenvectors_inv: object
    pi: object


class RateModel(NamedTuple):
    """subby/jax/types.py:21-23: reversible rate matrix and root distribution."""
    subRate: object
    rootProb: object


_TS = np.array([[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]], dtype=bool)


def hky85_diag(kappa, pi) -> DiagModel:
    """Closed-form HKY85 eigensystem of S = Pi^1/2 Q Pi^-1/2 (models.py:7-73)."""
    pi = np.asarray(pi, dtype=np.float64)
    pA, pC, pG, pT = pi
    pR, pY = pA + pG, pC + pT
    beta = 1.0 / (2.0 * pR * pY + 2.0 * kappa * (pA * pG + pC * pT))
    ev = np.array([0.0, -beta, -beta * (pR + kappa * pY), -beta * (pY + kappa * pR)])
    s = np.sqrt(pi)
    w1 = np.array([s[0] * pY, -s[1] * pR, s[2] * pY, -s[3] * pR]) / np.sqrt(pR * pY)
    w2 = np.array([0.0, s[1] * pT, 0.0, -s[3] * pC]) / np.sqrt(pC * pT * pY)
    w3 = np.array([s[0] * pG, 0.0, -s[2] * pA, 0.0]) / np.sqrt(pA * pG * pR)
    return DiagModel(ev, np.column_stack([s, w1, w2, w3]), pi)


def hky85_rate_matrix(kappa, pi) -> np.ndarray:
    """Q_ij = beta kappa^{ts(i,j)} pi_j (i != j), rows sum to 0, mean rate 1."""
    pi = np.asarray(pi, dtype=np.float64)
    Q = np.where(_TS, kappa, 1.0) * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q / -(pi @ np.diag(Q))


def hky85_log_kappa_weights(kappa, pi) -> np.ndarray:
    """Fisher weights W for theta = log(kappa): W_ij = dlog Q_ij/dtheta (i != j),
    W_ii = dQ_ii/dtheta, so dlogL/dtheta = sum_ij W_ij Counts_ij (unmasked)
    (Holmes & Rubin 2002, eqs. 3-4)."""
    pi = np.asarray(pi, dtype=np.float64)
    pA, pC, pG, pT = pi
    pR, pY = pA + pG, pC + pT
    den = 2.0 * pR * pY + 2.0 * kappa * (pA * pG + pC * pT)
    dlog_beta = -2.0 * kappa * (pA * pG + pC * pT) / den
    Q = hky85_rate_matrix(kappa, pi)
    dlogQ = dlog_beta + _TS.astype(np.float64)
    W = dlogQ.copy()
    off = ~np.eye(4, dtype=bool)
    dQ = np.where(off, Q * dlogQ, 0.0)
    np.fill_diagonal(W, -dQ.sum(1))
    return W


def jukes_cantor_model(A) -> DiagModel:
    """JC: pi uniform, eigenvalues (0, -A/(A-1) x (A-1)), QR basis (models.py:76-97)."""
    pi = np.ones(A) / A
    ev = np.zeros(A)
    ev[1:] = -float(A) / (A - 1.0)
    Q, _ = np.linalg.qr(np.column_stack([np.ones(A) / np.sqrt(A), np.eye(A)]))
    return DiagModel(ev, Q[:, :A], pi)


def f81_model(pi) -> DiagModel:
    """F81: eigenvalues (0, -mu, ...), mu = 1/(1 - sum pi^2) (models.py:100-123)."""
    pi = np.asarray(pi, dtype=np.float64)
    A = len(pi)
    ev = np.zeros(A)
    ev[1:] = -1.0 / (1.0 - np.sum(pi ** 2))
    Q, _ = np.linalg.qr(np.column_stack([np.sqrt(pi), np.eye(A)]))
    return DiagModel(ev, Q[:, :A], pi)


def gamma_rate_categories(alpha, K):
    """Yang (1994) discretised gamma rate categories (models.py:126-158): the
    medians of K equal-probability bins of Gamma(alpha, 1/alpha), found by 64
    bisection steps on the regularised incomplete gamma function, normalised to
    mean 1; uniform weights.  Returns (rates (K,), weights (K,))."""
    from scipy.special import gammainc
    p = (np.arange(K) + 0.5) / K
    lo = np.zeros(K)
    hi = np.full(K, max(50.0, 10.0 / float(alpha)))
    for _ in range(64):
        mid = 0.5 * (lo + hi)
        below = gammainc(alpha, mid * alpha) < p
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    r = 0.5 * (lo + hi)
    return r * K / r.sum(), np.full(K, 1.0 / K)


def scale_model(model, rate_multiplier):
    """Eigenvalues times a rate multiplier (models.py:161-186).  A (K,) multiplier
    gives a model stacked along a leading K axis, which LogLike / Counts /
    RootProb evaluate as K models over the same alignment."""
    r = np.asarray(rate_multiplier, dtype=np.float64)
    if isinstance(model, IrrevDiagModel):
        if r.ndim:
            raise ValueError('irreversible models take a scalar rate multiplier')
        return model._replace(eigenvalues=np.asarray(model.eigenvalues) * r)
    m = ensure_diag(model)
    if r.ndim == 0:
        return DiagModel(m.eigenvalues * r, m.eigenvectors, m.pi)
    K = r.shape[0]
    return DiagModel(m.eigenvalues[None, ...] * r.reshape((K,) + (1,) * m.eigenvalues.ndim),
                     np.broadcast_to(m.eigenvectors[None], (K,) + m.eigenvectors.shape),
                     np.broadcast_to(m.pi[None], (K,) + m.pi.shape))


_BASES = 'ACGT'
# Standard genetic code (NCBI translation table 1), codon -> amino acid ('*' = stop),
# written per first base in ACGT order of the second and third bases.
_CODE = {
    'A': 'KNKNTTTTRSRSIIMI',
    'C': 'QHQHPPPPRRRRLLLL',
    'G': 'EDEDAAAAGGGGVVVV',
    'T': '*Y*YSSSS*CWCLFLF',
}


def genetic_code():
    """The 64 codons in ACGT product order (index 16*b1 + 4*b2 + b3), their amino
    acids under the standard code, and the 61 sense codons (formats.py:879-919):
    {'codons', 'amino_acids', 'sense_mask', 'sense_indices', 'codon_to_sense',
     'sense_codons'}."""
    codons = [a + b + c for a in _BASES for b in _BASES for c in _BASES]
    aas = [_CODE[c[0]][4 * _BASES.index(c[1]) + _BASES.index(c[2])] for c in codons]
    sense = np.array([a != '*' for a in aas])
    idx = np.nonzero(sense)[0]
    to_sense = np.full(64, -1, dtype=np.int64)
    to_sense[idx] = np.arange(len(idx))
    return {'codons': codons, 'amino_acids': aas, 'sense_mask': sense, 'sense_indices': idx,
            'codon_to_sense': to_sense, 'sense_codons': [codons[i] for i in idx]}


def codon_to_sense(alignment, A=64):
    """Remap 64-codon tokens to 61 sense-codon tokens: stop codons and gaps become
    the gap token -1, the unobserved token A becomes 61 (formats.py:922-960).
    Returns (alignment, sense codon strings)."""
    gc = genetic_code()
    a = np.asarray(alignment)
    out = np.full(a.shape, -1, dtype=np.int64)
    obs = (a >= 0) & (a < A)
    out[obs] = gc['codon_to_sense'][a[obs]]
    out[a == A] = 61
    return out, gc['sense_codons']


def gy94_rate_matrix(omega, kappa, pi=None) -> np.ndarray:
    """Goldman-Yang (1994) codon rates on the 61 sense codons (models.py:189-232):
    Q_ij = pi_j kappa^[transition] omega^[nonsynonymous] for codons one nucleotide
    apart, 0 otherwise; rows sum to 0; -sum_i pi_i Q_ii = 1."""
    gc = genetic_code()
    sc = gc['sense_codons']
    aa = [gc['amino_acids'][i] for i in gc['sense_indices']]
    pi = np.full(61, 1.0 / 61) if pi is None else np.asarray(pi, dtype=np.float64)
    ts = {('A', 'G'), ('G', 'A'), ('C', 'T'), ('T', 'C')}
    Q = np.zeros((61, 61))
    for i, ci in enumerate(sc):
        for j, cj in enumerate(sc):
            d = [(x, y) for x, y in zip(ci, cj) if x != y]
            if len(d) != 1:
                continue
            Q[i, j] = pi[j] * (kappa if d[0] in ts else 1.0) * (omega if aa[i] != aa[j] else 1.0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q / -(pi @ np.diag(Q))


def gy94_model(omega, kappa, pi=None) -> DiagModel:
    """GY94 codon model as a DiagModel with A = 61 (models.py:189-232)."""
    pi = np.full(61, 1.0 / 61) if pi is None else np.asarray(pi, dtype=np.float64)
    return diagonalize_rate_matrix(gy94_rate_matrix(omega, kappa, pi), pi)


def diagonalize_rate_matrix(subRate, rootProb) -> DiagModel:
    """eigh of the symmetrised S_ij = Q_ij sqrt(pi_i/pi_j) (diagonalize.py:6-26)."""
    Q = np.asarray(subRate, dtype=np.float64)
    pi = np.asarray(rootProb, dtype=np.float64)
    s = np.sqrt(pi)
    S = Q * s[:, None] / s[None, :]
    ev, V = np.linalg.eigh(0.5 * (S + S.T))
    return DiagModel(ev, V, pi)


def reversible_rate_matrix(exchangeabilities, pi) -> np.ndarray:
    """Q_ij = s_ij pi_j (i != j), normalised so that -sum_i pi_i Q_ii = 1."""
    s = np.asarray(exchangeabilities, dtype=np.float64)
    pi = np.asarray(pi, dtype=np.float64)
    s = 0.5 * (s + s.T)
    Q = s * pi[None, :]
    np.fill_diagonal(Q, 0.0)
    np.fill_diagonal(Q, -Q.sum(1))
    return Q / -(pi @ np.diag(Q))


def diagonalize_irreversible(subRate, rootProb) -> IrrevDiagModel:
    """Q = V diag(mu) V^{-1} by eig, no symmetrisation (diagonalize.py:49-67)."""
    Q = np.asarray(subRate, dtype=np.float64)
    mu, V = np.linalg.eig(Q)
    return IrrevDiagModel(mu.astype(np.complex128), V.astype(np.complex128),
                          np.linalg.inv(V).astype(np.complex128),
                          np.asarray(rootProb, dtype=np.float64))


def check_detailed_balance(subRate, pi, tol=1e-10) -> bool:
    """|pi_i Q_ij - pi_j Q_ji| < tol for all i, j (diagonalize.py:70-86)."""
    Q = np.asarray(subRate, dtype=np.float64)
    pi = np.asarray(pi, dtype=np.float64)
    F = pi[:, None] * Q
    return bool(np.all(np.abs(F - F.T) < tol))


def irrev_model_from_rate_matrix(subRate, pi) -> IrrevDiagModel:
    """Always the irreversible decomposition (models.py:234-248)."""
    return diagonalize_irreversible(subRate, pi)


def model_from_rate_matrix(subRate, pi, reversible=None, tol=1e-10):
    """DiagModel when reversible (eigh of the symmetrised matrix), IrrevDiagModel
    otherwise; reversible=None detects detailed balance (models.py:251-271)."""
    if reversible is None:
        reversible = check_detailed_balance(subRate, pi, tol)
    if reversible:
        return diagonalize_rate_matrix(subRate, pi)
    return diagonalize_irreversible(subRate, pi)


def diagonalize_rate_matrix_auto(subRate, rootProb, reversible=None, tol=1e-10):
    """The reference's name for model_from_rate_matrix (diagonalize.py:89-120):
    eigh when reversible (detected by detailed balance when ``reversible`` is
    None), else the irreversible decomposition."""
    return model_from_rate_matrix(subRate, rootProb, reversible=reversible, tol=tol)


def ensure_diag(model):
    """RateModel -> DiagModel (subby/jax/__init__.py:44-50); IrrevDiagModel passes
    through with complex128 arrays."""
    if isinstance(model, RateModel):
        return diagonalize_rate_matrix(model.subRate, model.rootProb)
    if isinstance(model, IrrevDiagModel):
        return IrrevDiagModel(np.asarray(model.eigenvalues, dtype=np.complex128),
                              np.asarray(model.eigenvectors, dtype=np.complex128),
                              np.asarray(model.eigenvectors_inv, dtype=np.complex128),
                              np.asarray(model.pi, dtype=np.float64))
    if isinstance(model, DiagModel):
        return DiagModel(np.asarray(model.eigenvalues, dtype=np.float64),
                         np.asarray(model.eigenvectors, dtype=np.float64),
                         np.asarray(model.pi, dtype=np.float64))
    if isinstance(model, dict):
        return DiagModel(np.asarray(model['eigenvalues'], dtype=np.float64),
                         np.asarray(model['eigenvectors'], dtype=np.float64),
                         np.asarray(model['pi'], dtype=np.float64))
    raise TypeError(f'unsupported model type {type(model).__name__}')
