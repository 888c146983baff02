"""Effective Hamiltonians over a basis, their eigenstates, and basis changes.

* ``Hamiltonian`` (molmot/states/hamiltonian.py:43-160): a sum of
  value * operator terms over a basis; ``evaluate`` fills the matrix from the
  upper triangle <b_i|O|b_j>, j >= i, mirrored as its conjugate, dropping
  elements below 1e-15; ``solve`` diagonalises (LAPACK ``eigh`` when the
  matrix is Hermitian, ``eig`` sorted by real part otherwise) into ``State``
  objects sorted by energy.
* ``operator_to_matrix`` (:296-320): every element <b_i|O(args)|b_j>.
* ``overlap_caseb_casea`` (molmot/states/overlaps.py:32-89): the case (a) to
  case (b) transformation <N K S J | J P S Sigma> =
  (-1)^(-J+P+2S) sqrt(2N+1) (J N S; P -K -Sigma), diagonal in J F M ell
  Lambda S I (Hirota eq. 2.3.3).
* ``convert_basis`` (:92-137) and ``compute_tdms`` (molmot/states/tdm.py:
  17-101): eigenvector coefficients re-expressed in another basis, and the
  basis-level dipole operator carried into the eigenbasis,
  d[g, e, q] = C_g^* d_basis[:, :, q] C_e^T, with the excited-to-ground block
  the conjugate transpose stored under the opposite polarisation index.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, List

import numpy as np

from .case_b import phase2
from .wigner import _3j, two

__all__ = ["State", "Hamiltonian", "operator_to_matrix", "overlap_caseb_casea",
           "convert_basis", "compute_tdms"]

_DROP = 1e-15


@dataclass
class State:
    """One eigenstate: energy, coefficients over ``basis``, sorted index."""
    E: float = 0.0
    coeffs: np.ndarray = field(default_factory=lambda: np.zeros(0, complex))
    basis: list = field(default_factory=list, repr=False)
    index: int = 0

    def dominant_state(self):
        i = int(np.argmax(np.abs(self.coeffs)))
        return self.basis[i], self.coeffs[i]


class Hamiltonian:
    def __init__(self, basis: list):
        self.basis = list(basis)
        self.n = len(self.basis)
        self.matrix = np.zeros((self.n, self.n), dtype=complex)
        self.operators: List[tuple] = []
        self.states: List[State] = []
        self._eigenvalues = None
        self._eigenvectors = None

    def add_operator(self, name: str, value: float, op: Callable) -> None:
        self.operators.append((name, value, op))

    def evaluate(self) -> np.ndarray:
        n = self.n
        H = np.zeros((n, n), dtype=complex)
        for _, value, op in self.operators:
            for i in range(n):
                bi = self.basis[i]
                for j in range(i, n):
                    x = op(bi, self.basis[j])
                    if abs(x) > _DROP:
                        H[i, j] += value * x
                        if j != i:
                            H[j, i] += value * np.conj(x)
        self.matrix = H
        return H

    def solve(self) -> List[State]:
        if self.matrix is None or not np.any(self.matrix):
            self.evaluate()
        H = self.matrix
        if np.allclose(H, H.conj().T):
            w, V = np.linalg.eigh(H)
        else:
            w, V = np.linalg.eig(H)
            order = np.argsort(w.real)
            w, V = w[order], V[:, order]
        self._eigenvalues = np.asarray(w.real, dtype=float)
        self._eigenvectors = V
        self.states = [State(E=float(self._eigenvalues[k]), coeffs=V[:, k].copy(),
                             basis=self.basis, index=k) for k in range(self.n)]
        return self.states

    @property
    def energies(self) -> np.ndarray:
        if self._eigenvalues is None:
            self.solve()
        return self._eigenvalues

    @property
    def eigenvectors(self) -> np.ndarray:
        if self._eigenvectors is None:
            self.solve()
        return self._eigenvectors


def operator_to_matrix(basis: list, op: Callable, *args) -> np.ndarray:
    n = len(basis)
    out = np.zeros((n, n), dtype=complex)
    for i, bi in enumerate(basis):
        for j, bj in enumerate(basis):
            x = op(bi, bj, *args)
            if abs(x) > _DROP:
                out[i, j] = x
    return out


def overlap_caseb_casea(sb, sa) -> complex:
    """<case (b) state sb | case (a) state sa>."""
    for a, b in (("J", "J"), ("F", "F"), ("M", "M"), ("ell", "ell"), ("Lambda", "Lambda"),
                 ("S", "S"), ("I", "I")):
        if two(getattr(sb, a)) != two(getattr(sa, b)):
            return 0j
    J2, N2, S2 = two(sb.J), two(sb.N), two(sb.S)
    K2, P2, G2 = two(sb.K), two(sa.P), two(sa.Sigma)
    e = -J2 + P2 + 2 * S2
    if e & 1:
        return 0j
    return complex(phase2(e) * np.sqrt(N2 + 1.0) * _3j(J2, N2, S2, P2, -K2, -G2))


def convert_basis(eigenstates, eigencoeffs: np.ndarray, old_basis: list, new_basis: list,
                  overlap_func) -> np.ndarray:
    """Coefficients (n_eig, n_new) of eigenstates given as (n_eig, n_old)
    rows over ``old_basis``: c_new = c_old S^H with S[i, j] = <new_i|old_j>."""
    S = np.array([[overlap_func(nb, ob) for ob in old_basis] for nb in new_basis],
                 dtype=complex).reshape(len(new_basis), len(old_basis))
    return np.asarray(eigencoeffs) @ S.conj().T


def compute_tdms(ground_states: list, excited_states: list, ground_basis: list,
                 excited_basis: list, tdm_func: Callable, ground_coeffs=None,
                 excited_coeffs=None) -> np.ndarray:
    """(n, n, 3) dipole table, q index 0, 1, 2 = p -1, 0, +1."""
    ng, ne = len(ground_states), len(excited_states)
    Cg = np.array([s.coeffs for s in ground_states]) if ground_coeffs is None else ground_coeffs
    Ce = np.array([s.coeffs for s in excited_states]) if excited_coeffs is None else excited_coeffs
    db = np.zeros((len(ground_basis), len(excited_basis), 3), dtype=complex)
    for m, bg in enumerate(ground_basis):
        for k, be in enumerate(excited_basis):
            for qi, p in enumerate((-1, 0, 1)):
                x = tdm_func(bg, be, p)
                if abs(x) > _DROP:
                    db[m, k, qi] = x
    n = ng + ne
    d = np.zeros((n, n, 3), dtype=complex)
    for qi in range(3):
        blk = Cg.conj() @ db[:, :, qi] @ Ce.T
        d[:ng, ng:, qi] = blk
        d[ng:, :ng, 2 - qi] = blk.conj().T
    return d
