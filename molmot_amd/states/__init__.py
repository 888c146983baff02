"""Angular-momentum bases, spectroscopic operators and effective
Hamiltonians from which the molecular tables are built
(molmot/states/: basis.py, case_a.py, case_b.py, hamiltonian.py,
overlaps.py, tdm.py; molmot/wigner.py)."""
from .wigner import wigner3j, wigner6j, wigner9j
from .basis import (BasisState, HundsCaseA_LinearMolecule, HundsCaseB_LinearMolecule,
                    delta, enumerate_states, order_basis_by_m)
from .hamiltonian import (Hamiltonian, State, compute_tdms, convert_basis,
                          operator_to_matrix, overlap_caseb_casea)

__all__ = ["wigner3j", "wigner6j", "wigner9j", "BasisState", "HundsCaseA_LinearMolecule",
           "HundsCaseB_LinearMolecule", "delta", "enumerate_states", "order_basis_by_m",
           "Hamiltonian", "State", "compute_tdms", "convert_basis", "operator_to_matrix",
           "overlap_caseb_casea"]
