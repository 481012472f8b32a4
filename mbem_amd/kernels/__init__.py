"""Kernel identifiers and coefficient helpers (mbem's src/mbem/kernels/__init__.py)."""
# Kernel identifiers, as in the reference (kernels/__init__.py:4-12).
KERNEL_T = "H"     # T-kernel: displacement jump (slip / u_p) -> displacement
KERNEL_U = "G"     # U-kernel: traction -> displacement
KERNEL_HT = "Ht"   # jump -> elastic traction (FAULT_TRACTION rows: basis.assemble_traction_matrix)
KERNEL_GT = "Gt"   # traction -> traction (FAULT_TRACTION rows: basis.assemble_traction_matrix)
KERNELS = (KERNEL_T, KERNEL_U, KERNEL_HT, KERNEL_GT)
DISPLACEMENT_KERNELS = (KERNEL_T, KERNEL_U)

from .basis import (  # noqa: E402
    TBasis, UBasis, as_eps_array, assemble_t_basis, assemble_t_matrix, assemble_traction_matrix,
    assemble_u_basis, assemble_u_matrix, condense_columns, element_sizes, identity_ids,
    resolve_eps, resolve_patch_eps, t_coeffs, u_coeffs,
)
from .tri_nodal import n_nodes  # noqa: E402


def kernel_n_basis(kernel: str) -> int:
    """Basis count of a kernel: 6 for T, 3 for U (kernels/__init__.py:30-41)."""
    if kernel == KERNEL_T:
        return 6
    if kernel == KERNEL_U:
        return 3
    if kernel in (KERNEL_HT, KERNEL_GT):
        raise ValueError(f"kernel {kernel!r} (a traction kernel) has no basis stack")
    raise ValueError(f"unknown kernel {kernel!r}; expected one of {KERNELS}")


def kernel_coeffs(kernel: str, material):
    """Material coefficient vector of a kernel (kernels/__init__.py:44-53)."""
    if kernel == KERNEL_T:
        return t_coeffs(material.mu, material.lam)
    if kernel == KERNEL_U:
        return u_coeffs(material.mu, material.lam)
    if kernel in (KERNEL_HT, KERNEL_GT):
        raise ValueError(f"kernel {kernel!r} (a traction kernel) has no coefficient vector")
    raise ValueError(f"unknown kernel {kernel!r}; expected one of {KERNELS}")


__all__ = ["KERNEL_T", "KERNEL_U", "KERNEL_HT", "KERNEL_GT", "KERNELS", "DISPLACEMENT_KERNELS",
           "kernel_n_basis", "kernel_coeffs", "UBasis", "TBasis", "assemble_u_basis",
           "assemble_t_basis", "assemble_u_matrix", "assemble_t_matrix", "u_coeffs", "t_coeffs",
           "as_eps_array", "element_sizes", "resolve_eps", "resolve_patch_eps",
           "assemble_traction_matrix", "condense_columns", "identity_ids", "n_nodes"]
