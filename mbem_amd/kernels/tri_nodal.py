"""The nodal kernel entry points (mbem's src/mbem/kernels/tri_nodal.py:51-56,
1257-1404) with their trailing ``order``: order 0 (one node per element, the
centroid) is what every configuration reaches, and the reference routes it to
tri_kernels' pair code (tri_nodal.py:1259-1355); here it routes to this
package's device kernels (``kernels.tri_kernels``) with the same arguments.
Lagrange P1 / P2 densities (order 1 / 2) and the traction rows of solved
faults (``traction_matrix``, FAULT_TRACTION) are not in mbem_amd and raise
NotImplementedError; a malformed order raises ValueError as the reference's
``n_nodes`` does.
"""
from __future__ import annotations

from . import tri_kernels as tk


def n_nodes(order: int) -> int:
    """Lagrange nodes per triangle, 1 / 3 / 6 for order 0 / 1 / 2
    (tri_nodal.py:51-55)."""
    if order not in (0, 1, 2):
        raise ValueError(f"order must be 0, 1 or 2 (got {order!r})")
    return (order + 1) * (order + 2) // 2


def _order0(order: int, name: str) -> None:
    if n_nodes(order) != 1:
        raise NotImplementedError(f"{name}: order {order} (Lagrange P{order} nodal densities) "
                                  f"is not in mbem_amd; order 0 is")


def u_basis_matrices(x_field, tri_verts, eps_arr, order):
    _order0(order, "u_basis_matrices")
    return tk.u_basis_matrices(x_field, tri_verts, eps_arr)


def t_basis_matrices(x_field, tri_verts, normals, eps_arr, order):
    _order0(order, "t_basis_matrices")
    return tk.t_basis_matrices(x_field, tri_verts, normals, eps_arr)


def u_basis_matrices_serial(x_field, tri_verts, eps_arr, order):
    _order0(order, "u_basis_matrices_serial")
    return tk.u_basis_matrices_serial(x_field, tri_verts, eps_arr)


def t_basis_matrices_serial(x_field, tri_verts, normals, eps_arr, order):
    _order0(order, "t_basis_matrices_serial")
    return tk.t_basis_matrices_serial(x_field, tri_verts, normals, eps_arr)


def u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3, order):
    _order0(order, "u_matrix_direct")
    return tk.u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3)


def t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6, order):
    _order0(order, "t_matrix_direct")
    return tk.t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6)


def u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3, order):
    _order0(order, "u_disp_contract")
    return tk.u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3)


def t_disp_contract(x_field, tri_verts, normals, eps_arr, density, c1, c2, c3, c4, c5, c6,
                    order):
    _order0(order, "t_disp_contract")
    return tk.t_disp_contract(x_field, tri_verts, normals, eps_arr, density,
                              c1, c2, c3, c4, c5, c6)


def kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam, order):
    _order0(order, "kelvin_stress_contract")
    return tk.kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam)


def dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam, order):
    _order0(order, "dd_stress_contract")
    return tk.dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam)


def eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam, order):
    _order0(order, "eigenstress_contract")
    return tk.eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam)


def traction_matrix(x_field, n_field, tri_verts, normals, eps_arr, ids, n_cols, mu, lam, order,
                    which="slip"):
    """Traction rows of a nodal density (tri_nodal.py:1370-1404): the rows of
    solved (FAULT_TRACTION) faults, not in mbem_amd at any order."""
    n_nodes(order)
    if which not in ("slip", "force", "eigen", "total"):
        raise KeyError(which)
    raise NotImplementedError("traction_matrix: the traction rows of solved (FAULT_TRACTION) "
                              "faults are not in mbem_amd")
