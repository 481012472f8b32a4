"""The nodal kernel entry points (mbem's src/mbem/kernels/tri_nodal.py:51-56,
1257-1404) with their trailing ``order``, on the GPU.

Order 0 (one node per element, the centroid) routes to this package's P0
device kernels (``kernels.tri_kernels``) with the same arguments, as the
reference routes it to tri_kernels' pair code (tri_nodal.py:1259-1355).
Orders 1 and 2 (Lagrange P1 / P2 densities: nodes at the vertices, then the
edge midpoints 12, 23, 31) run the nodal kernels of csrc/nodal.hip
(``kernels.nodal``).  ``traction_matrix`` (the traction rows of solved,
FAULT_TRACTION faults) runs the nodal kernels at every order, order 0
included, as the reference does.  Numpy in, numpy out; columns
3 (K s + k) + j (element s, node k, component j); densities (K N_s, 3).
Needles (height / L < NODAL_MIN_HEIGHT_OVER_L) raise for orders >= 1, a
malformed order raises ValueError.  There is no host fallback.
"""
from __future__ import annotations

import numpy as np

from . import nodal as nd
from . import tri_kernels as tk


def n_nodes(order: int) -> int:
    """Lagrange nodes per triangle, 1 / 3 / 6 for order 0 / 1 / 2
    (tri_nodal.py:51-55)."""
    return nd.n_nodes(order)


def _host(t):
    import torch
    torch.cuda.current_stream().synchronize()
    return t.cpu().numpy()


def _nodal(tri_verts, normals, eps_arr):
    nd.check_shape(tri_verts)
    return nd.Elements(tri_verts, normals, eps_arr)


def u_basis_matrices(x_field, tri_verts, eps_arr, order):
    """U-kernel basis stack (3, 3 N_f, 3 K N_s) (tri_nodal.py:1257-1262)."""
    if n_nodes(order) == 1:
        return tk.u_basis_matrices(x_field, tri_verts, eps_arr)
    el = _nodal(tri_verts, None, eps_arr)
    return _host(nd.matrix(0, x_field, el, nd.spec(order, nd.KU)))


def t_basis_matrices(x_field, tri_verts, normals, eps_arr, order):
    """T-kernel basis stack (6, 3 N_f, 3 K N_s) (tri_nodal.py:1265-1270)."""
    if n_nodes(order) == 1:
        return tk.t_basis_matrices(x_field, tri_verts, normals, eps_arr)
    el = _nodal(tri_verts, normals, eps_arr)
    return _host(nd.matrix(1, x_field, el, nd.spec(order, nd.KT)))


def u_basis_matrices_serial(x_field, tri_verts, eps_arr, order):
    return u_basis_matrices(x_field, tri_verts, eps_arr, order)


def t_basis_matrices_serial(x_field, tri_verts, normals, eps_arr, order):
    return t_basis_matrices(x_field, tri_verts, normals, eps_arr, order)


def u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3, order):
    """U-kernel matrix (3 N_f, 3 K N_s) with u_coeffs applied (tri_nodal.py:1289-1295)."""
    if n_nodes(order) == 1:
        return tk.u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3)
    el = _nodal(tri_verts, None, eps_arr)
    return _host(nd.matrix(0, x_field, el, nd.spec(order, nd.KU), (g1, g2, g3)))


def t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6, order):
    """T-kernel matrix (3 N_f, 3 K N_s) with t_coeffs applied (tri_nodal.py:1298-1306)."""
    if n_nodes(order) == 1:
        return tk.t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6)
    el = _nodal(tri_verts, normals, eps_arr)
    return _host(nd.matrix(1, x_field, el, nd.spec(order, nd.KT), (c1, c2, c3, c4, c5, c6)))


def u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3, order):
    """Displacement (N_f, 3) of a nodal force density (tri_nodal.py:1309-1315)."""
    if n_nodes(order) == 1:
        return tk.u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3)
    el = _nodal(tri_verts, None, eps_arr)
    return _host(nd.contract(0, x_field, el, nd.spec(order, nd.KU), density, (g1, g2, g3)))


def t_disp_contract(x_field, tri_verts, normals, eps_arr, density, c1, c2, c3, c4, c5, c6,
                    order):
    """Displacement (N_f, 3) of a nodal slip density (tri_nodal.py:1318-1326)."""
    if n_nodes(order) == 1:
        return tk.t_disp_contract(x_field, tri_verts, normals, eps_arr, density,
                                  c1, c2, c3, c4, c5, c6)
    el = _nodal(tri_verts, normals, eps_arr)
    return _host(nd.contract(1, x_field, el, nd.spec(order, nd.KT), density,
                             (c1, c2, c3, c4, c5, c6)))


def kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam, order):
    """Stress (N_f, 3, 3) of a nodal force density (tri_nodal.py:1334-1341)."""
    if n_nodes(order) == 1:
        return tk.kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam)
    el = _nodal(tri_verts, None, eps_arr)
    return _host(nd.contract(2, x_field, el, nd.spec(order, nd.KS), density, mu=mu, lam=lam))


def dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam, order):
    """TOTAL stress (N_f, 3, 3) of a nodal slip density, eigenstress
    included (tri_nodal.py:1344-1352)."""
    if n_nodes(order) == 1:
        return tk.dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam)
    el = _nodal(tri_verts, normals, eps_arr)
    return _host(nd.contract(3, x_field, el, nd.spec(order, nd.KH), density, mu=mu, lam=lam))


def eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam, order):
    """Eigenstress +C:eps* (N_f, 3, 3) of a nodal slip density (tri_nodal.py:1355-1363)."""
    if n_nodes(order) == 1:
        return tk.eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam)
    el = _nodal(tri_verts, normals, eps_arr)
    return _host(nd.contract(4, x_field, el, nd.spec(order, nd.KE), density, mu=mu, lam=lam))


_TRACTION_KINDS = {"slip": (nd.KH, 0), "force": (nd.KS, 1), "eigen": (nd.KE, 2),
                   "total": (nd.KH, 3)}


def traction_matrix(x_field, n_field, tri_verts, normals, eps_arr, ids, n_cols, mu, lam, order,
                    which="slip"):
    """Traction rows (3 N_f, 3 n_cols) of a nodal density: the stress at each
    field point dotted with that point's ``n_field`` (N_f, 3), columns
    scattered by ``ids`` (N_s, K) so that entry (3 f + i, 3 ids[s, k] + j)
    accumulates sigma_im n_m per unit density_j of node k of element s
    (tri_nodal.py:1370-1404).  ``which``: "slip" the ELASTIC traction of a
    slip density (total minus eigenstress), "force" the Kelvin single
    layer's, "eigen" the eigenstress traction alone, "total" the slip
    density's TOTAL traction.  Every order runs the nodal kernels."""
    flags, w = _TRACTION_KINDS[which]
    K = n_nodes(order)
    tv = np.ascontiguousarray(np.asarray(tri_verts, float))
    ns = tv.shape[0]
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int64))
    if ids.shape != (ns, K):
        raise ValueError(f"ids shape {ids.shape} != ({ns}, {K})")
    if ids.size and (ids.min() < 0 or ids.max() >= n_cols):
        raise ValueError(f"ids must index 0 .. n_cols - 1 = {int(n_cols) - 1}")
    if order:
        nd.check_shape(tv)
    nrm = np.zeros((ns, 3)) if which == "force" else normals
    el = nd.Elements(tv, nrm, eps_arr)
    return _host(nd.traction(w, x_field, n_field, el, nd.spec(order, flags), ids, int(n_cols),
                             float(mu), float(lam)))
