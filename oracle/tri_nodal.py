"""ORACLE -- test infrastructure only (never the product path).

Restatement of mbem's ``src/mbem/kernels/tri_nodal.py`` (Lagrange P0/P1/P2
nodal pair kernels).  The spec the kernels consume -- node lattice, the
moment degrees each kernel needs, the edge-primitive degrees the recursion
closes over, the series coefficients C(-m/2, j) and the two far-field
product Gauss rules -- is built here in numpy (tri_nodal.py:51-221); the pair
machinery and the drivers are the plain-C restatement ``tri_nodal_ref.c``.
Same public names and positional signatures as the reference module
(``:1257-1404``), numpy in, numpy out.  Order 0 takes the same route as the
reference (``tri_kernels`` for the contracts, the nodal machinery for
``traction_matrix``).

Only ``tests/`` may import this module.  Pinning: tests/test_nodal_oracle.py
(Gauss product quadrature of the pointwise kernels weighted by the Lagrange
basis, the partition of unity sum_k N_k = 1 against the P0 restatement, exact
linear and quadratic densities, the order-0 nodal traction rows against the
P0 stress contractions, and a planted error that must fail).
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from fractions import Fraction
from math import comb

import numpy as np

from . import tri_kernels as tk

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_nodal.so")
_lib = None

# the nodal constants of defaults.py:916-932
NODAL_D_STAR = 10.0
NODAL_FAR_GAUSS_N = 12
NODAL_FAR_GAUSS_N_DISTANT = 8
NODAL_D_STAR_DISTANT = 40.0
NODAL_SERIES_U_OVER_RHO = 2.0
NODAL_SERIES_TERMS = 32
NODAL_SMALL_U_OVER_RHO = 0.5
NODAL_MIN_HEIGHT_OVER_L = 1e-3

KT, KU, KH, KS, KE = 1, 2, 4, 8, 16         # tri_nodal.py:35-39
_M_ROWS = (-1, 1, 3, 5)
_MAX_EDGE_DEG = 5

_D = ctypes.POINTER(ctypes.c_double)
_L = ctypes.POINTER(ctypes.c_int64)
_I64 = ctypes.c_int64
_DBL = ctypes.c_double


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE, "_build/liboracle_nodal.so"], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        spec = [_L, _D, _L, _D, _I64, _D, _D, _D, _I64]
        lib.ornd_u_matrices.argtypes = [_D, _I64, _D, _I64, _D, *spec, _D, _D]
        lib.ornd_t_matrices.argtypes = [_D, _I64, _D, _D, _I64, _D, *spec, _D, _D]
        lib.ornd_disp_contract.argtypes = [_D, _I64, _D, _D, _I64, _D, *spec, _D, _D,
                                           ctypes.c_int, _D]
        lib.ornd_stress_contract.argtypes = [_D, _I64, _D, _D, _I64, _D, *spec, _D, _DBL, _DBL,
                                             _DBL, ctypes.c_int, _D]
        lib.ornd_traction_matrix.argtypes = [_D, _D, _I64, _D, _D, _I64, _D, *spec, _L, _I64,
                                             _DBL, _DBL, _DBL, ctypes.c_int, _D]
        lib.ornd_pair_tables.argtypes = [_D, _D, _DBL, *spec, _D]
        lib.ornd_pair_tables.restype = _DBL
        for f in (lib.ornd_u_matrices, lib.ornd_t_matrices, lib.ornd_disp_contract,
                  lib.ornd_stress_contract, lib.ornd_traction_matrix):
            f.restype = None
        _lib = lib
    return _lib


# ---------------------------------------------------------------------------
# spec: tri_nodal.py:51-221

def n_nodes(order: int) -> int:
    if order not in (0, 1, 2):
        raise ValueError(f"order must be 0, 1 or 2 (got {order!r})")
    return (order + 1) * (order + 2) // 2


def lattice(order: int) -> np.ndarray:
    """Multi-indices (K, 3): vertices, edges 12, 23, 31, interior (:58-74)."""
    p = order
    if p == 0:
        return np.zeros((1, 3), dtype=np.int64)
    out = [[p, 0, 0], [0, p, 0], [0, 0, p]]
    for i, j in ((0, 1), (1, 2), (2, 0)):
        for s in range(1, p):
            a = [0, 0, 0]
            a[i], a[j] = p - s, s
            out.append(a)
    for a1 in range(1, p):
        for a2 in range(1, p - a1):
            if p - a1 - a2 >= 1:
                out.append([a1, a2, p - a1 - a2])
    return np.array(out, dtype=np.int64)


def kernel_degrees(order: int, flags: int) -> dict:
    """{n: highest monomial degree of 1/R^n} a kernel reads (:77-93)."""
    need: dict = {}

    def add(n, d):
        need[n] = max(need.get(n, -1), d)
    if flags & KT:
        add(3, 1 + order); add(5, 3 + order)
    if flags & KH:
        add(3, 0 + order); add(5, 2 + order); add(7, 4 + order)
    if flags & KE:
        add(7, 0 + order)
    if flags & KU:
        add(1, 0 + order); add(3, 2 + order)
    if flags & KS:
        add(3, 1 + order); add(5, 3 + order)
    return need


def close_degrees(need: dict) -> dict:
    """The lower-n degrees the recursion reads, and the vertical seeds (:96-115)."""
    deg = dict(need)
    n = max(deg)
    while True:
        d = deg.get(n, -1)
        if d >= 2:
            deg[n - 2] = max(deg.get(n - 2, -1), d - 2)
        if n - 2 < min(deg):
            break
        n -= 2
    deg[3] = max(deg.get(3, -1), 0)
    for n in list(deg):
        if n >= 5:
            for m in range(3, n, 2):
                deg[m] = max(deg.get(m, -1), 0)
        if n <= 1:
            for m in range(n, 4, 2):
                deg[m] = max(deg.get(m, -1), 0)
    return {n: d for n, d in deg.items() if d >= 0}


def edge_degrees(deg: dict) -> dict:
    """{m: highest power of u in the edge primitives of 1/R^m} (:118-128)."""
    edge: dict = {}
    for n, d in deg.items():
        if d >= 1:
            edge[n - 2] = max(edge.get(n - 2, -1), d - 1)
    for n in deg:
        if n >= 5:
            edge[n - 2] = max(edge.get(n - 2, -1), 0)
        if n <= 1:
            edge[n] = max(edge.get(n, -1), 0)
    return edge


def binom_half_table(n_terms: int) -> np.ndarray:
    """C(-m/2, j), exact then rounded once (:131-140)."""
    out = np.empty((len(_M_ROWS), n_terms))
    for mi, m in enumerate(_M_ROWS):
        a, c = Fraction(-m, 2), Fraction(1)
        for j in range(n_terms):
            out[mi, j] = float(c)
            c *= (a - j) / (j + 1)
    return out


def gauss_triangle(n: int):
    """Collapsed product Gauss-Legendre rule on the unit triangle (:143-153)."""
    g, w = np.polynomial.legendre.leggauss(n)
    g = 0.5 * (g + 1.0)
    w = 0.5 * w
    gi, gj = np.meshgrid(g, g, indexing="ij")
    wi, wj = np.meshgrid(w, w, indexing="ij")
    return gi.ravel(), (gj * (1.0 - gi)).ravel(), (wi * wj * (1.0 - gi)).ravel()


def lagrange_values(order: int, lam: np.ndarray) -> np.ndarray:
    """N_k(lam) (Q, K) by the lattice product formula (:156-165)."""
    lat = lattice(order)
    out = np.ones((lam.shape[0], lat.shape[0]))
    for k, alpha in enumerate(lat):
        for i in range(3):
            for j in range(int(alpha[i])):
                out[:, k] *= (order * lam[:, i] - j) / (j + 1)
    return out


def far_rules(order: int):
    lams, ws = [], []
    for n in (NODAL_FAR_GAUSS_N, NODAL_FAR_GAUSS_N_DISTANT):
        x1, x2, w = gauss_triangle(n)
        lams.append(np.stack([1.0 - x1 - x2, x1, x2], axis=1))
        ws.append(w)
    lam = np.ascontiguousarray(np.concatenate(lams))
    w = np.ascontiguousarray(np.concatenate(ws))
    return lam, w, np.ascontiguousarray(lagrange_values(order, lam)), ws[0].size


_SPECS: dict = {}


def spec(order: int, flags: int):
    """(ispec, fspec, lattice, binom, q_lam, q_w, q_N) as tri_nodal.py:193-221."""
    key = (int(order), int(flags))
    if key in _SPECS:
        return _SPECS[key]
    K = n_nodes(order)
    need0 = kernel_degrees(0, flags)
    deg = close_degrees(kernel_degrees(order, flags))
    edge = edge_degrees(deg)
    if deg.get(-1, -1) >= 1 or min(edge) < -1 or max(edge.values()) > _MAX_EDGE_DEG:
        raise ValueError(f"order {order} kernel {flags}: degrees beyond the P2 spec")
    for m, kmax in edge.items():
        for k in range(2, kmax + 1):
            if edge.get(m - 2, -1) < k - 2:
                raise ValueError("edge-primitive spec is not closed under the recursion")
    q_lam, q_w, q_N, q_split = far_rules(order)
    binom = binom_half_table(NODAL_SERIES_TERMS)
    ispec = np.full(16, -1, dtype=np.int64)
    ispec[0], ispec[1], ispec[2] = order, K, q_split
    for n, d in deg.items():
        ispec[3 + (n + 1) // 2] = d
    for m, d in edge.items():
        ispec[8 + (m + 1) // 2] = d
    for n, d in need0.items():
        ispec[12 + (n - 1) // 2] = d
    fspec = np.array([NODAL_D_STAR, NODAL_D_STAR_DISTANT, NODAL_SMALL_U_OVER_RHO,
                      NODAL_SERIES_U_OVER_RHO])
    s = (ispec, fspec, lattice(order), np.ascontiguousarray(binom), q_lam, q_w, q_N)
    _SPECS[key] = s
    return s


def _spec_args(order, flags):
    ispec, fspec, lat, binom, q_lam, q_w, q_N = spec(order, flags)
    return (ispec.ctypes.data_as(_L), _p(fspec), lat.ctypes.data_as(_L), _p(binom),
            binom.shape[1], _p(q_lam), _p(q_w), _p(q_N), q_w.shape[0])


def _c(a, tail=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if tail is not None and a.shape[1:] != tail:
        raise ValueError(f"expected trailing shape {tail}, got {a.shape}")
    return a


def _p(a):
    return a.ctypes.data_as(_D)


def check_shape(tri_verts):
    """Refuse needles, height / L < NODAL_MIN_HEIGHT_OVER_L (:1239-1254)."""
    tv = np.asarray(tri_verts, float)
    e = np.stack([tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 1], tv[:, 0] - tv[:, 2]])
    L = np.linalg.norm(e, axis=2).max(axis=0)
    area2 = np.linalg.norm(np.cross(e[0], -e[2]), axis=1)
    ratio = area2 / np.where(L > 0.0, L * L, 1.0)
    bad = ratio < NODAL_MIN_HEIGHT_OVER_L
    if bad.any():
        raise ValueError(f"{int(bad.sum())} source triangle(s) have height / L < "
                         f"{NODAL_MIN_HEIGHT_OVER_L:g}")


def pair_tables(tv, obs, eps, order, flags):
    """(W (4, K, 5, 5), z) of one pair (the weighted tables the kernels read)."""
    tv, obs = _c(tv).reshape(9), _c(obs).reshape(3)
    K = n_nodes(order)
    W = np.zeros((4, K, 5, 5))
    z = _load().ornd_pair_tables(_p(tv), _p(obs), float(eps), *_spec_args(order, flags), _p(W))
    return W, z


# ---------------------------------------------------------------------------
# public drivers: tri_nodal.py:1257-1404

def _u(x_field, tri_verts, eps_arr, order, coeffs):
    x, tv, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(eps_arr)
    K = n_nodes(order)
    shape = (3 * x.shape[0], 3 * K * tv.shape[0])
    out = np.zeros(shape if coeffs is not None else (3,) + shape)
    c = None if coeffs is None else _c(coeffs)
    _load().ornd_u_matrices(_p(x), x.shape[0], _p(tv), tv.shape[0], _p(e), *_spec_args(order, KU),
                            None if c is None else _p(c), _p(out))
    return out


def _t(x_field, tri_verts, normals, eps_arr, order, coeffs):
    x, tv, n, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,)), _c(eps_arr)
    K = n_nodes(order)
    shape = (3 * x.shape[0], 3 * K * tv.shape[0])
    out = np.zeros(shape if coeffs is not None else (6,) + shape)
    c = None if coeffs is None else _c(coeffs)
    _load().ornd_t_matrices(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e),
                            *_spec_args(order, KT), None if c is None else _p(c), _p(out))
    return out


def u_basis_matrices(x_field, tri_verts, eps_arr, order):
    if n_nodes(order) == 1:
        return tk.u_basis_matrices(x_field, tri_verts, eps_arr)
    check_shape(tri_verts)
    return _u(x_field, tri_verts, eps_arr, order, None)


def t_basis_matrices(x_field, tri_verts, normals, eps_arr, order):
    if n_nodes(order) == 1:
        return tk.t_basis_matrices(x_field, tri_verts, normals, eps_arr)
    check_shape(tri_verts)
    return _t(x_field, tri_verts, normals, eps_arr, order, None)


def u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3, order):
    if n_nodes(order) == 1:
        return tk.u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3)
    check_shape(tri_verts)
    return _u(x_field, tri_verts, eps_arr, order, [g1, g2, g3])


def t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6, order):
    if n_nodes(order) == 1:
        return tk.t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6)
    check_shape(tri_verts)
    return _t(x_field, tri_verts, normals, eps_arr, order, [c1, c2, c3, c4, c5, c6])


def _disp(x_field, tri_verts, normals, eps_arr, density, coeffs, order, kind, flags):
    x, tv, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(eps_arr)
    n = _c(normals, (3,))
    d = _c(density, (3,))
    c = _c(coeffs)
    out = np.zeros((x.shape[0], 3))
    _load().ornd_disp_contract(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e),
                               *_spec_args(order, flags), _p(d), _p(c), kind, _p(out))
    return out


def u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3, order):
    if n_nodes(order) == 1:
        return tk.u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3)
    check_shape(tri_verts)
    return _disp(x_field, tri_verts, np.zeros((len(tri_verts), 3)), eps_arr, density,
                 [g1, g2, g3], order, 0, KU)


def t_disp_contract(x_field, tri_verts, normals, eps_arr, density, c1, c2, c3, c4, c5, c6,
                    order):
    if n_nodes(order) == 1:
        return tk.t_disp_contract(x_field, tri_verts, normals, eps_arr, density,
                                  c1, c2, c3, c4, c5, c6)
    check_shape(tri_verts)
    return _disp(x_field, tri_verts, normals, eps_arr, density, [c1, c2, c3, c4, c5, c6],
                 order, 1, KT)


def _nu(mu, lam):
    return lam / (2.0 * (lam + mu))


def _stress(x_field, tri_verts, normals, eps_arr, density, mu, lam, order, flags, which):
    x, tv, n, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,)), _c(eps_arr)
    d = _c(density, (3,))
    out = np.zeros((x.shape[0], 3, 3))
    _load().ornd_stress_contract(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e),
                                 *_spec_args(order, flags), _p(d), float(mu), _nu(mu, lam),
                                 float(lam), which, _p(out))
    return out


def kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam, order):
    if n_nodes(order) == 1:
        return tk.kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam)
    check_shape(tri_verts)
    return _stress(x_field, tri_verts, np.zeros((len(tri_verts), 3)), eps_arr, density, mu, lam,
                   order, KS, 1)


def dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam, order):
    if n_nodes(order) == 1:
        return tk.dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam)
    check_shape(tri_verts)
    return _stress(x_field, tri_verts, normals, eps_arr, density, mu, lam, order, KH, 0)


def eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam, order):
    if n_nodes(order) == 1:
        return tk.eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam)
    check_shape(tri_verts)
    return _stress(x_field, tri_verts, normals, eps_arr, density, mu, lam, order, KE, 2)


TRACTION_KINDS = {"slip": (KH, 0), "force": (KS, 1), "eigen": (KE, 2), "total": (KH, 3)}


def traction_matrix(x_field, n_field, tri_verts, normals, eps_arr, ids, n_cols, mu, lam, order,
                    which="slip"):
    """(3 N_f, 3 n_cols) traction rows of a nodal density (:1370-1404)."""
    flags, w = TRACTION_KINDS[which]
    K = n_nodes(order)
    tv = _c(tri_verts, (3, 3))
    ns = tv.shape[0]
    ids = np.ascontiguousarray(np.asarray(ids, dtype=np.int64))
    if ids.shape != (ns, K):
        raise ValueError(f"ids shape {ids.shape} != ({ns}, {K})")
    if ids.size and (ids.min() < 0 or ids.max() >= n_cols):
        raise ValueError(f"ids must index 0 .. n_cols - 1 = {int(n_cols) - 1}")
    if order:
        check_shape(tv)
    nrm = np.zeros((ns, 3)) if which == "force" else _c(normals, (3,))
    x, nf_ = _c(x_field, (3,)), _c(n_field, (3,))
    e = _c(eps_arr)
    out = np.zeros((3 * x.shape[0], 3 * int(n_cols)))
    _load().ornd_traction_matrix(_p(x), _p(nf_), x.shape[0], _p(tv), _p(nrm), ns, _p(e),
                                 *_spec_args(order, flags), ids.ctypes.data_as(_L), int(n_cols),
                                 float(mu), _nu(mu, lam), float(lam), w, _p(out))
    return out
