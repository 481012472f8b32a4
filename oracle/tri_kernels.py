"""ORACLE -- test infrastructure only (never the product path).

ctypes front end of ``tri_kernels_ref.c``, the plain-C restatement of mbem's
``src/mbem/kernels/tri_kernels.py``.  Same public names and positional
signatures as the reference module (``tri_kernels.py:451-591`` assemblers,
``:976-1260`` contractions), numpy in, freshly allocated numpy out.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module.  Pinning (what checks each function against
something other than its own formula) is listed in the oracle's package
docstring, ``oracle/__init__.py``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle_tri.so")
_lib = None

_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.c_int64


def build() -> None:
    """Compile the C restatement (``oracle/Makefile``)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        lib = ctypes.CDLL(_LIB_PATH)
        dbl = ctypes.c_double
        lib.orc_frames.argtypes = [_D, _I64, _D, ctypes.POINTER(ctypes.c_int32)]
        lib.orc_u_basis_matrices.argtypes = [_D, _I64, _D, _I64, _D, _D]
        lib.orc_t_basis_matrices.argtypes = [_D, _I64, _D, _D, _I64, _D, _D]
        lib.orc_u_matrix_direct.argtypes = [_D, _I64, _D, _I64, _D, dbl, dbl, dbl, _D]
        lib.orc_t_matrix_direct.argtypes = [_D, _I64, _D, _D, _I64, _D, _D, _D]
        lib.orc_u_disp_contract.argtypes = [_D, _I64, _D, _I64, _D, _D, dbl, dbl, dbl, _D]
        lib.orc_t_disp_contract.argtypes = [_D, _I64, _D, _D, _I64, _D, _D, _D, _D]
        lib.orc_kelvin_stress_contract.argtypes = [_D, _I64, _D, _I64, _D, _D, dbl, dbl, _D]
        lib.orc_dd_stress_contract.argtypes = [_D, _I64, _D, _D, _I64, _D, _D, dbl, dbl, _D]
        lib.orc_eigenstress_contract.argtypes = [_D, _I64, _D, _D, _I64, _D, _D, dbl, dbl, _D]
        lib.orc_moments.argtypes = [_D, _D, ctypes.c_double, _D]
        for f in (lib.orc_frames, lib.orc_u_basis_matrices, lib.orc_t_basis_matrices,
                  lib.orc_u_matrix_direct, lib.orc_t_matrix_direct, lib.orc_u_disp_contract,
                  lib.orc_t_disp_contract, lib.orc_kelvin_stress_contract,
                  lib.orc_dd_stress_contract, lib.orc_eigenstress_contract, lib.orc_moments):
            f.restype = None
        _lib = lib
    return _lib


def _c(a, shape_tail=None):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if shape_tail is not None and a.shape[1:] != shape_tail:
        raise ValueError(f"expected trailing shape {shape_tail}, got {a.shape}")
    return a


def _p(a):
    return a.ctypes.data_as(_D)


def precompute_frames(tri_verts):
    """(EX, EY, NH, OK) like ``tri_kernels._precompute_frames`` (:1052-1065)."""
    tv = _c(tri_verts, (3, 3))
    ns = tv.shape[0]
    fr = np.zeros((ns, 9))
    ok = np.zeros(ns, np.int32)
    _load().orc_frames(_p(tv), ns, _p(fr), ok.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return fr[:, 0:3].copy(), fr[:, 3:6].copy(), fr[:, 6:9].copy(), ok.astype(bool)


def u_basis_matrices(x_field, tri_verts, eps_arr):
    """(3, 3N_f, 3N_s): tri_kernels.py:451-469."""
    x, tv, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(eps_arr)
    out = np.empty((3, 3 * x.shape[0], 3 * tv.shape[0]))
    _load().orc_u_basis_matrices(_p(x), x.shape[0], _p(tv), tv.shape[0], _p(e), _p(out))
    return out


def t_basis_matrices(x_field, tri_verts, normals, eps_arr):
    """(6, 3N_f, 3N_s): tri_kernels.py:472-492."""
    x, tv, n, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,)), _c(eps_arr)
    out = np.empty((6, 3 * x.shape[0], 3 * tv.shape[0]))
    _load().orc_t_basis_matrices(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e), _p(out))
    return out


u_basis_matrices_serial = u_basis_matrices
t_basis_matrices_serial = t_basis_matrices


def u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3):
    """(3N_f, 3N_s): tri_kernels.py:546-565."""
    x, tv, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(eps_arr)
    out = np.empty((3 * x.shape[0], 3 * tv.shape[0]))
    _load().orc_u_matrix_direct(_p(x), x.shape[0], _p(tv), tv.shape[0], _p(e),
                                float(g1), float(g2), float(g3), _p(out))
    return out


def t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6):
    """(3N_f, 3N_s): tri_kernels.py:568-591."""
    x, tv, n, e = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,)), _c(eps_arr)
    c = np.array([c1, c2, c3, c4, c5, c6], dtype=np.float64)
    out = np.empty((3 * x.shape[0], 3 * tv.shape[0]))
    _load().orc_t_matrix_direct(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e), _p(c), _p(out))
    return out


def u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3):
    """(N_f, 3): tri_kernels.py:976-1009."""
    x, tv, e, d = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(eps_arr), _c(density, (3,))
    out = np.empty((x.shape[0], 3))
    _load().orc_u_disp_contract(_p(x), x.shape[0], _p(tv), tv.shape[0], _p(e), _p(d),
                                float(g1), float(g2), float(g3), _p(out))
    return out


def t_disp_contract(x_field, tri_verts, normals, eps_arr, density, c1, c2, c3, c4, c5, c6):
    """(N_f, 3): tri_kernels.py:1012-1049."""
    x, tv, n = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,))
    e, d = _c(eps_arr), _c(density, (3,))
    c = np.array([c1, c2, c3, c4, c5, c6], dtype=np.float64)
    out = np.empty((x.shape[0], 3))
    _load().orc_t_disp_contract(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e), _p(d),
                                _p(c), _p(out))
    return out


def kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam):
    """(N_f, 3, 3): tri_kernels.py:1068-1108."""
    x, tv, e, d = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(eps_arr), _c(density, (3,))
    out = np.empty((x.shape[0], 3, 3))
    _load().orc_kelvin_stress_contract(_p(x), x.shape[0], _p(tv), tv.shape[0], _p(e), _p(d),
                                       float(mu), float(lam), _p(out))
    return out


def dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam):
    """(N_f, 3, 3): tri_kernels.py:1111-1154 (total stress, eigenstress included)."""
    x, tv, n = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,))
    e, d = _c(eps_arr), _c(density, (3,))
    out = np.empty((x.shape[0], 3, 3))
    _load().orc_dd_stress_contract(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e), _p(d),
                                   float(mu), float(lam), _p(out))
    return out


def eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam):
    """(N_f, 3, 3): tri_kernels.py:1217-1260 (+C:eps*, full sum)."""
    x, tv, n = _c(x_field, (3,)), _c(tri_verts, (3, 3)), _c(normals, (3,))
    e, d = _c(eps_arr), _c(density, (3,))
    out = np.empty((x.shape[0], 3, 3))
    _load().orc_eigenstress_contract(_p(x), x.shape[0], _p(tv), _p(n), tv.shape[0], _p(e),
                                     _p(d), float(mu), float(lam), _p(out))
    return out


def moments(tri_vert, obs, eps):
    """(I3, I5, I7, z) of one (source, point) pair: tri_kernels.py:655-769."""
    tv, o = _c(tri_vert).reshape(9), _c(obs).reshape(3)
    out = np.empty(4)
    _load().orc_moments(_p(tv), _p(o), float(eps), _p(out))
    return tuple(float(v) for v in out)
