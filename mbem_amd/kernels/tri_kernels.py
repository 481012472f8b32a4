"""Drop-in for mbem's ``mbem.kernels.tri_kernels`` (P0 pair kernels), on the GPU.

Same public names, positional signatures and conventions as the reference
module (src/mbem/kernels/tri_kernels.py): C-contiguous float64 numpy inputs,
a freshly allocated caller-owned numpy output, matrix layout
``out[3f+i, 3s+j]`` (:40-41), degenerate sources giving zero blocks instead
of raising, zero-density sources skipped by the contractions.  Validation
(``ValueError`` on a bad eps) stays one layer up, in ``basis.as_eps_array``,
as in the reference.  Each call moves its inputs to HBM, runs the HIP kernel
(libmbem_hip.so) and copies the result back; the device-resident API the
dense backend uses is ``mbem_amd.kernels.device``.

There is no host fallback: without a GPU every function raises.
"""
from __future__ import annotations

import numpy as np
import torch

from .. import _lib
from . import device as dv

F64 = torch.float64


def _t(a, tail=None, name="array"):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if tail is not None and a.shape[1:] != tail:
        raise ValueError(f"{name}: expected shape (N, {', '.join(map(str, tail))}), got {a.shape}")
    return torch.from_numpy(a).to(_lib.device())


def _src(tri_verts, normals, eps_arr):
    tv = _t(tri_verts, (3, 3), "tri_verts")
    nrm = None if normals is None else _t(normals, (3,), "normals")
    e = _t(eps_arr, None, "eps_arr").reshape(-1)
    return dv.Sources(tv, nrm, e)


def _host(t):
    torch.cuda.current_stream().synchronize()
    return t.cpu().numpy()


def _precompute_frames(tri_verts):
    """(EX, EY, NH, OK) per source, tri_kernels.py:1052-1065."""
    s = _src(tri_verts, None, np.ones(np.asarray(tri_verts).shape[0]))
    b = _host(s.buf)
    return (b[9:12].T.copy(), b[12:15].T.copy(), b[15:18].T.copy(), b[22] != 0.0)


def u_basis_matrices(x_field, tri_verts, eps_arr):
    """U-kernel basis stack, shape (3, 3*N_f, 3*N_s) (tri_kernels.py:451-469)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, None, eps_arr)
    out = torch.zeros((3, 3 * x.shape[0], 3 * s.ns), dtype=F64, device=x.device)
    if x.shape[0] and s.ns:
        dv.u_basis(x, s, out)
    return _host(out)


def t_basis_matrices(x_field, tri_verts, normals, eps_arr):
    """T-kernel basis stack, shape (6, 3*N_f, 3*N_s) (tri_kernels.py:472-492)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, normals, eps_arr)
    out = torch.zeros((6, 3 * x.shape[0], 3 * s.ns), dtype=F64, device=x.device)
    if x.shape[0] and s.ns:
        dv.t_basis(x, s, out)
    return _host(out)


def u_basis_matrices_serial(x_field, tri_verts, eps_arr):
    """Serial twin (tri_kernels.py:502-520): the same GPU kernel; safe to call
    from Python threads (the call is stream-ordered)."""
    return u_basis_matrices(x_field, tri_verts, eps_arr)


def t_basis_matrices_serial(x_field, tri_verts, normals, eps_arr):
    """Serial twin (tri_kernels.py:523-543)."""
    return t_basis_matrices(x_field, tri_verts, normals, eps_arr)


def u_matrix_direct(x_field, tri_verts, eps_arr, g1, g2, g3):
    """U-kernel matrix (3N_f, 3N_s), coefficients in-loop (tri_kernels.py:546-565)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, None, eps_arr)
    out = torch.zeros((3 * x.shape[0], 3 * s.ns), dtype=F64, device=x.device)
    if x.shape[0] and s.ns:
        dv.u_block(x, s, (g1, g2, g3), out, out.stride(0))
    return _host(out)


def t_matrix_direct(x_field, tri_verts, normals, eps_arr, c1, c2, c3, c4, c5, c6):
    """T-kernel matrix (3N_f, 3N_s), coefficients in-loop (tri_kernels.py:568-591)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, normals, eps_arr)
    out = torch.zeros((3 * x.shape[0], 3 * s.ns), dtype=F64, device=x.device)
    if x.shape[0] and s.ns:
        dv.t_block(x, s, (c1, c2, c3, c4, c5, c6), out, out.stride(0))
    return _host(out)


def u_disp_contract(x_field, tri_verts, eps_arr, density, g1, g2, g3):
    """Displacement (N_f,3) of a FORCE density (tri_kernels.py:976-1009)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, None, eps_arr)
    d = _t(density, (3,), "density")
    out = torch.zeros((x.shape[0], 3), dtype=F64, device=x.device)
    dv.u_disp(x, s, d, (g1, g2, g3), out)
    return _host(out)


def t_disp_contract(x_field, tri_verts, normals, eps_arr, density, c1, c2, c3, c4, c5, c6):
    """Displacement (N_f,3) of a SLIP density (tri_kernels.py:1012-1049)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, normals, eps_arr)
    d = _t(density, (3,), "density")
    out = torch.zeros((x.shape[0], 3), dtype=F64, device=x.device)
    dv.t_disp(x, s, d, (c1, c2, c3, c4, c5, c6), out)
    return _host(out)


def kelvin_stress_contract(x_field, tri_verts, eps_arr, density, mu, lam):
    """Stress (N_f,3,3) of a FORCE density (tri_kernels.py:1068-1108)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, None, eps_arr)
    d = _t(density, (3,), "density")
    out = torch.zeros((x.shape[0], 3, 3), dtype=F64, device=x.device)
    dv.stress("force", x, s, d, mu, lam, out)
    return _host(out)


def dd_stress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam):
    """TOTAL stress (N_f,3,3) of a SLIP density (tri_kernels.py:1111-1154)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, normals, eps_arr)
    d = _t(density, (3,), "density")
    out = torch.zeros((x.shape[0], 3, 3), dtype=F64, device=x.device)
    dv.stress("dd", x, s, d, mu, lam, out)
    return _host(out)


def eigenstress_contract(x_field, tri_verts, normals, eps_arr, density, mu, lam):
    """Eigenstress +C:eps* (N_f,3,3) summed over ALL sources (tri_kernels.py:1217-1260)."""
    x = _t(x_field, (3,), "x_field")
    s = _src(tri_verts, normals, eps_arr)
    d = _t(density, (3,), "density")
    out = torch.zeros((x.shape[0], 3, 3), dtype=F64, device=x.device)
    dv.stress("eigen", x, s, d, mu, lam, out)
    return _host(out)
