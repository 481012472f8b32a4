"""The oracle is pinned before it is trusted (CPU, no GPU needed).

The oracle (oracle/tri_kernels_ref.c) restates mbem's tri_kernels.py.  These
checks hold it against things its own formulas cannot fake:

* Gauss product quadrature of the POINTWISE mollified Kelvin kernel and of
  its symbolic first and second derivatives (tests/_quad.py) -- U, T,
  force->stress, dislocation->stress and eigenstress, on generic triangles,
  at points on the element, just above it, near a vertex, outside and far,
  including a small eps and nu up to 0.45: agreement to 1e-11 relative, with
  the quadrature itself converged (two subdivision depths agree) first;
* closed forms: I3 = -Omega/h (the solid angle at the lifted point), the
  infinite-plane limit of the blob integral (3/4) eps^4 / (d^2+eps^2)^(5/2)
  (clq/pointwise.py:24-28), the far field -> point kernel x area;
* invariants: Gauss closure of a closed cube (sum of T over a closed surface
  is -I inside and 0 outside for a uniform density, verify_dd_pairing.py),
  rigid-motion covariance of the U and T blocks, and the basis dependency
  N2 = N1 + 3 N3 of the T stack (verify_viscoelastic.py clause b);
* a planted mistake: flipping one sign inside the oracle's formula is caught
  by the quadrature comparison (the oracle is rebuilt with -DPLANT).
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import _quad
from oracle import tri_kernels as ok

TOL = 1e-11


def _tri():
    return np.array([[0.0, 0.0, 0.0], [2.0, 0.3, 0.1], [0.4, 1.7, -0.2]])


def _normal(tv):
    c = np.cross(tv[1] - tv[0], tv[2] - tv[0])
    return c / np.linalg.norm(c)


def _coeffs(mu, lam):
    from mbem_amd.kernels.basis import t_coeffs, u_coeffs
    return u_coeffs(mu, lam), t_coeffs(mu, lam)


def _points(tv, n):
    c = tv.mean(axis=0)
    return [c, c + 0.3 * n, tv[0] + 0.05 * n, tv[1] + np.array([0.5, 0, 0]),
            np.array([3.0, 2.0, 1.0])]


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("mu,lam,eps", [(1.3, 2.1, 0.4), (30.0, 30.0, 0.5), (1.0, 9.0, 0.15)])
def test_closed_forms_match_quadrature(mu, lam, eps):
    tv = _tri()
    n = _normal(tv)
    g, c = _coeffs(mu, lam)
    d = np.array([0.3, -0.7, 0.5])
    depth, ng = (2, 16) if eps >= 0.3 else (3, 20)
    for obs in _points(tv, n):
        qa, q = _quad.pair("G", tv, obs, eps, mu, lam, depth=depth, n=ng)
        assert _rel(qa, q) < 1e-12
        assert _rel(ok.u_matrix_direct(obs[None], tv[None], [eps], *g), q) < TOL
        qa, q = _quad.pair("T", tv, obs, eps, mu, lam, depth=depth, n=ng)
        assert _rel(qa, q) < 1e-12
        assert _rel(ok.t_matrix_direct(obs[None], tv[None], n[None], [eps], *c), q) < TOL
        qa, q = _quad.pair("S", tv, obs, eps, mu, lam, depth=depth, n=ng)
        assert _rel(qa, q) < 1e-12
        S = ok.kelvin_stress_contract(obs[None], tv[None], [eps], d[None], mu, lam)[0]
        assert _rel(S, q @ d) < TOL
        qa, q = _quad.pair("H", tv, obs, eps, mu, lam, depth=depth, n=ng)
        assert _rel(qa, q) < 1e-12
        H = ok.dd_stress_contract(obs[None], tv[None], n[None], [eps], d[None], mu, lam)[0]
        assert _rel(H, q @ d) < TOL
        qa, q = _quad.pair("E", tv, obs, eps, mu, lam, depth=depth, n=ng)
        E = ok.eigenstress_contract(obs[None], tv[None], n[None], [eps], d[None], mu, lam)[0]
        assert _rel(E, q @ d) < TOL


def test_moments_match_quadrature():
    tv = _tri()
    n = _normal(tv)
    for obs in _points(tv, n):
        _, q = _quad.pair("I", tv, obs, 0.3, depth=3)
        I3, I5, I7, z = ok.moments(tv, obs, 0.3)
        assert abs(I3 - q[1]) < 1e-12 * abs(q[1])
        assert abs(I5 - q[2]) < 1e-12 * abs(q[2])
        assert abs(I7 - q[3]) < 1e-11 * abs(q[3])


def test_I3_is_minus_solid_angle_over_h():
    """I3 = int 1/R^3 = -Omega(lifted point)/h, h = sqrt(z^2 + eps^2) (a closed form)."""
    tv = _tri()
    n = _normal(tv)
    obs = tv.mean(axis=0) + 0.2 * n
    eps = 0.35
    z = float((obs - tv[0]) @ n)
    h = np.sqrt(z * z + eps * eps)
    lifted = obs - z * n + h * n
    a, b, cc = tv - lifted
    num = a @ np.cross(b, cc)
    ra, rb, rc = np.linalg.norm(a), np.linalg.norm(b), np.linalg.norm(cc)
    den = ra * rb * rc + rc * (a @ b) + ra * (b @ cc) + rb * (a @ cc)
    omega = 2.0 * np.arctan2(num, den)
    I3 = ok.moments(tv, obs, eps)[0]
    assert abs(I3 - (-omega / h)) < 1e-14 * abs(I3)


def test_infinite_plane_blob_limit():
    """Deep inside a large element the eigenstress density (15 eps^4/8pi) I7 is
    the infinite-plane marginal (3/4) eps^4 / (d^2+eps^2)^(5/2)."""
    L = 1e3
    tv = np.array([[-L, -L, 0.0], [L, -L, 0.0], [0.0, L, 0.0]])
    eps = 0.5
    for d in (0.0, 0.3, 1.0):
        I7 = ok.moments(tv, np.array([0.0, 0.0, d]), eps)[2]
        phi = 15.0 * eps ** 4 / (8.0 * np.pi) * I7
        assert abs(phi - 0.75 * eps ** 4 / (d * d + eps * eps) ** 2.5) < 1e-9


def test_far_field_is_point_kernel_times_area():
    tv = _tri() * 1e-3
    n = _normal(tv)
    area = 0.5 * np.linalg.norm(np.cross(tv[1] - tv[0], tv[2] - tv[0]))
    mu, lam = 30.0, 30.0
    g, _ = _coeffs(mu, lam)
    obs = np.array([5.0, -3.0, 2.0])
    eps = 0.1
    U = ok.u_matrix_direct(obs[None], tv[None], [eps], *g)
    d = obs - tv.mean(axis=0)
    nu = lam / (2 * (lam + mu))
    R = np.sqrt(d @ d + eps ** 2)
    C1 = 1 / (16 * np.pi * mu * (1 - nu))
    G = C1 * ((3 - 4 * nu) * np.eye(3) / R + np.outer(d, d) / R ** 3
              + 2 * (1 - nu) * eps ** 2 * np.eye(3) / R ** 3)
    assert _rel(U, G * area) < 1e-6


def _cube():
    """Closed unit cube [-1,1]^3, 2 triangles per face, outward normals."""
    from mbem_amd.mesh import concatenate, rectangular_patch, vertical_panel
    top = rectangular_patch((-1, 1), (-1, 1), 1.0, 3, 3, normal_up=True)
    bot = rectangular_patch((-1, 1), (-1, 1), -1.0, 3, 3, normal_up=False)
    walls = [vertical_panel("x", 1.0, (-1, 1), (-1, 1), 3, 3, +1),
             vertical_panel("x", -1.0, (-1, 1), (-1, 1), 3, 3, -1),
             vertical_panel("y", 1.0, (-1, 1), (-1, 1), 3, 3, +1),
             vertical_panel("y", -1.0, (-1, 1), (-1, 1), 3, 3, -1)]
    return concatenate([top, bot] + walls)


def test_gauss_closure_of_a_cube():
    """sum_s T[x, s] over a closed surface with a uniform density = -I inside,
    0 outside (the mollified identity holds up to the blob leak, ~(eps/d)^4)."""
    m = _cube()
    tv = m.vertices[m.triangles]
    nrm, _ = m.normals_and_areas()
    eps = np.full(tv.shape[0], 1e-3)     # the blob's leak decays like (eps / d)^4
    _, c = _coeffs(30.0, 30.0)
    pts = np.array([[0.1, -0.2, 0.3], [0.0, 0.0, 0.0], [3.0, 0.5, -0.2]])
    T = ok.t_matrix_direct(pts, tv, nrm, eps, *c)
    S = T.reshape(3, 3, -1, 3).sum(axis=2)
    assert np.max(np.abs(S[0] + np.eye(3))) < 1e-9
    assert np.max(np.abs(S[1] + np.eye(3))) < 1e-9
    assert np.max(np.abs(S[2])) < 1e-9


def test_rigid_motion_covariance():
    """Rotating and translating source and point rotates the blocks: Q K Q^T."""
    rng = np.random.default_rng(0)
    tv = rng.normal(size=(5, 3, 3))
    n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    x = rng.normal(size=(4, 3)) * 2
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] *= -1
    t = np.array([1.0, -2.0, 0.5])
    g, c = _coeffs(30.0, 45.0)
    eps = np.full(5, 0.3)
    for fn, args in ((ok.u_matrix_direct, g), (None, c)):
        if fn is None:
            K = ok.t_matrix_direct(x, tv, n, eps, *c)
            K2 = ok.t_matrix_direct(x @ Q.T + t, tv @ Q.T + t, n @ Q.T, eps, *c)
        else:
            K = fn(x, tv, eps, *args)
            K2 = fn(x @ Q.T + t, tv @ Q.T + t, eps, *args)
        B = K.reshape(4, 3, 5, 3)
        B2 = K2.reshape(4, 3, 5, 3)
        rot = np.einsum("ia,fasb,jb->fisj", Q, B, Q)
        assert np.max(np.abs(rot - B2)) < 1e-12 * np.max(np.abs(B))


def test_basis_dependency_N2_N1_N3():
    """The T basis satisfies N2 = N1 + 3 N3 - ... exactly: with P2's trace
    4 V3 - 3 t and |d|^2 = |d|^2 + eps^2 - eps^2, N2 = N1 + 3 N3 (the
    dependency verify_viscoelastic.py clause b states)."""
    rng = np.random.default_rng(1)
    tv = rng.normal(size=(4, 3, 3))
    n = np.cross(tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    x = rng.normal(size=(3, 3)) * 2
    B = ok.t_basis_matrices(x, tv, n, np.full(4, 0.4))
    assert np.max(np.abs(B[1] - (B[0] + 3.0 * B[2]))) < 1e-12 * np.max(np.abs(B[1]))


def test_planted_sign_error_is_caught(tmp_path):
    """Rebuild the oracle with one sign flipped (-DPLANT) and show the
    quadrature comparison fails: the check is sensitive to a plausible slip."""
    src = os.path.join(os.path.dirname(ok.__file__), "tri_kernels_ref.c")
    text = open(src).read()
    planted = text.replace("double m21 = ", "double m21 = -", 0)
    planted = text.replace("double m21_5 = (m01_3 - BN1_113) / 3.0;",
                           "double m21_5 = (m01_3 + BN1_113) / 3.0;")
    assert planted != text
    csrc = tmp_path / "planted.c"
    csrc.write_text(planted)
    so = tmp_path / "planted.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off", "-o",
                    str(so), str(csrc), "-lm"], check=True)
    import ctypes
    lib = ctypes.CDLL(str(so))
    D = ctypes.POINTER(ctypes.c_double)
    lib.orc_t_matrix_direct.argtypes = [D, ctypes.c_int64, D, D, ctypes.c_int64, D, D, D]
    tv = _tri()
    n = _normal(tv)
    obs = tv.mean(axis=0) + 0.3 * n
    _, c = _coeffs(1.3, 2.1)
    x = np.ascontiguousarray(obs[None])
    tvc = np.ascontiguousarray(tv[None])
    nc = np.ascontiguousarray(n[None])
    e = np.array([0.4])
    cc = np.ascontiguousarray(c)
    out = np.empty((3, 3))
    p = lambda a: a.ctypes.data_as(D)  # noqa: E731
    lib.orc_t_matrix_direct(p(x), 1, p(tvc), p(nc), 1, p(e), p(cc), p(out))
    _, q = _quad.pair("T", tv, obs, 0.4, 1.3, 2.1, depth=2)
    assert _rel(out, q) > 1e-6
    assert _rel(ok.t_matrix_direct(obs[None], tv[None], n[None], [0.4], *c), q) < TOL
