"""GPU parity of the P0 pair kernels against the oracle (C restatement of
mbem's tri_kernels.py), entry for entry, on random and adversarial
geometries: points on the element, on edges and vertices, in the source plane
outside it, far away; per-element eps; needles; nu = 0.25 / 0.30 / 0.45;
degenerate sources and zero densities.

Tolerance (FP64 results within 1e-10 relative): every entry
within 1e-10 of the largest entry of its matrix, and the Frobenius norm of the
difference within 1e-12 of the matrix's.
"""
import numpy as np
import pytest

from _geom import adversarial_points, materials, normals_of, random_triangles
from oracle import tri_kernels as ok

pytestmark = pytest.mark.gpu

ENTRY_TOL = 1e-10
FRO_TOL = 1e-12
# needles (height / L = 1e-3) lose digits in the closed forms on both sides
FRO_TOL_NEEDLE = 1e-10


def _close(gpu, ref, what, fro_tol=FRO_TOL):
    scale = np.max(np.abs(ref))
    assert scale > 0, what
    err = np.max(np.abs(gpu - ref))
    fro = np.linalg.norm(gpu - ref) / np.linalg.norm(ref)
    print(f"{what}: max|d|/max|ref| = {err / scale:.2e}, fro = {fro:.2e}", flush=True)
    assert err <= ENTRY_TOL * scale, (what, err / scale)
    assert fro <= fro_tol, (what, fro)


def _case(seed, n_src=24, needle=False):
    rng = np.random.default_rng(seed)
    tv = random_triangles(rng, n_src, needle=needle)
    nrm = normals_of(tv)
    pts = adversarial_points(rng, tv[:8])
    eps = rng.uniform(0.05, 0.8, size=n_src)
    return tv, nrm, pts, eps


def _coeffs(mu, lam):
    from mbem_amd.kernels.basis import t_coeffs, u_coeffs
    return u_coeffs(mu, lam), t_coeffs(mu, lam)


@pytest.mark.parametrize("seed,needle", [(1, False), (2, False), (3, True)])
def test_u_t_matrix_direct(seed, needle):
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, pts, eps = _case(seed, needle=needle)
    tol = FRO_TOL_NEEDLE if needle else FRO_TOL
    for mu, lam in materials():
        g, c = _coeffs(mu, lam)
        _close(gk.u_matrix_direct(pts, tv, eps, *g), ok.u_matrix_direct(pts, tv, eps, *g),
               f"U seed={seed} nu={lam / (2 * (lam + mu)):.2f}", tol)
        _close(gk.t_matrix_direct(pts, tv, nrm, eps, *c), ok.t_matrix_direct(pts, tv, nrm, eps, *c),
               f"T seed={seed} nu={lam / (2 * (lam + mu)):.2f}", tol)


def test_basis_stacks():
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, pts, eps = _case(4, n_src=10)
    U = gk.u_basis_matrices(pts, tv, eps)
    Ur = ok.u_basis_matrices(pts, tv, eps)
    T = gk.t_basis_matrices(pts, tv, nrm, eps)
    Tr = ok.t_basis_matrices(pts, tv, nrm, eps)
    for b in range(3):
        _close(U[b], Ur[b], f"U basis {b}")
    for b in range(6):
        _close(T[b], Tr[b], f"T basis {b}")


def test_degenerate_source_gives_zero_block():
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, pts, eps = _case(5, n_src=6)
    tv[2, 2] = tv[2, 1]                                  # repeated vertex: zero area
    nrm[2] = 0.0
    g, c = _coeffs(30.0, 30.0)
    U = gk.u_matrix_direct(pts, tv, eps, *g)
    T = gk.t_matrix_direct(pts, tv, nrm, eps, *c)
    assert np.all(U[:, 6:9] == 0.0) and np.all(T[:, 6:9] == 0.0)
    _close(U, ok.u_matrix_direct(pts, tv, eps, *g), "U with degenerate")
    _close(T, ok.t_matrix_direct(pts, tv, nrm, eps, *c), "T with degenerate")


@pytest.mark.parametrize("seed", [6, 7])
def test_contractions(seed):
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, pts, eps = _case(seed, n_src=20)
    rng = np.random.default_rng(seed + 100)
    dens = rng.normal(size=(tv.shape[0], 3))
    dens[3] = 0.0                      # a zero-density source is skipped
    dens[5, 1] = 0.0
    for mu, lam in materials():
        g, c = _coeffs(mu, lam)
        _close(gk.u_disp_contract(pts, tv, eps, dens, *g),
               ok.u_disp_contract(pts, tv, eps, dens, *g), "u_disp")
        _close(gk.t_disp_contract(pts, tv, nrm, eps, dens, *c),
               ok.t_disp_contract(pts, tv, nrm, eps, dens, *c), "t_disp")
        _close(gk.kelvin_stress_contract(pts, tv, eps, dens, mu, lam),
               ok.kelvin_stress_contract(pts, tv, eps, dens, mu, lam), "kelvin stress")
        _close(gk.dd_stress_contract(pts, tv, nrm, eps, dens, mu, lam),
               ok.dd_stress_contract(pts, tv, nrm, eps, dens, mu, lam), "dd stress")
        _close(gk.eigenstress_contract(pts, tv, nrm, eps, dens, mu, lam),
               ok.eigenstress_contract(pts, tv, nrm, eps, dens, mu, lam), "eigenstress")


def test_contraction_forms_agree():
    """The two contraction kernels (a lane per point for 65,536 points or
    more, a wave per point below) give the same sums, in a different order:
    70,000 points at once against the same points in two calls of 35,000, for
    all five contractions, with a zero-density and a degenerate source
    skipped by both.  Bar 1e-12 of the largest entry."""
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, _, eps = _case(21, n_src=12)
    tv[7, 2] = tv[7, 0]                       # degenerate (zero area)
    nrm[7] = 0.0
    rng = np.random.default_rng(22)
    lo, hi = tv.reshape(-1, 3).min(0) - 1.0, tv.reshape(-1, 3).max(0) + 1.0
    pts = rng.uniform(lo, hi, size=(70_000, 3))
    dens = rng.normal(size=(12, 3))
    dens[3] = 0.0
    mu, lam = 30.0, 30.0
    g, c = _coeffs(mu, lam)
    calls = {
        "u_disp": lambda x: gk.u_disp_contract(x, tv, eps, dens, *g),
        "t_disp": lambda x: gk.t_disp_contract(x, tv, nrm, eps, dens, *c),
        "kelvin": lambda x: gk.kelvin_stress_contract(x, tv, eps, dens, mu, lam),
        "dd": lambda x: gk.dd_stress_contract(x, tv, nrm, eps, dens, mu, lam),
        "eigen": lambda x: gk.eigenstress_contract(x, tv, nrm, eps, dens, mu, lam),
    }
    for name, f in calls.items():
        whole = f(pts)
        parts = np.concatenate([f(pts[:35_000]), f(pts[35_000:])])
        scale = np.abs(whole).max()
        err = np.abs(whole - parts).max() / scale
        print(f"{name}: lane-per-point vs wave-per-point max|d|/max = {err:.1e}", flush=True)
        assert scale > 0 and np.all(np.isfinite(whole)) and err < 1e-12, name


def test_contraction_equals_matrix_times_density():
    """u_disp_contract == u_matrix_direct @ density (verify_disp_contract.py:50-98)."""
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, pts, eps = _case(8, n_src=16)
    dens = np.random.default_rng(9).normal(size=(16, 3))
    g, c = _coeffs(30.0, 30.0)
    U = gk.u_matrix_direct(pts, tv, eps, *g)
    T = gk.t_matrix_direct(pts, tv, nrm, eps, *c)
    _close(gk.u_disp_contract(pts, tv, eps, dens, *g).ravel(), U @ dens.ravel(), "U contract vs matrix")
    _close(gk.t_disp_contract(pts, tv, nrm, eps, dens, *c).ravel(), T @ dens.ravel(), "T contract vs matrix")


def test_frames():
    from mbem_amd.kernels import tri_kernels as gk
    tv, nrm, pts, eps = _case(10, n_src=12)
    a = gk._precompute_frames(tv)
    b = ok.precompute_frames(tv)
    for x, y in zip(a[:3], b[:3]):
        assert np.max(np.abs(x - y)) < 1e-15
    assert np.array_equal(a[3], b[3])
