"""ORACLE -- test infrastructure only (never the product path).

CPU restatement of mbem's interior readout (src/mbem/evaluate.py:194-253,
:256-407, :598-695) on the oracle pair kernels: the representation formula
with the sigma sign rule, every double layer's eigenstress added back (the
near list of EIGEN_NEAR_RADIUS_EPS * (eps + h) by default, the full sum with
``near_list=False``), and a point ON a boundary reading half the interior
limit as the reference does.  Pinned by the identity elastic - total = sum
sigma C:eps* and by the near-list-vs-full agreement (tests), and by the
manufactured solutions of tests/test_readout_pinning.py.

``near`` (an object with ``recovers(patch)``, ``recover(patch, values)``,
``eps_factor`` and ``radius_h``, as recover.NearField) adds the near-face
readout of evaluate.py:410-472 on the nodal restatement (oracle/tri_nodal.py):
the recovered P2 density it is given is taken as input, so this checks the
readout, not the recovery (which tests/test_recover.py pins by itself).
"""
from __future__ import annotations

import numpy as np

from . import tri_kernels as ok
from . import tri_nodal as on
from .dense import _arrays, _eps_of, _t_coeffs, _u_coeffs

EIGEN_NEAR_RADIUS_EPS = 20.0   # defaults.py:846


def _density(model, patch, kind, solution):
    from mbem_amd.model.core import BCType
    if kind == "slip":
        return patch.density()
    fixed = {"u": BCType.PRESCRIBED_DISPLACEMENT, "t": BCType.FREE_TRACTION}
    if patch.bc is fixed[kind]:
        return model.value_of(patch)
    return np.asarray(solution[f"{kind}:{patch.name}"], float)


def _near_blocks(points, cen, radius):
    """evaluate.py:278-295."""
    from scipy.spatial import cKDTree
    if points.shape[0] == 0 or cen.shape[0] == 0:
        return
    tree = cKDTree(points)
    for S in _source_chunks(cen, radius):
        c0 = cen[S].mean(axis=0)
        reach = float(np.max(np.linalg.norm(cen[S] - c0, axis=1) + radius[S]))
        P = np.asarray(tree.query_ball_point(c0, reach), dtype=np.intp)
        if P.size:
            yield P, S


def _near_field(points, patch, u0, u2, t0, mu, lam, e, near, stress, eigen=True):
    """(double, single, eigen) of a patch's near pairs read from the
    recovered density u2 at the narrowed width (evaluate.py:410-472)."""
    from mbem_amd.kernels.basis import element_sizes
    tv, nrm = _arrays(patch)
    e = np.ascontiguousarray(np.asarray(e, float))
    e1 = near.eps_factor * e
    radius = near.radius_h * element_sizes(patch.mesh)
    cen = tv.mean(axis=1)
    shape = (points.shape[0], 3, 3) if stress else (points.shape[0], 3)
    dbl, sgl = np.zeros(shape), np.zeros(shape)
    eig = np.zeros(shape) if stress else None
    do_u, do_t = bool(np.any(u0)), bool(np.any(t0))
    if not (do_u or do_t):
        return dbl, sgl, eig
    u0 = np.asarray(u0, float).reshape(-1, 3)
    t0 = np.asarray(t0, float).reshape(-1, 3)
    u2 = np.asarray(u2, float).reshape(-1, 6, 3)
    c, g = _t_coeffs(mu, lam), _u_coeffs(mu, lam)
    for P, S in _near_blocks(points, cen, radius):
        x = np.ascontiguousarray(points[P])
        v, n = np.ascontiguousarray(tv[S]), np.ascontiguousarray(nrm[S])
        a0, a1 = np.ascontiguousarray(e[S]), np.ascontiguousarray(e1[S])
        if do_u:
            q0 = np.ascontiguousarray(u0[S])
            q2 = np.ascontiguousarray(u2[S].reshape(-1, 3))
            if stress:
                dbl[P] += (on.dd_stress_contract(x, v, n, a1, q2, mu, lam, 2)
                           - ok.dd_stress_contract(x, v, n, a0, q0, mu, lam))
                if eigen:
                    eig[P] += (on.eigenstress_contract(x, v, n, a1, q2, mu, lam, 2)
                               - ok.eigenstress_contract(x, v, n, a0, q0, mu, lam))
            else:
                dbl[P] += (on.t_disp_contract(x, v, n, a1, q2, *c, 2)
                           - ok.t_disp_contract(x, v, n, a0, q0, *c))
        if do_t:
            f0 = np.ascontiguousarray(t0[S])
            if stress:
                sgl[P] += (ok.kelvin_stress_contract(x, v, a1, f0, mu, lam)
                           - ok.kelvin_stress_contract(x, v, a0, f0, mu, lam))
            else:
                sgl[P] += (ok.u_disp_contract(x, v, a1, f0, *g) - ok.u_disp_contract(x, v, a0, f0, *g))
    return dbl, sgl, eig


def evaluate_displacement(model, region, solution, points, eps, near=None):
    """evaluate.py:194-253."""
    pts = np.ascontiguousarray(points, dtype=float)
    u = np.zeros((pts.shape[0], 3))
    m = region.material
    g, c = _u_coeffs(m.mu, m.lam), _t_coeffs(m.mu, m.lam)
    for p in region.patches:
        sigma = float(model.orientation(region, p))
        tv, nrm = _arrays(p)
        e = _eps_of(eps, p)
        u_p = _density(model, p, "u", solution)
        if np.any(u_p):
            u -= sigma * ok.t_disp_contract(pts, tv, nrm, e, u_p, *c)
        t_p = _density(model, p, "t", solution)
        if np.any(t_p):
            u += sigma * ok.u_disp_contract(pts, tv, e, t_p, *g)
        if near is not None and near.recovers(p):
            dbl, sgl, _ = _near_field(pts, p, u_p, near.recover(p, u_p), t_p, m.mu, m.lam, e,
                                      near, stress=False)
            u += sigma * (sgl - dbl)
    for f in region.faults:
        sigma = float(model.orientation(region, f))
        tv, nrm = _arrays(f)
        slip = _density(model, f, "slip", solution)
        if np.any(slip):
            u -= sigma * ok.t_disp_contract(pts, tv, nrm, _eps_of(eps, f), slip, *c)
    return u


def _source_chunks(cen, radius):
    """evaluate.py:256-275."""
    chunks, stack = [], [np.arange(cen.shape[0])]
    while stack:
        idx = stack.pop()
        c = cen[idx]
        span = c.max(axis=0) - c.min(axis=0)
        if idx.size <= 1 or np.linalg.norm(span) <= radius[idx].min():
            chunks.append(idx)
            continue
        order = idx[np.argsort(c[:, int(span.argmax())], kind="stable")]
        stack.append(order[: idx.size // 2])
        stack.append(order[idx.size // 2:])
    return chunks


def _eigen_near(points, tv, nrm, e, dens, h, mu, lam):
    """evaluate.py:278-326."""
    from scipy.spatial import cKDTree
    sig = np.zeros((points.shape[0], 3, 3))
    live = np.flatnonzero((e > 0.0) & np.any(dens != 0.0, axis=1))
    if live.size == 0 or points.shape[0] == 0:
        return sig
    cen = tv.mean(axis=1)[live]
    radius = EIGEN_NEAR_RADIUS_EPS * (e + h)[live]
    tree = cKDTree(points)
    for S in _source_chunks(cen, radius):
        c0 = cen[S].mean(axis=0)
        reach = float(np.max(np.linalg.norm(cen[S] - c0, axis=1) + radius[S]))
        P = np.asarray(tree.query_ball_point(c0, reach), dtype=np.intp)
        if P.size:
            s = live[S]
            sig[P] += ok.eigenstress_contract(points[P], tv[s], nrm[s], e[s], dens[s], mu, lam)
    return sig


def evaluate_stress(model, region, solution, points, eps, subtract_anelastic=True,
                    near_list=True, near=None):
    """evaluate.py:598-695."""
    from mbem_amd.kernels.basis import element_sizes
    pts = np.ascontiguousarray(points, dtype=float)
    sig = np.zeros((pts.shape[0], 3, 3))
    mu, lam = region.material.mu, region.material.lam

    def double_layer(patch, jump, sigma):
        tv, nrm = _arrays(patch)
        e = _eps_of(eps, patch)
        out = -sigma * ok.dd_stress_contract(pts, tv, nrm, e, jump, mu, lam)
        if subtract_anelastic:
            if near_list:
                out += sigma * _eigen_near(pts, tv, nrm, e, np.asarray(jump, float).reshape(-1, 3),
                                           element_sizes(patch.mesh), mu, lam)
            else:
                out += sigma * ok.eigenstress_contract(pts, tv, nrm, e, jump, mu, lam)
        return out

    for p in region.patches:
        sigma = float(model.orientation(region, p))
        u_p = _density(model, p, "u", solution)
        if np.any(u_p):
            sig += double_layer(p, u_p, sigma)
        t_p = _density(model, p, "t", solution)
        if np.any(t_p):
            tv, _ = _arrays(p)
            sig += sigma * ok.kelvin_stress_contract(pts, tv, _eps_of(eps, p), t_p, mu, lam)
        if near is not None and near.recovers(p):
            dbl, sgl, eig = _near_field(pts, p, u_p, near.recover(p, u_p), t_p, mu, lam,
                                        _eps_of(eps, p), near, stress=True,
                                        eigen=subtract_anelastic)
            sig += sigma * (sgl - dbl)
            if subtract_anelastic:
                sig += sigma * eig
    for f in region.faults:
        sigma = float(model.orientation(region, f))
        slip = _density(model, f, "slip", solution)
        if np.any(slip):
            sig += double_layer(f, slip, sigma)
    return sig
