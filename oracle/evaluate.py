"""ORACLE -- test infrastructure only (never the product path).

CPU restatement of mbem's interior readout (src/mbem/evaluate.py:194-253,
:256-407, :598-695) on the oracle pair kernels: the representation formula
with the sigma sign rule, every double layer's eigenstress added back (the
near list of EIGEN_NEAR_RADIUS_EPS * (eps + h) by default, the full sum with
``near_list=False``), and a point ON a boundary reading half the interior
limit as the reference does.  Pinned by the identity elastic - total = sum
sigma C:eps* and by the near-list-vs-full agreement (tests).
"""
from __future__ import annotations

import numpy as np

from . import tri_kernels as ok
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


def evaluate_displacement(model, region, solution, points, eps):
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
                    near_list=True):
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
    for f in region.faults:
        sigma = float(model.orientation(region, f))
        slip = _density(model, f, "slip", solution)
        if np.any(slip):
            sig += double_layer(f, slip, sigma)
    return sig
