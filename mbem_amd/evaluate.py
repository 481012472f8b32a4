"""Interior displacement and stress readout on the GPU (mbem's
src/mbem/evaluate.py:194-253, :256-407, :598-695).

Representation for x inside region R (evaluate.py:1-31):

    u(x) = sum_p sigma G t_p - sum_p sigma H u_p - sum_f sigma_f H slip_f
    sigma(x) = sum_p sigma SG t_p - sum_p sigma SH u_p - sum_f sigma_f SH slip_f

with every double layer's eigenstress +sigma C:eps* added back when
``subtract_anelastic`` (the exact finite-triangle form, summed over the
near list of EIGEN_NEAR_RADIUS_EPS * (eps + h) per element, or over every
element with ``near_list=False``).  A point ON a boundary patch reads HALF
the interior limit (evaluate.py:629-636).  All accumulation happens in one
device buffer; the host sees the result once.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch

from . import _lib, defaults
from .kernels import basis as kb
from .kernels import device as dv
from .model.core import BCType, RegionModel

F64 = torch.float64


def _region(model: RegionModel, region):
    """The MODEL's region of that name (evaluate.py:74-88)."""
    name = region if isinstance(region, str) else region.name
    for r in model.regions:
        if r.name == name:
            return r
    raise KeyError(f"no region named {name!r} (regions: {[r.name for r in model.regions]})")


def _density(model: RegionModel, patch, kind: str, solution: dict, region=None) -> np.ndarray:
    """evaluate.py:91-119: the prescribed value where the BC fixes it, else
    the solved slot; "slip" is a fault's prescribed Burgers vector."""
    if kind == "slip":
        return patch.density()
    fixed = {"u": BCType.PRESCRIBED_DISPLACEMENT, "t": BCType.FREE_TRACTION}
    if kind not in fixed:
        raise ValueError(f"unknown density kind {kind!r}")
    if patch.bc is fixed[kind]:
        return model.value_of(patch)
    return np.asarray(solution[f"{kind}:{patch.name}"], float)


def _warn_near_boundary(points, region):
    """evaluate.py:135-167."""
    from .geometry import distance_to_mesh
    ratio_max = defaults.NEAR_BOUNDARY_H_RATIO
    points = np.asarray(points, float).reshape(-1, 3)
    close = np.zeros(points.shape[0], bool)
    worst, names = np.inf, []
    for p in region.patches:
        d, idx = distance_to_mesh(points, p.mesh)
        ratio = d / kb.element_sizes(p.mesh)[idx]
        if ratio.size and ratio.min() < ratio_max:
            close |= ratio < ratio_max
            worst = min(worst, float(ratio.min()))
            names.append(p.name)
    if close.any():
        warnings.warn(f"{int(close.sum())} of {close.size} observation point(s) lie within "
                      f"{ratio_max}*local-h of boundary patch(es) {', '.join(names)} (min d/h = "
                      f"{worst:.2f}): the piecewise-constant boundary density limits the volume "
                      f"representation there")


class _Ctx:
    """Device copies shared by one readout call."""

    def __init__(self, points, eps):
        self.dev = _lib.device()
        self.x = torch.from_numpy(np.ascontiguousarray(points, dtype=np.float64)).to(self.dev)
        self.eps = eps
        self._src = {}

    def sources(self, patch):
        k = id(patch)
        if k not in self._src:
            tv, nrm = kb._source_arrays(patch.mesh)
            e = kb.resolve_patch_eps(self.eps, patch)
            self._src[k] = (dv.Sources(torch.from_numpy(tv).to(self.dev),
                                       torch.from_numpy(nrm).to(self.dev),
                                       torch.from_numpy(e).to(self.dev)), tv, nrm, e)
        return self._src[k]

    def dens(self, d):
        return torch.from_numpy(np.ascontiguousarray(np.asarray(d, float).reshape(-1, 3))).to(self.dev)


def evaluate_displacement(model: RegionModel, region, solution: dict, points, eps,
                          warn_near: bool = True, near=None) -> np.ndarray:
    """Displacement (N,3) at interior ``points`` of ``region`` (evaluate.py:194-253).
    ``near`` (a ``recover.NearField`` of this model) reads the pairs close to
    a boundary patch from the recovered continuous displacement at the
    narrowed mollification."""
    region = _region(model, region)
    if region.faults:
        from .selfcheck import ensure_fault_convention
        ensure_fault_convention()
    points = np.ascontiguousarray(np.asarray(points, dtype=float).reshape(-1, 3))
    if warn_near and near is None:
        _warn_near_boundary(points, region)
    ctx = _Ctx(points, eps)
    out = torch.zeros((points.shape[0], 3), dtype=F64, device=ctx.dev)
    mat = region.material
    g = kb.u_coeffs(mat.mu, mat.lam)
    c = kb.t_coeffs(mat.mu, mat.lam)
    if points.shape[0]:
        for p in region.patches:
            sigma = float(model.orientation(region, p))
            src = ctx.sources(p)[0]
            u_p = _density(model, p, "u", solution, region)
            if np.any(u_p):
                dv.t_disp(ctx.x, src, ctx.dens(u_p), c, out, scale=-sigma, beta=1.0)
            t_p = _density(model, p, "t", solution)
            if np.any(t_p):
                dv.u_disp(ctx.x, src, ctx.dens(t_p), g, out, scale=sigma, beta=1.0)
            if near is not None and near.recovers(p):
                dbl, sgl, _ = _near_field(points, p, u_p, near.recover(p, u_p), t_p, mat.mu,
                                          mat.lam, kb.resolve_patch_eps(eps, p), near,
                                          stress=False)
                out += sigma * (sgl - dbl)
        for f in region.faults:
            sigma = float(model.orientation(region, f))
            slip = _density(model, f, "slip", solution)
            if np.any(slip):
                dv.t_disp(ctx.x, ctx.sources(f)[0], ctx.dens(slip), c, out, scale=-sigma, beta=1.0)
    torch.cuda.current_stream().synchronize()
    return out.cpu().numpy()


def _source_chunks(cen, radius):
    """Spatially compact source chunks for the near list (evaluate.py:256-275):
    median bisection along the longest box axis until a chunk's box diagonal
    is within the smallest near radius in it; stack order."""
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


def near_list(points, cen, radius):
    """CSR near list (evaluate.py:278-295): for every chunk S the points P
    within the one ball covering its members' balls.  Returns (chunks,
    pt_ptr, pt_chunk): point p's chunks, in chunk order, are
    pt_chunk[pt_ptr[p]:pt_ptr[p+1]]."""
    from scipy.spatial import cKDTree
    n = points.shape[0]
    chunks = []
    if n == 0 or cen.shape[0] == 0:
        return chunks, np.zeros(n + 1, np.int64), np.zeros(0, np.int64)
    tree = cKDTree(points)
    pts_l, ch_l = [], []
    for S in _source_chunks(cen, radius):
        c0 = cen[S].mean(axis=0)
        reach = float(np.max(np.linalg.norm(cen[S] - c0, axis=1) + radius[S]))
        P = np.asarray(tree.query_ball_point(c0, reach), dtype=np.int64)
        if P.size:
            pts_l.append(P)
            ch_l.append(np.full(P.size, len(chunks), np.int64))
            chunks.append(S)
    if not chunks:
        return chunks, np.zeros(n + 1, np.int64), np.zeros(0, np.int64)
    P = np.concatenate(pts_l)
    C = np.concatenate(ch_l)
    order = np.lexsort((C, P))
    counts = np.bincount(P, minlength=n)
    pt_ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return chunks, pt_ptr, C[order]


def _near_blocks(points, cen, radius):
    """(P, S) per source chunk with observation points in reach: the points
    within the one ball covering the chunk members' balls (evaluate.py:278-295)."""
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


def _near_field(points, patch, u0, u2, t0, mu, lam, eps_arr, near, stress, eigen=True):
    """What reading a patch's NEAR pairs from the recovered density adds to
    the element-constant readout: (double, single, eigen), unsigned, on the
    device (evaluate.py:410-472).  For the elements within near.radius_h of
    their own h of a point: the double layer of the recovered P2 displacement
    u2 at the narrowed width in place of the element values u0 at the
    solve's; the single layer of the same element traction t0 at the
    narrowed width in place of the solve's; for stress the eigenstress of
    each of the two double layers."""
    from .kernels import nodal as nd
    dev = _lib.device()
    tv, nrm = kb._source_arrays(patch.mesh)
    eps_arr = np.ascontiguousarray(np.asarray(eps_arr, float))
    eps_near = near.eps_factor * eps_arr
    radius = near.radius_h * kb.element_sizes(patch.mesh)
    cen = tv.mean(axis=1)
    n = points.shape[0]
    shape = (n, 3, 3) if stress else (n, 3)
    dbl = torch.zeros(shape, dtype=F64, device=dev)
    sgl = torch.zeros(shape, dtype=F64, device=dev)
    eig = torch.zeros(shape, dtype=F64, device=dev) if stress else None
    do_u, do_t = bool(np.any(u0)), bool(np.any(t0))
    if not (do_u or do_t):
        return dbl, sgl, eig
    u0 = np.asarray(u0, float).reshape(-1, 3)
    t0 = np.asarray(t0, float).reshape(-1, 3)
    u2 = np.asarray(u2, float).reshape(-1, 6, 3)
    c = kb.t_coeffs(mu, lam)
    g = kb.u_coeffs(mu, lam)
    sp = {k: nd.spec(2, f) for k, f in (("t", nd.KT), ("h", nd.KH), ("e", nd.KE))}
    for P, S in _near_blocks(points, cen, radius):
        xn = np.ascontiguousarray(points[P])
        x = torch.from_numpy(xn).to(dev)
        Pt = torch.from_numpy(P.astype(np.int64)).to(dev)
        src0 = dv.Sources(torch.from_numpy(tv[S]).to(dev), torch.from_numpy(nrm[S]).to(dev),
                          torch.from_numpy(eps_arr[S]).to(dev))
        src1 = dv.Sources(src0.tri_verts, src0.normals, torch.from_numpy(eps_near[S]).to(dev))
        if do_u:
            q0 = torch.from_numpy(np.ascontiguousarray(u0[S])).to(dev)
            q2 = np.ascontiguousarray(u2[S].reshape(-1, 3))
            el = nd.Elements(tv[S], nrm[S], eps_near[S], dev)
            if stress:
                blk = nd.contract(3, xn, el, sp["h"], q2, mu=mu, lam=lam)
                dv.stress("dd", x, src0, q0, mu, lam, blk, scale=-1.0, beta=1.0)
                dbl.index_add_(0, Pt, blk)
                if eigen:
                    blk = nd.contract(4, xn, el, sp["e"], q2, mu=mu, lam=lam)
                    dv.stress("eigen", x, src0, q0, mu, lam, blk, scale=-1.0, beta=1.0)
                    eig.index_add_(0, Pt, blk)
            else:
                blk = nd.contract(1, xn, el, sp["t"], q2, coeffs=c)
                dv.t_disp(x, src0, q0, c, blk, scale=-1.0, beta=1.0)
                dbl.index_add_(0, Pt, blk)
        if do_t:
            f0 = torch.from_numpy(np.ascontiguousarray(t0[S])).to(dev)
            blk = torch.zeros((P.size,) + shape[1:], dtype=F64, device=dev)
            if stress:
                dv.stress("force", x, src1, f0, mu, lam, blk, scale=1.0, beta=1.0)
                dv.stress("force", x, src0, f0, mu, lam, blk, scale=-1.0, beta=1.0)
            else:
                dv.u_disp(x, src1, f0, g, blk, scale=1.0, beta=1.0)
                dv.u_disp(x, src0, f0, g, blk, scale=-1.0, beta=1.0)
            sgl.index_add_(0, Pt, blk)
    return dbl, sgl, eig


def _eigen_near(ctx, points, patch, density, sigma, mu, lam, out):
    """+sigma * eigenstress over the near pairs (evaluate.py:298-326)."""
    src, tv, nrm, eps_arr = ctx.sources(patch)
    dens = np.asarray(density, float).reshape(-1, 3)
    live = np.flatnonzero((eps_arr > 0.0) & np.any(dens != 0.0, axis=1))
    if live.size == 0 or points.shape[0] == 0:
        return
    cen = tv.mean(axis=1)
    radius = defaults.EIGEN_NEAR_RADIUS_EPS * (eps_arr + kb.element_sizes(patch.mesh))
    chunks, pt_ptr, pt_chunk = near_list(points, cen[live], radius[live])
    if not chunks:
        return
    ch_src = np.concatenate([live[S] for S in chunks]).astype(np.int64)
    ch_ptr = np.concatenate([[0], np.cumsum([S.size for S in chunks])]).astype(np.int64)
    dv.eigen_near(ctx.x, src, ctx.dens(dens), pt_ptr, pt_chunk, ch_ptr, ch_src, mu, lam, out,
                  scale=sigma, beta=1.0)


def evaluate_stress(model: RegionModel, region, solution: dict, points, eps,
                    subtract_anelastic: bool = True, warn_near: bool = True,
                    parts: bool = False, near=None, near_list: bool = True):
    """Stress (N,3,3) at ``points`` of ``region`` (evaluate.py:598-695);
    ``parts`` also returns the eigenstress removed.  ``near`` as in
    ``evaluate_displacement``: the near pairs of every boundary patch read
    from the recovered displacement, each double layer with its own
    eigenstress."""
    region = _region(model, region)
    if region.faults:
        from .selfcheck import ensure_fault_convention
        ensure_fault_convention("stress")
    points = np.ascontiguousarray(np.asarray(points, dtype=float).reshape(-1, 3))
    if warn_near and near is None:
        _warn_near_boundary(points, region)
    ctx = _Ctx(points, eps)
    N = points.shape[0]
    sig = torch.zeros((N, 3, 3), dtype=F64, device=ctx.dev)
    eig = torch.zeros((N, 3, 3), dtype=F64, device=ctx.dev) if parts else None
    mu, lam = region.material.mu, region.material.lam

    def double_layer(patch, jump, sigma):
        src = ctx.sources(patch)[0]
        d = ctx.dens(jump)
        dv.stress("dd", ctx.x, src, d, mu, lam, sig, scale=-sigma, beta=1.0)
        if subtract_anelastic or parts:
            target = eig if parts else sig
            if near_list:
                _eigen_near(ctx, points, patch, jump, sigma, mu, lam, target)
            else:
                dv.stress("eigen", ctx.x, src, d, mu, lam, target, scale=sigma, beta=1.0)

    if N:
        for p in region.patches:
            sigma = float(model.orientation(region, p))
            u_p = _density(model, p, "u", solution, region)
            if np.any(u_p):
                double_layer(p, u_p, sigma)
            t_p = _density(model, p, "t", solution)
            if np.any(t_p):
                dv.stress("force", ctx.x, ctx.sources(p)[0], ctx.dens(t_p), mu, lam, sig,
                          scale=sigma, beta=1.0)
            if near is not None and near.recovers(p):
                dbl, sgl, e_near = _near_field(points, p, u_p, near.recover(p, u_p), t_p, mu, lam,
                                               kb.resolve_patch_eps(eps, p), near, stress=True,
                                               eigen=subtract_anelastic or parts)
                sig += sigma * (sgl - dbl)
                if subtract_anelastic or parts:
                    (eig if parts else sig).add_(sigma * e_near)
        for f in region.faults:
            sigma = float(model.orientation(region, f))
            slip = _density(model, f, "slip", solution)
            if np.any(slip):
                double_layer(f, slip, sigma)
    if parts and subtract_anelastic:
        sig += eig
    torch.cuda.current_stream().synchronize()
    s = sig.cpu().numpy()
    return (s, eig.cpu().numpy()) if parts else s


def evaluate_fault_traction(model: RegionModel, region, solution: dict, fault, eps,
                            subtract_anelastic: bool = True) -> np.ndarray:
    """Elastic traction (n_dofs, 3) at a fault's collocation points: the
    ``evaluate_stress`` field there dotted with the collocation normals
    (evaluate.py:475-492).  ``fault`` is a fault of ``region`` or its name.
    On a prescribed-slip fault this is the on-fault stress readout of the
    slip."""
    region = _region(model, region)
    if isinstance(fault, str):
        fault = next((f for f in region.faults if f.name == fault), None)
        if fault is None:
            raise KeyError(f"region {region.name!r} has no fault of that name")
    sig = evaluate_stress(model, region, solution, fault.collocation_points(), eps,
                          subtract_anelastic=subtract_anelastic, warn_near=False)
    return np.einsum("nij,nj->ni", sig, fault.collocation_normals())

