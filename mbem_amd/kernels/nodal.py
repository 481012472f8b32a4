"""Device side of the Lagrange P1/P2 nodal kernels (csrc/nodal.hip): the
kernel spec each launch consumes and the calls into libmbem_hip.

The spec of a (order, kernel) pair (tri_nodal.py:77-221): which monomial
degrees of which 1/R^n moment tables the kernel's contraction reads (at
order p every degree grows by p, because the shape functions multiply the
integrand by a degree-p polynomial), the lower tables the divergence-theorem
recursion consumes on the way up, the edge-primitive degrees those tables
need, the series coefficients C(-m/2, j) of the small- and large-|u| edge
series, the node lattice, and the two far-field Gauss rules with the shape
function values at their points.  Built once per (order, kernel) on the
host, uploaded once per device.
"""
from __future__ import annotations

import ctypes
from fractions import Fraction

import numpy as np
import torch

from .. import _lib, defaults
from .._lib import call, ptr, stream

F64 = torch.float64

# the moment tables each kernel's node contraction reads at order 0:
# (n, highest monomial degree of xi^a / R^n) -- tri_nodal.py:77-93
KT, KU, KH, KS, KE = 1, 2, 4, 8, 16
_READS = {KT: ((3, 1), (5, 3)),            # slip -> displacement
          KU: ((1, 0), (3, 2)),            # force -> displacement
          KH: ((3, 0), (5, 2), (7, 4)),    # slip -> total stress
          KS: ((3, 1), (5, 3)),            # force -> stress
          KE: ((7, 0),)}                   # eigenstress
_N_ROWS = (-1, 1, 3, 5, 7)
_M_ROWS = (-1, 1, 3, 5)
_MAX_EDGE = 5
_MAX_TABLE = 6


def n_nodes(order: int) -> int:
    if order not in (0, 1, 2):
        raise ValueError(f"order must be 0, 1 or 2 (got {order!r})")
    return (order + 1) * (order + 2) // 2


def node_lattice(order: int) -> np.ndarray:
    """Barycentric exponents of the nodes (K, 3): vertices, then the edges
    12, 23, 31, then the interior (tri_nodal.py:58-74)."""
    if order == 0:
        return np.zeros((1, 3), np.int64)
    rows = [[order if c == v else 0 for c in range(3)] for v in range(3)]
    for i, j in ((0, 1), (1, 2), (2, 0)):
        for s in range(1, order):
            r = [0, 0, 0]
            r[i], r[j] = order - s, s
            rows.append(r)
    rows += [[a, b, order - a - b] for a in range(1, order) for b in range(1, order - a)
             if order - a - b >= 1]
    return np.array(rows, np.int64)


def _reads(order, flags):
    out = {}
    for f, reads in _READS.items():
        if flags & f:
            for n, d in reads:
                out[n] = max(out.get(n, -1), d + order)
    return out


def table_degrees(order: int, flags: int) -> dict:
    """{n: degree} of every moment table the recursion fills (tri_nodal.py:96-115)."""
    deg = _reads(order, flags)
    # going down: the horizontal step to degree d at n reads degree d - 2 at n - 2
    n = max(deg)
    while True:
        if deg.get(n, -1) >= 2:
            deg[n - 2] = max(deg.get(n - 2, -1), deg[n] - 2)
        if n - 2 < min(deg):
            break
        n -= 2
    # the vertical seeds: I3 from the solid angle, up to I5, I7, down to I1, I_-1
    deg[3] = max(deg.get(3, -1), 0)
    for n in list(deg):
        lo_hi = range(3, n, 2) if n >= 5 else (range(n, 4, 2) if n <= 1 else ())
        for m in lo_hi:
            deg[m] = max(deg.get(m, -1), 0)
    return {n: d for n, d in deg.items() if d >= 0}


def edge_degrees(deg: dict) -> dict:
    """{m: highest u power} of the edge primitives int u^k / R^m the tables
    read (tri_nodal.py:118-128)."""
    edge = {}
    for n, d in deg.items():
        if d >= 1:
            edge[n - 2] = max(edge.get(n - 2, -1), d - 1)
        if n >= 5:
            edge[n - 2] = max(edge.get(n - 2, -1), 0)
        if n <= 1:
            edge[n] = max(edge.get(n, -1), 0)
    return edge


def series_coefficients(n_terms: int) -> np.ndarray:
    """C(-m/2, j) for m = -1, 1, 3, 5 and j < n_terms, exact rationals rounded once."""
    out = np.empty((len(_M_ROWS), n_terms))
    for r, m in enumerate(_M_ROWS):
        c = Fraction(1)
        for j in range(n_terms):
            out[r, j] = float(c)
            c = c * (Fraction(-m, 2) - j) / (j + 1)
    return out


def _collapsed_gauss(n: int):
    """Collapsed product Gauss-Legendre rule on (0,0), (1,0), (0,1): points
    (xi1, xi2) and weights summing to 1/2 (tri_nodal.py:143-153)."""
    g, w = np.polynomial.legendre.leggauss(n)
    g, w = 0.5 * (g + 1.0), 0.5 * w
    a = np.repeat(g, n)
    b = np.tile(g, n)
    return a, b * (1.0 - a), np.repeat(w, n) * np.tile(w, n) * (1.0 - a)


def shape_values(order: int, lam: np.ndarray) -> np.ndarray:
    """N_k at barycentric points lam (Q, 3) by the lattice product formula."""
    lat = node_lattice(order)
    out = np.ones((lam.shape[0], lat.shape[0]))
    for k, alpha in enumerate(lat):
        for i in range(3):
            for j in range(int(alpha[i])):
                out[:, k] *= (order * lam[:, i] - j) / (j + 1)
    return out


class Spec:
    """Host record and device tables of one (order, kernel) spec."""

    def __init__(self, order: int, flags: int):
        self.order, self.flags = int(order), int(flags)
        self.K = n_nodes(order)
        deg = table_degrees(order, flags)
        edge = edge_degrees(deg)
        need0 = _reads(0, flags)
        if deg.get(-1, -1) >= 1 or max(deg.values()) > _MAX_TABLE or max(edge.values()) > _MAX_EDGE:
            raise ValueError(f"order {order}, kernel {flags}: moment degrees beyond P2")
        rules = [_collapsed_gauss(n) for n in (defaults.NODAL_FAR_GAUSS_N,
                                               defaults.NODAL_FAR_GAUSS_N_DISTANT)]
        lam = np.concatenate([np.column_stack([1.0 - a - b, a, b]) for a, b, _ in rules])
        w = np.concatenate([r[2] for r in rules])
        nq = w.size
        nt = defaults.NODAL_SERIES_TERMS
        self.ispec = np.full(18, -1, np.int64)
        self.ispec[:5] = [self.order, self.K, rules[0][2].size, nt, nq]
        for i, n in enumerate(_N_ROWS):
            self.ispec[5 + i] = deg.get(n, -1)
        for i, m in enumerate(_M_ROWS):
            self.ispec[10 + i] = edge.get(m, -1)
        for i, n in enumerate((1, 3, 5, 7)):
            self.ispec[14 + i] = need0.get(n, -1)
        self.fspec = np.array([defaults.NODAL_D_STAR, defaults.NODAL_D_STAR_DISTANT,
                               defaults.NODAL_SMALL_U_OVER_RHO, defaults.NODAL_SERIES_U_OVER_RHO])
        tables = np.concatenate([series_coefficients(nt).ravel(),
                                 node_lattice(order).astype(np.float64).ravel(), lam.ravel(), w,
                                 shape_values(order, lam).ravel()])
        if tables.size != _lib.load().mbem_nodal_table_doubles(self.order, nt, nq):
            raise RuntimeError("nodal table layout mismatch")
        self._host_tables = tables
        self._dev = {}
        self._ispec_c = (ctypes.c_int64 * 18)(*[int(v) for v in self.ispec])
        self._fspec_c = (ctypes.c_double * 4)(*[float(v) for v in self.fspec])

    def tables(self, dev):
        key = str(dev)
        if key not in self._dev:
            self._dev[key] = torch.as_tensor(self._host_tables, dtype=F64, device=dev)
        return self._dev[key]

    def args(self, dev):
        return self._ispec_c, self._fspec_c, ptr(self.tables(dev))


_SPECS: dict = {}


def spec(order: int, flags: int) -> Spec:
    key = (int(order), int(flags))
    if key not in _SPECS:
        _SPECS[key] = Spec(*key)
    return _SPECS[key]


def check_shape(tri_verts) -> None:
    """Refuse needles for a nodal density: height / L below
    NODAL_MIN_HEIGHT_OVER_L loses (L / height)^2 digits in the closed form
    (tri_nodal.py:1239-1254)."""
    tv = np.asarray(tri_verts, float)
    e = np.stack([tv[:, 1] - tv[:, 0], tv[:, 2] - tv[:, 1], tv[:, 0] - tv[:, 2]])
    L = np.linalg.norm(e, axis=2).max(axis=0)
    ratio = np.linalg.norm(np.cross(e[0], -e[2]), axis=1) / np.where(L > 0.0, L * L, 1.0)
    bad = ratio < defaults.NODAL_MIN_HEIGHT_OVER_L
    if bad.any():
        raise ValueError(
            f"{int(bad.sum())} source triangle(s) have height / L < "
            f"{defaults.NODAL_MIN_HEIGHT_OVER_L:g} (min {ratio.min():.2e}): "
            "a nodal (order >= 1) density is not accurate on needles")


class Elements:
    """Source triangles on the device with their nodal frames."""

    def __init__(self, tri_verts, normals, eps_arr, dev=None):
        dev = dev or _lib.device()
        tv = np.ascontiguousarray(np.asarray(tri_verts, np.float64))
        if tv.ndim != 3 or tv.shape[1:] != (3, 3):
            raise ValueError(f"tri_verts: expected shape (N, 3, 3), got {tv.shape}")
        self.ns = tv.shape[0]
        self.tv = torch.as_tensor(tv, device=dev).reshape(self.ns, 9).contiguous()
        nrm = np.zeros((self.ns, 3)) if normals is None else np.asarray(normals, np.float64)
        if nrm.shape != (self.ns, 3):
            raise ValueError(f"normals: expected shape ({self.ns}, 3), got {nrm.shape}")
        self.nrm = torch.as_tensor(np.ascontiguousarray(nrm), device=dev)
        e = np.ascontiguousarray(np.asarray(eps_arr, np.float64).reshape(-1))
        if e.size != self.ns:
            raise ValueError(f"eps_arr: expected {self.ns} values, got {e.size}")
        self.eps = torch.as_tensor(e, device=dev)
        self.geo = torch.zeros(max(1, _lib.load().mbem_nodal_geo_doubles(self.ns)), dtype=F64,
                               device=dev)
        call("mbem_nodal_frames", ptr(self.tv), self.ns, ptr(self.geo), stream())


def _points(x, name="x_field"):
    x = np.ascontiguousarray(np.asarray(x, np.float64))
    if x.ndim != 2 or x.shape[1] != 3:
        raise ValueError(f"{name}: expected shape (N, 3), got {x.shape}")
    return torch.as_tensor(x, device=_lib.device())


def matrix(kind: int, x_field, el: Elements, sp: Spec, coeffs=None) -> torch.Tensor:
    """kind 0 (U) / 1 (T): the direct matrix (3 nf, 3 K ns) with the
    coefficients applied, or with ``coeffs=None`` the basis stack
    (3 or 6, 3 nf, 3 K ns)."""
    x = _points(x_field)
    nf, K = x.shape[0], sp.K
    nb = 3 if kind == 0 else 6
    shape = (3 * nf, 3 * K * el.ns)
    out = torch.zeros(shape if coeffs is not None else (nb,) + shape, dtype=F64, device=x.device)
    if nf and el.ns:
        c = None if coeffs is None else _lib.host_doubles(coeffs)
        ld = shape[1]
        call("mbem_nodal_matrix", kind, ptr(x), nf, ptr(el.tv), ptr(el.nrm), ptr(el.eps), el.ns,
             ptr(el.geo), *sp.args(x.device), c, ptr(out), ld, shape[0] * ld, stream())
    return out


def contract(kind: int, x_field, el: Elements, sp: Spec, density, coeffs=None, mu=1.0,
             lam=1.0) -> torch.Tensor:
    """kind 0 / 1: displacement (nf, 3) of a nodal force / slip density;
    2 / 3 / 4: force -> stress, slip -> total stress, eigenstress (nf, 3, 3).
    density (K ns, 3)."""
    x = _points(x_field)
    nf = x.shape[0]
    d = np.ascontiguousarray(np.asarray(density, np.float64))
    if d.shape != (sp.K * el.ns, 3):
        raise ValueError(f"density: expected shape ({sp.K * el.ns}, 3), got {d.shape}")
    dd = torch.as_tensor(d, device=x.device)
    no = 3 if kind <= 1 else 9
    out = torch.zeros((nf, no), dtype=F64, device=x.device)
    if nf:
        c = _lib.host_doubles(coeffs) if coeffs is not None else None
        nu = lam / (2.0 * (lam + mu))
        call("mbem_nodal_contract", kind, ptr(x), nf, ptr(el.tv), ptr(el.nrm), ptr(el.eps), el.ns,
             ptr(el.geo), *sp.args(x.device), ptr(dd), c, float(mu), float(nu), float(lam),
             ptr(out), stream())
    return out if kind <= 1 else out.reshape(nf, 3, 3)


def traction(which: int, x_field, n_field, el: Elements, sp: Spec, ids, n_cols: int, mu: float,
             lam: float) -> torch.Tensor:
    """Traction rows (3 nf, 3 n_cols): which 0 slip -> elastic, 1 force,
    2 eigenstress, 3 slip -> total; ``ids`` (ns, K) already range-checked."""
    x = _points(x_field)
    nfld = _points(n_field, "n_field")
    if nfld.shape != x.shape:
        raise ValueError(f"n_field: expected shape {tuple(x.shape)}, got {tuple(nfld.shape)}")
    nf = x.shape[0]
    idt = torch.as_tensor(np.ascontiguousarray(ids, np.int64), device=x.device)
    out = torch.zeros((3 * nf, 3 * int(n_cols)), dtype=F64, device=x.device)
    if nf and el.ns:
        nu = lam / (2.0 * (lam + mu))
        call("mbem_nodal_traction", which, ptr(x), ptr(nfld), nf, ptr(el.tv), ptr(el.nrm),
             ptr(el.eps), el.ns, ptr(el.geo), *sp.args(x.device), ptr(idt), int(n_cols), float(mu),
             float(nu), float(lam), ptr(out), 3 * int(n_cols), stream())
    return out
