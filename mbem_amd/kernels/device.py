"""Device-resident front end of the pair kernels (torch tensors in HBM).

Thin wrappers over the C ABI: they check shapes and dtypes on the host,
then enqueue on torch's current stream.  Every tensor is float64, contiguous
and on the current device; nothing here copies to the host.
"""
from __future__ import annotations

import torch

from .. import _lib
from .._lib import call, host_doubles, ptr, stream

F64 = torch.float64


def _dev(a, shape_tail=None, name="array"):
    if not isinstance(a, torch.Tensor):
        a = torch.as_tensor(a, dtype=F64)
    if a.dtype != F64:
        raise TypeError(f"{name}: expected float64, got {a.dtype}")
    if a.device.type != "cuda":
        a = a.to(_lib.device())
    a = a.contiguous()
    if shape_tail is not None and tuple(a.shape[1:]) != tuple(shape_tail):
        raise ValueError(f"{name}: expected shape (N, {', '.join(map(str, shape_tail))}), "
                         f"got {tuple(a.shape)}")
    return a


class Sources:
    """Packed source records (include/mbem_hip.h, mbem_prepare_sources):
    column s of a (24, ld) float64 device array."""

    def __init__(self, tri_verts, normals=None, eps=None):
        _lib.require_gpu()
        tv = _dev(tri_verts, (3, 3), "tri_verts")
        ns = tv.shape[0]
        if normals is None:
            normals = torch.zeros((ns, 3), dtype=F64, device=tv.device)
        nrm = _dev(normals, (3,), "normals")
        e = _dev(eps, None, "eps").reshape(-1)
        if nrm.shape[0] != ns or e.shape[0] != ns:
            raise ValueError(f"sources: {ns} triangles but {nrm.shape[0]} normals and "
                             f"{e.shape[0]} eps")
        self.ns = ns
        self.ld = max(ns, 1)
        self.buf = torch.empty((_lib.SRC_FIELDS, self.ld), dtype=F64, device=tv.device)
        self.eps = e
        self.tri_verts = tv
        self.normals = nrm
        call("mbem_prepare_sources", ptr(tv), ptr(nrm), ptr(e), ns, ptr(self.buf), self.ld,
             stream())

    def subset(self, idx):
        """Records of the sources ``idx`` (a device or host index array)."""
        idx = torch.as_tensor(idx, device=self.tri_verts.device, dtype=torch.long)
        return Sources(self.tri_verts[idx], self.normals[idx], self.eps[idx])


def u_block(x_field, srcs: Sources, g, out, ld_out, scale=1.0, beta=0.0):
    x = _dev(x_field, (3,), "x_field")
    call("mbem_u_block", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns,
         host_doubles(g), ptr(out), int(ld_out), float(scale), float(beta), stream())


def t_block(x_field, srcs: Sources, c, out, ld_out, scale=1.0, beta=0.0, rowsum=None,
            rs_scale=0.0, store=True):
    x = _dev(x_field, (3,), "x_field")
    call("mbem_t_block", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns,
         host_doubles(c), ptr(out if store else None), int(ld_out) if store else 3 * srcs.ns,
         float(scale), float(beta), ptr(rowsum), float(rs_scale), int(bool(store)), stream())


def u_basis(x_field, srcs: Sources, out):
    x = _dev(x_field, (3,), "x_field")
    nb, rows, cols = out.shape
    call("mbem_u_basis", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns, ptr(out), cols,
         rows * cols, stream())


def t_basis(x_field, srcs: Sources, out):
    x = _dev(x_field, (3,), "x_field")
    nb, rows, cols = out.shape
    call("mbem_t_basis", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns, ptr(out), cols,
         rows * cols, stream())


def u_disp(x_field, srcs: Sources, density, g, out, scale=1.0, beta=0.0):
    x = _dev(x_field, (3,), "x_field")
    d = _dev(density, (3,), "density")
    call("mbem_u_disp_contract", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns, ptr(d),
         host_doubles(g), ptr(out), float(scale), float(beta), stream())


def t_disp(x_field, srcs: Sources, density, c, out, scale=1.0, beta=0.0):
    x = _dev(x_field, (3,), "x_field")
    d = _dev(density, (3,), "density")
    call("mbem_t_disp_contract", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns, ptr(d),
         host_doubles(c), ptr(out), float(scale), float(beta), stream())


def stress(kind: str, x_field, srcs: Sources, density, mu, lam, out, scale=1.0, beta=0.0):
    """kind: "force" (Kelvin), "dd" (total dislocation stress) or "eigen"."""
    name = {"force": "mbem_kelvin_stress_contract", "dd": "mbem_dd_stress_contract",
            "eigen": "mbem_eigenstress_contract"}[kind]
    x = _dev(x_field, (3,), "x_field")
    d = _dev(density, (3,), "density")
    call(name, ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns, ptr(d), float(mu),
         float(lam), ptr(out), float(scale), float(beta), stream())


def eigen_near(x_field, srcs: Sources, density, pt_ptr, pt_chunk, ch_ptr, ch_src, mu, lam, out,
               scale=1.0, beta=0.0):
    x = _dev(x_field, (3,), "x_field")
    d = _dev(density, (3,), "density")
    dev = x.device
    ts = [torch.as_tensor(a, dtype=torch.int64, device=dev).contiguous()
          for a in (pt_ptr, pt_chunk, ch_ptr, ch_src)]
    call("mbem_eigenstress_near", ptr(x), x.shape[0], ptr(srcs.buf), srcs.ld, srcs.ns, ptr(d),
         ptr(ts[0]), ptr(ts[1]), ptr(ts[2]), ptr(ts[3]), float(mu), float(lam), ptr(out),
         float(scale), float(beta), stream())


def add_block_diag(A_block, lda, D, scale=1.0):
    Dd = _dev(D, (3, 3), "D")
    call("mbem_add_block_diag", ptr(A_block), int(lda), Dd.shape[0], ptr(Dd), float(scale),
         stream())
