"""Learned FGMRES preconditioner: a point-cloud network over the boundary
elements that proposes a low-rank correction of block-Jacobi.

    M(r) = BJ(r) + D Phi (Psi^T r) / n,   Phi, Psi: n x K  (K = 16)

with D the diagonal of the block-Jacobi inverse (so Phi is dimensionless
whatever the slot kind and the moduli).

One point per boundary triangle (its three unknowns are one slot triple).
The network sees geometry only — per point the centroid (scaled to the unit
box of the mesh), the normal, log(area / mean area) and the slot kind (u or
t, one-hot), zero-padded to 16 inputs — so M is linear in r and fixed for a
mesh: the network runs once per mesh, and each FGMRES application costs the
block-Jacobi apply plus two n x K products.  Network (width 128): a point
layer, one k-nearest-neighbour message step (k = 16, mean of transformed
neighbours), a layer on [h1 | m], a layer on [h3 | mean over all points of
h3] (global context), and a linear head with 6K outputs per point, the rows
of Phi and Psi at the point's three unknowns; ReLU between.

Training: residual/correction pairs from our own solves.  For a system
A x = b, restarted FGMRES cycles with block-Jacobi give iterates x_c; with
the converged x*, r_c = b - A x_c and e_c = x* - x_c satisfy A e_c = r_c.  The
network learns the part of e_c that block-Jacobi misses: with s = RMS(r),
r^ = r / (s n) and t^ = (e - BJ(r)) / (s D), the loss over a geometry's
pairs is (1/B) sum_b |Phi Psi^T r^_b - t^_b|^2 / |t^_b|^2.  Features, forward,
backward and Adam run in the HIP kernels of csrc/learned.hip (FP32);
inference runs the dense layers on the BF16 matrix cores (``mbem_hgemm``).
The network has no counterpart in the reference; block-Jacobi FGMRES
(preconditioner.py:314-543) is its fallback and the measure of its gain.
"""
from __future__ import annotations

import math
import os
import time
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import call, ptr, stream

F32 = torch.float32
F64 = torch.float64
I32 = torch.int32

WIDTH = 128
N_IN = 16
N_GEO = 9
K_NBR = 16
RANK = 16
N_OUT = 6 * RANK
SHAPES = [("W1", (WIDTH, N_IN)), ("b1", (WIDTH,)), ("W2", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
          ("W3", (WIDTH, 2 * WIDTH)), ("b3", (WIDTH,)), ("W4", (WIDTH, 2 * WIDTH)),
          ("b4", (WIDTH,)), ("W5", (N_OUT, WIDTH)), ("b5", (N_OUT,))]
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "learned_precond.npz")


def n_params() -> int:
    return int(sum(math.prod(s) for _, s in SHAPES))


def init_params(seed: int = 0) -> np.ndarray:
    """He-uniform weights, zero biases, a small head (the correction starts
    near zero, so an untrained network is close to block-Jacobi)."""
    rng = np.random.default_rng(seed)
    out = []
    for name, shp in SHAPES:
        if name.startswith("b"):
            out.append(np.zeros(shp))
        else:
            lim = math.sqrt(6.0 / shp[1]) * (0.1 if name == "W5" else 1.0)
            out.append(rng.uniform(-lim, lim, size=shp))
    return np.concatenate([a.reshape(-1) for a in out]).astype(np.float32)


# ---------------------------------------------------------------------------
class Cloud:
    """The points of one boundary mesh: geometry features, kNN graph and its
    reverse (for the backward pass), activation buffers."""

    def __init__(self, layout, dev=None):
        dev = dev or _lib.device()
        cen, nrm, area, kind = [], [], [], []
        pos = 0
        for s in layout.slots:
            if s.offset != pos:
                raise NotImplementedError("the learned preconditioner needs one slot per patch, "
                                          "in point order")
            m = s.patch.mesh
            v = m.vertices[m.triangles]
            cen.append(v.mean(axis=1))
            cr = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
            a = 0.5 * np.linalg.norm(cr, axis=1)
            nrm.append(cr / np.maximum(2 * a, 1e-300)[:, None])
            area.append(a)
            kind.append(np.full(m.n_triangles, 1.0 if s.kind == "u" else 0.0))
            pos = s.stop
        if pos != layout.n_unknowns:
            raise NotImplementedError("unknowns beyond the boundary slots")
        cen, nrm, area, kind = map(np.concatenate, (cen, nrm, area, kind))
        lo, hi = cen.min(0), cen.max(0)
        mid, half = 0.5 * (lo + hi), 0.5 * float(np.linalg.norm(hi - lo)) + 1e-30
        geo = np.column_stack([(cen - mid) / half, nrm, np.log(area / area.mean()), kind,
                               1.0 - kind]).astype(np.float32)
        self.P = geo.shape[0]
        self.dev = dev
        self.geo = torch.as_tensor(geo, device=dev).contiguous()
        feats = np.zeros((self.P, N_IN), dtype=np.float32)
        feats[:, :N_GEO] = geo
        pts = torch.as_tensor(cen.astype(np.float32), device=dev).contiguous()
        self.nbr = torch.empty((self.P, K_NBR), dtype=I32, device=dev)
        call("mbem_knn", ptr(pts), self.P, K_NBR, ptr(self.nbr), stream())
        nb = self.nbr.cpu().numpy().astype(np.int64)
        order = np.argsort(nb.reshape(-1), kind="stable")
        src = np.repeat(np.arange(self.P), K_NBR)[order]
        counts = np.bincount(nb.reshape(-1), minlength=self.P)
        rptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
        self.rptr = torch.as_tensor(rptr, device=dev)
        self.ridx = torch.as_tensor(src.astype(np.int32), device=dev)
        P = self.P
        e = dict(dtype=F32, device=dev)
        self.X = torch.as_tensor(feats, device=dev).contiguous()
        self.h1 = torch.zeros((P, WIDTH), **e)
        self.a = torch.zeros((P, WIDTH), **e)
        self.m = torch.zeros((P, WIDTH), **e)
        self.c1 = torch.zeros((P, 2 * WIDTH), **e)
        self.h3 = torch.zeros((P, WIDTH), **e)
        self.h4 = torch.zeros((P, WIDTH), **e)
        self.y = torch.zeros((P, N_OUT), **e)


class PointNet:
    """Parameters (one flat FP32 device vector), gradients, Adam state, and the
    forward / backward / inference kernel sequences."""

    def __init__(self, theta=None, dev=None):
        _lib.require_gpu()
        dev = dev or _lib.device()
        self.dev = dev
        theta = init_params() if theta is None else np.asarray(theta, np.float32)
        if theta.size != n_params():
            raise ValueError(f"{theta.size} parameters, expected {n_params()}")
        self.theta = torch.as_tensor(theta, device=dev).contiguous()
        self.grad = torch.zeros_like(self.theta)
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.step = 0
        self.view = {}
        o = 0
        for name, shp in SHAPES:
            k = math.prod(shp)
            self.view[name] = (self.theta[o:o + k].view(shp), self.grad[o:o + k].view(shp))
            o += k
        self.W4a = None
        self._work = torch.zeros(1, dtype=F32, device=dev)
        self._g = torch.zeros(WIDTH, dtype=F32, device=dev)
        self._u = torch.zeros(WIDTH, dtype=F32, device=dev)
        self._dc = torch.zeros(WIDTH, dtype=F32, device=dev)
        self._dg = torch.zeros(WIDTH, dtype=F32, device=dev)
        self._loss = torch.zeros(1, dtype=F32, device=dev)

    @classmethod
    def load(cls, path=WEIGHTS, dev=None):
        with np.load(path, allow_pickle=False) as z:
            return cls(z["theta"], dev=dev)

    def save(self, path, **meta):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, theta=self.theta.cpu().numpy(),
                 **{k: np.asarray(v) for k, v in meta.items()})

    def W(self, name):
        return self.view[name][0]

    def G(self, name):
        return self.view[name][1]

    def _ws(self, n):
        if self._work.numel() < n:
            self._work = torch.zeros(n, dtype=F32, device=self.dev)
        return self._work

    # -- dense layer helpers ----------------------------------------------------
    def _linear(self, X, in_dim, W, bias, out, relu, ld_out=None, beta=0.0):
        """out (P x out) = act(X W^T + bias) in FP32 (training path)."""
        P = X.shape[0]
        nout = W.shape[0]
        call("mbem_sgemm", ptr(X), X.stride(0), 1, ptr(W), 1, in_dim, ptr(out),
             ld_out or out.stride(0), P, nout, in_dim, 1.0, float(beta), ptr(bias), int(relu), 1,
             None, stream())

    def _linear_bf16(self, X, in_dim, W, bias, out, relu):
        P = X.shape[0]
        call("mbem_hgemm", ptr(X), X.stride(0), ptr(W), in_dim, W.shape[0], P, ptr(bias), int(relu),
             ptr(out), out.stride(0), stream())

    # -- forward ---------------------------------------------------------------------
    def forward(self, cl: Cloud, bf16: bool = False):
        """cl.X -> cl.y, activations kept in the cloud's buffers."""
        lin = self._linear_bf16 if bf16 else self._linear
        W4 = self.W("W4")
        lin(cl.X, N_IN, self.W("W1"), self.W("b1"), cl.h1, True)
        lin(cl.h1, WIDTH, self.W("W2"), self.W("b2"), cl.a, True)
        call("mbem_knn_mean", ptr(cl.a), cl.P, WIDTH, ptr(cl.nbr), K_NBR, ptr(cl.m), stream())
        call("mbem_concat", ptr(cl.h1), WIDTH, ptr(cl.m), WIDTH, cl.P, ptr(cl.c1), stream())
        lin(cl.c1, 2 * WIDTH, self.W("W3"), self.W("b3"), cl.h3, True)
        # g = mean_i h3_i; u = W4b g + b4 (the global half of layer 4 as a bias)
        P = cl.P
        ns = min(256, max(1, P // 64))
        call("mbem_colsum", ptr(cl.h3), P, WIDTH, ptr(self._g), ptr(self._ws(ns * WIDTH)), ns, 0.0,
             stream())
        call("mbem_sgemm", ptr(W4[:, WIDTH:]), 2 * WIDTH, 1, ptr(self._g), 1, 1, ptr(self._u), 1,
             WIDTH, 1, WIDTH, 1.0 / P, 0.0, None, 0, 1, None, stream())
        call("mbem_add_row", ptr(self._u), 1, WIDTH, ptr(self.W("b4")), 1.0, stream())
        # layer 4 on the local half: a strided view of W4 (ld 2W) through the FP32
        # GEMM; the BF16 path takes a contiguous copy of W4a
        if bf16:
            if self.W4a is None:
                self.W4a = W4[:, :WIDTH].contiguous()
            self._linear_bf16(cl.h3, WIDTH, self.W4a, self._u, cl.h4, True)
        else:
            call("mbem_sgemm", ptr(cl.h3), WIDTH, 1, ptr(W4), 1, 2 * WIDTH, ptr(cl.h4), WIDTH, P,
                 WIDTH, WIDTH, 1.0, 0.0, ptr(self._u), 1, 1, None, stream())
        lin(cl.h4, WIDTH, self.W("W5"), self.W("b5"), cl.y, False)

    # -- backward ------------------------------------------------------------------------
    def backward(self, cl: Cloud, dy):
        """Accumulate d loss / d theta into self.grad (dy: P x 3)."""
        P = cl.P
        W4 = self.W("W4")
        e = dict(dtype=F32, device=self.dev)
        splits = max(1, min(64, P // 256))
        ws = self._ws(splits * WIDTH * 2 * WIDTH)

        def wgrad(dY, Xin, in_dim, G, nout):
            # G (nout x in_dim) += dY^T Xin   (split-K over the points)
            call("mbem_sgemm", ptr(dY), 1, dY.stride(0), ptr(Xin), Xin.stride(0), 1, ptr(G), in_dim,
                 nout, in_dim, P, 1.0, 1.0, None, 0, splits, ptr(ws), stream())

        def bgrad(dY, Gb):
            ns = min(256, max(1, P // 64))
            tmp = self._ws2(ns * dY.shape[1])
            call("mbem_colsum", ptr(dY), P, dY.shape[1], ptr(Gb), ptr(tmp), ns, 1.0, stream())

        # head
        nout = dy.shape[1]
        wgrad(dy, cl.h4, WIDTH, self.G("W5"), nout)
        bgrad(dy, self.G("b5"))
        dh4 = torch.empty((P, WIDTH), **e)
        call("mbem_sgemm", ptr(dy), nout, 1, ptr(self.W("W5")), WIDTH, 1, ptr(dh4), WIDTH, P, WIDTH,
             nout, 1.0, 0.0, None, 0, 1, None, stream())
        call("mbem_relu_bwd", ptr(cl.h4), ptr(dh4), P * WIDTH, stream())
        # layer 4: local half, bias, global half (dW4b = dc g^T / ..., dg = W4b^T dc)
        G4 = self.G("W4")
        call("mbem_sgemm", ptr(dh4), 1, WIDTH, ptr(cl.h3), WIDTH, 1, ptr(G4), 2 * WIDTH, WIDTH,
             WIDTH, P, 1.0, 1.0, None, 0, 1, None, stream())
        ns = min(256, max(1, P // 64))
        call("mbem_colsum", ptr(dh4), P, WIDTH, ptr(self._dc), ptr(self._ws2(ns * WIDTH)), ns, 0.0,
             stream())
        call("mbem_add_row", ptr(self.G("b4")), 1, WIDTH, ptr(self._dc), 1.0, stream())
        # dW4b[i, j] += dc_i * gmean_j, gmean = g / P (self._g holds the column sum)
        call("mbem_sgemm", ptr(self._dc), 1, 1, ptr(self._g), 1, 1, ptr(G4[:, WIDTH:]), 2 * WIDTH,
             WIDTH, WIDTH, 1, 1.0 / P, 1.0, None, 0, 1, None, stream())
        # dg = W4b^T dc (per point share 1/P)
        call("mbem_sgemm", ptr(W4[:, WIDTH:]), 1, 2 * WIDTH, ptr(self._dc), 1, 1, ptr(self._dg), 1,
             WIDTH, 1, WIDTH, 1.0 / P, 0.0, None, 0, 1, None, stream())
        dh3 = torch.empty((P, WIDTH), **e)
        call("mbem_sgemm", ptr(dh4), WIDTH, 1, ptr(W4), 2 * WIDTH, 1, ptr(dh3), WIDTH, P, WIDTH, WIDTH,
             1.0, 0.0, None, 0, 1, None, stream())
        call("mbem_add_row", ptr(dh3), P, WIDTH, ptr(self._dg), 1.0, stream())
        call("mbem_relu_bwd", ptr(cl.h3), ptr(dh3), P * WIDTH, stream())
        # layer 3 on [h1 | m]
        wgrad(dh3, cl.c1, 2 * WIDTH, self.G("W3"), WIDTH)
        bgrad(dh3, self.G("b3"))
        dc1 = torch.empty((P, 2 * WIDTH), **e)
        call("mbem_sgemm", ptr(dh3), WIDTH, 1, ptr(self.W("W3")), 2 * WIDTH, 1, ptr(dc1), 2 * WIDTH, P,
             2 * WIDTH, WIDTH, 1.0, 0.0, None, 0, 1, None, stream())
        dh1 = torch.empty((P, WIDTH), **e)
        dm = torch.empty((P, WIDTH), **e)
        call("mbem_split", ptr(dc1), WIDTH, WIDTH, P, ptr(dh1), 0, ptr(dm), stream())
        # message step
        da = torch.empty((P, WIDTH), **e)
        call("mbem_knn_mean_bwd", ptr(dm), P, WIDTH, ptr(cl.rptr), ptr(cl.ridx), K_NBR, ptr(da),
             stream())
        call("mbem_relu_bwd", ptr(cl.a), ptr(da), P * WIDTH, stream())
        wgrad(da, cl.h1, WIDTH, self.G("W2"), WIDTH)
        bgrad(da, self.G("b2"))
        call("mbem_sgemm", ptr(da), WIDTH, 1, ptr(self.W("W2")), WIDTH, 1, ptr(dh1), WIDTH, P, WIDTH,
             WIDTH, 1.0, 1.0, None, 0, 1, None, stream())
        call("mbem_relu_bwd", ptr(cl.h1), ptr(dh1), P * WIDTH, stream())
        # layer 1
        wgrad(dh1, cl.X, N_IN, self.G("W1"), WIDTH)
        bgrad(dh1, self.G("b1"))

    def _ws2(self, n):
        if not hasattr(self, "_work2") or self._work2.numel() < n:
            self._work2 = torch.zeros(max(n, 1), dtype=F32, device=self.dev)
        return self._work2

    def loss_grad(self, cl: Cloud, tgt, dy, s: float = 1.0):
        """dy = d(|s y - t|^2/|t|^2)/dy over the whole output; the loss value
        stays on the device."""
        n = cl.y.numel()
        part = self._ws2(2 * ((n + 255) // 256))
        call("mbem_loss", ptr(cl.y), ptr(tgt), n, float(s), ptr(part), ptr(dy), ptr(self._loss),
             stream())
        return self._loss

    def lowrank_loss_grad(self, cl: Cloud, R, T, dy):
        """Batch loss (1/B) sum_b |Phi Psi^T r_b - t_b|^2 / |t_b|^2 of the
        output cl.y and its gradient dy (P x 6K); R, T: B x 3P FP32."""
        B, n = R.shape
        P, K = cl.P, RANK
        ns = min(256, max(1, P // 64))
        part = self._part(ns * B * K)
        C = torch.empty((B, K), dtype=F64, device=self.dev)
        dC = torch.empty((B, K), dtype=F64, device=self.dev)
        pred = torch.empty((B, n), dtype=F32, device=self.dev)
        dpred = torch.empty((B, n), dtype=F32, device=self.dev)
        lb = torch.empty(B, dtype=F32, device=self.dev)
        call("mbem_lr_coef", ptr(cl.y), P, K, 3 * K, ptr(R), n, B, 0, ptr(part), ns, ptr(C), stream())
        call("mbem_lr_expand", ptr(cl.y), P, K, 0, ptr(C), B, ptr(pred), n, 0, 1.0, 0, None, stream())
        call("mbem_lr_loss", ptr(pred), ptr(T), n, B, ptr(dpred), ptr(lb), stream())
        call("mbem_lr_outer", ptr(dpred), n, ptr(C), B, P, K, 0, ptr(dy), stream())
        call("mbem_lr_coef", ptr(cl.y), P, K, 0, ptr(dpred), n, B, 0, ptr(part), ns, ptr(dC),
             stream())
        call("mbem_lr_outer", ptr(R), n, ptr(dC), B, P, K, 3 * K, ptr(dy), stream())
        return lb.sum()

    def _part(self, n):
        if not hasattr(self, "_part_t") or self._part_t.numel() < n:
            self._part_t = torch.zeros(max(n, 1), dtype=F64, device=self.dev)
        return self._part_t

    def zero_grad(self):
        self.grad.zero_()

    def adam(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
        self.step += 1
        t = self.step
        lr_t = lr * math.sqrt(1 - b2 ** t) / (1 - b1 ** t)
        call("mbem_adam", ptr(self.theta), ptr(self.grad), ptr(self.m), ptr(self.v),
             self.theta.numel(), float(lr_t), float(b1), float(b2), float(eps), float(wd), stream())
        self.W4a = None


# ---------------------------------------------------------------------------
class LearnedPreconditioner:
    """z = BJ(r) + D Phi (Psi^T r) / n: a linear right preconditioner for
    FGMRES (``apply_into(r, z)``).  The network runs once, on the BF16 matrix
    cores (or FP32), when the preconditioner is made; each application adds
    two n x K products to the block-Jacobi apply."""

    def __init__(self, base, layout, net: PointNet = None, bf16: bool = True):
        self.base = base
        self.cloud = Cloud(layout)
        self.net = net or PointNet.load()
        self.bf16 = bf16
        t0 = time.perf_counter()
        self.net.forward(self.cloud, bf16=bf16)
        self.Y = self.cloud.y.clone()
        self.D = base.inverse_diagonal().abs().clamp_min(1e-300)
        torch.cuda.current_stream().synchronize()
        self.infer_s = time.perf_counter() - t0
        P = self.cloud.P
        self.ns = min(256, max(1, P // 64))
        self._part = torch.zeros(self.ns * RANK, dtype=F64, device=self.cloud.dev)
        self._c = torch.zeros(RANK, dtype=F64, device=self.cloud.dev)
        self.applies = 0

    def apply_into(self, r, z):
        self.base.apply_into(r, z)
        P, K = self.cloud.P, RANK
        call("mbem_lr_coef", ptr(self.Y), P, K, 3 * K, ptr(r), 3 * P, 1, 1, ptr(self._part), self.ns,
             ptr(self._c), stream())
        call("mbem_lr_expand", ptr(self.Y), P, K, 0, ptr(self._c), 1, ptr(z), 3 * P, 1, 1.0 / (3 * P),
             1, ptr(self.D), stream())
        self.applies += 1
        return z

    def __call__(self, r):
        z = torch.empty_like(r)
        return self.apply_into(r, z)

    def summary(self):
        s = dict(self.base.summary()) if hasattr(self.base, "summary") else {}
        s.update(learned=True, points=self.cloud.P, rank=RANK, bf16=self.bf16,
                 inference_s=self.infer_s)
        return s


# ---------------------------------------------------------------------------
# training data and the training loop

@dataclass
class Sample:
    cloud: Cloud
    R: torch.Tensor          # B x 3P  normalised residuals r / |r|
    T: torch.Tensor          # B x 3P  targets (e - BJ(r)) / |r|


def residual_pairs(asm, M, cycles: int = 8, restart: int = 10, rtol: float = 1e-12):
    """[(r, e)] device FP64 pairs with A e = r from restarted FGMRES cycles of
    A x = b (x_0 = 0 included) and the converged solution."""
    from .la.solver import fgmres
    b = asm.b_dev
    xs = [torch.zeros_like(b)]
    x = xs[0]
    for _ in range(cycles):
        x, _ = fgmres(asm, b, M=M, x0=x, rtol=1e-14, restart=restart, maxiter=restart)
        xs.append(x.clone())
    xstar, rep = fgmres(asm, b, M=M, x0=x, rtol=rtol, restart=200, maxiter=600)
    pairs = []
    for xc in xs:
        r = b - asm.matvec(xc)
        pairs.append((r, xstar - xc))
    return pairs, rep


def make_samples(asm, M, cloud: Cloud, pairs):
    """One batch Sample of a geometry's (r, e) pairs (uses block-Jacobi M):
    r^ = r / (RMS(r) n), t^ = (e - BJ(r)) / (RMS(r) D)."""
    R, T = [], []
    D = M.inverse_diagonal().abs().clamp_min(1e-300)
    n = pairs[0][0].numel() if pairs else 0
    for r, e in pairs:
        rms = float(torch.linalg.vector_norm(r)) / math.sqrt(n)
        if rms == 0.0:
            continue
        z = torch.empty_like(r)
        M.apply_into(r, z)
        R.append((r / (rms * n)).float())
        T.append(((e - z) / (rms * D)).float())
    return [Sample(cloud, torch.stack(R).contiguous(), torch.stack(T).contiguous())]


def train(net: PointNet, samples, epochs: int = 10, lr: float = 1e-3, seed: int = 0,
          log=print, lr_final: float = 0.1):
    """Adam over the samples (one step per geometry batch, shuffled per
    epoch), rate on a cosine from lr to lr * lr_final; returns the mean loss
    per epoch."""
    rng = np.random.default_rng(seed)
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(samples))
        tot = torch.zeros(1, dtype=F32, device=net.dev)
        t0 = time.perf_counter()
        for i in order:
            s = samples[i]
            cl = s.cloud
            net.forward(cl)
            dy = torch.empty_like(cl.y)
            loss = net.lowrank_loss_grad(cl, s.R, s.T, dy)
            tot += loss
            net.zero_grad()
            net.backward(cl, dy)
            frac = ep / max(epochs - 1, 1)
            net.adam(lr=lr * (lr_final + (1 - lr_final) * 0.5 * (1 + math.cos(math.pi * frac))))
        mean = float(tot.item()) / max(len(samples), 1)
        hist.append(mean)
        log(f"[learned] epoch {ep}: mean loss {mean:.4f} ({time.perf_counter() - t0:.1f} s)")
    return hist


def evaluate_loss(net: PointNet, samples) -> float:
    tot = 0.0
    for s in samples:
        cl = s.cloud
        net.forward(cl)
        dy = torch.empty_like(cl.y)
        tot += float(net.lowrank_loss_grad(cl, s.R, s.T, dy).item())
    return tot / max(len(samples), 1)
