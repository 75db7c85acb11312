"""CPU restatement of the flood surrogate (test infrastructure only).

The surrogate has no counterpart in the reference; this module restates what
the gfx950 kernels of floodca_amd/csrc/fca_nn.hip compute, in float64 numpy,
written independently of the kernels' formulation:

* convolutions by explicit im2col (zero padding 1) and a matrix product, and
  their input gradients by scattering the column gradient back (col2im) --
  the kernels instead gather the transposed convolution in place;
* with ``emulate=True`` the values the kernels round to bfloat16 (weights,
  stored activations, the up-sampling sums, GEMM operands of the backward
  pass) are rounded the same way (float32, then bfloat16 round to nearest
  even); with ``emulate=False`` nothing is rounded, which is the form the
  finite-difference tests check;
* the terrain pyramid, the network inputs and the training targets as loops
  over records and cells.

Parity pins: the backward pass against central finite differences of a
linear functional of the output (tests/test_surrogate_oracle.py); the
convolution against a direct seven-loop sum; the loss gradient against
finite differences of its two terms; the inputs and targets against
closed forms on uniform fields.
"""
from __future__ import annotations

import math

import numpy as np

from floodca_amd.nn import NetCfg, layers

LS_LO, LS_HI = -12.0, 6.0


def bf16(x):
    """float64 -> float32 -> bfloat16 (round to nearest even) -> float64."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x):
    """bfloat16 bits (uint16) of float values, round to nearest even."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32).astype(
        np.float64)


def im2col(X, S):
    """[B, H, W, C] -> [B, Ho, Wo, 9, C] of the 3x3 taps, zero padding 1."""
    B, H, W, C = X.shape
    Ho, Wo = H // S, W // S
    Xp = np.zeros((B, H + 2, W + 2, C), dtype=X.dtype)
    Xp[:, 1:-1, 1:-1] = X
    cols = np.empty((B, Ho, Wo, 9, C), dtype=X.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, 3 * ky + kx] = Xp[:, ky:ky + S * Ho:S, kx:kx + S * Wo:S]
    return cols


def col2im(dcol, H, W, S):
    """Adjoint of im2col: [B, Ho, Wo, 9, C] -> [B, H, W, C]."""
    B, Ho, Wo, _, C = dcol.shape
    dXp = np.zeros((B, H + 2, W + 2, C), dtype=dcol.dtype)
    for ky in range(3):
        for kx in range(3):
            dXp[:, ky:ky + S * Ho:S, kx:kx + S * Wo:S] += dcol[:, :, :, 3 * ky + kx]
    return dXp[:, 1:-1, 1:-1]


def conv_direct(X, Wt, S):
    """Direct 3x3 convolution, padding 1: Y[b,y,x,o] = sum over taps and input
    channels of W[o, ky, kx, i] * X[b, y*S+ky-1, x*S+kx-1, i]."""
    B, H, W, Ci = X.shape
    Co = Wt.shape[0]
    Wr = Wt.reshape(Co, 3, 3, Ci)
    Y = np.zeros((B, H // S, W // S, Co))
    for b in range(B):
        for y in range(H // S):
            for x in range(W // S):
                for o in range(Co):
                    s = 0.0
                    for ky in range(3):
                        for kx in range(3):
                            ys, xs = y * S + ky - 1, x * S + kx - 1
                            if 0 <= ys < H and 0 <= xs < W:
                                for i in range(Ci):
                                    s += Wr[o, ky, kx, i] * X[b, ys, xs, i]
                    Y[b, y, x, o] = s
    return Y


def up2(Y):
    return np.repeat(np.repeat(Y, 2, axis=1), 2, axis=2)


def up2_back(dX):
    B, H, W, C = dX.shape
    return dX.reshape(B, H // 2, 2, W // 2, 2, C).sum(axis=(2, 4))


class RefNet:
    """The network of floodca_amd.nn in float64 over the same flat parameters."""

    def __init__(self, cfg: NetCfg):
        self.cfg = cfg
        self.L = layers(cfg)
        self.by = {ly.name: ly for ly in self.L}

    def _w(self, P, ly, r):
        return r(P[ly.off_w:ly.off_b].reshape(ly.cout, ly.k))

    def _bias(self, P, ly, s):
        b = P[ly.off_b:ly.off_f]
        Wf = P[ly.off_f:ly.off_f + 4 * ly.cout].reshape(ly.cout, 4)
        return b[None, :] + s @ Wf.T          # [B, cout]

    def forward(self, P, X0, s, emulate=True):
        """X0: [B, R/4, R/4, 64] (float values), s: [B, 4]. Returns the cache
        (activations and the head output O [B, R/4, R/4, 32])."""
        r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
        P = np.asarray(P, dtype=np.float64)
        X0 = np.asarray(X0, dtype=np.float64)
        s = np.asarray(s, dtype=np.float64)
        B = X0.shape[0]
        by = self.by
        c = {"X0": X0, "s": s}

        def act(ly, v):
            v = v + self._bias(P, ly, s)[:, None, None, :]
            if ly.relu:
                v = np.maximum(v, 0.0)
            return v

        e1 = by["E1"]
        c["E1"] = r(act(e1, X0 @ self._w(P, e1, r).T))
        prev = c["E1"]
        names = ["E2", "E3", "E4"] + [f"B{i + 1}" for i in range(self.cfg.nbott)]
        for name in names:
            ly = by[name]
            cols = im2col(prev, ly.stride).reshape(B, ly.hout, ly.hout, -1)
            c[name] = r(act(ly, cols @ self._w(P, ly, r).T))
            prev = c[name]
        for name, skip, u in (("D4", "E3", "U4"), ("D3", "E2", "U3"), ("D2", "E1", "U2")):
            ly = by[name]
            cols = im2col(prev, 1).reshape(B, ly.hout, ly.hout, -1)
            c[name] = r(act(ly, cols @ self._w(P, ly, r).T))
            c[u] = r(up2(c[name]) + c[skip])
            prev = c[u]
        hd = by["H"]
        c["O"] = act(hd, prev @ self._w(P, hd, r).T)
        return c

    def backward(self, P, c, dO, emulate=True):
        """Flat gradient of every parameter given dO = d(objective)/dO."""
        r = bf16 if emulate else (lambda v: np.asarray(v, dtype=np.float64))
        P = np.asarray(P, dtype=np.float64)
        G = np.zeros_like(P)
        s = c["s"]
        by = self.by

        def bias_grad(ly, gY):
            ps = gY.sum(axis=(1, 2))              # [B, cout]
            G[ly.off_b:ly.off_f] = ps.sum(axis=0)
            G[ly.off_f:ly.off_f + 4 * ly.cout] = (ps.T @ s).ravel()

        def conv_back(ly, X, gY, need_dx):
            gY = np.where(c[ly.name] > 0, gY, 0.0) if ly.relu else gY
            bias_grad(ly, gY)
            gb = r(gY)
            B = X.shape[0]
            cols = im2col(X, ly.stride).reshape(-1, 9 * ly.cin)
            G[ly.off_w:ly.off_b] = (gb.reshape(-1, ly.cout).T @ cols).ravel()
            if not need_dx:
                return None
            dcol = (gb.reshape(-1, ly.cout) @ self._w(P, ly, r)).reshape(
                B, ly.hout, ly.hout, 9, ly.cin)
            return col2im(dcol, ly.hin, ly.hin, ly.stride)

        hd = by["H"]
        dO = np.asarray(dO, dtype=np.float64)
        bias_grad(hd, dO)
        gb = r(dO).reshape(-1, 32)
        G[hd.off_w:hd.off_b] = (gb.T @ c["U2"].reshape(-1, hd.cin)).ravel()
        gU = (gb @ self._w(P, hd, r)).reshape(c["U2"].shape)
        g = {}
        for name, u, xin in (("D2", "U2", "U3"), ("D3", "U3", "U4"), ("D4", "U4", None)):
            ly = by[name]
            skip = {"U2": "E1", "U3": "E2", "U4": "E3"}[u]
            g[skip] = gU
            last = f"B{self.cfg.nbott}" if self.cfg.nbott else "E4"
            X = c[xin] if xin else c[last]
            gU = conv_back(ly, X, up2_back(gU), True)
        gprev = gU                                # gradient of the bottleneck output
        chain = ["E4"] + [f"B{i + 1}" for i in range(self.cfg.nbott)]
        for i in range(len(chain) - 1, 0, -1):
            gprev = conv_back(by[chain[i]], c[chain[i - 1]], gprev, True)
        for name, src in (("E4", "E3"), ("E3", "E2"), ("E2", "E1")):
            gprev = conv_back(by[name], c[src], gprev, True) + g[src]
        e1 = by["E1"]
        gY = np.where(c["E1"] > 0, gprev, 0.0)
        bias_grad(e1, gY)
        G[e1.off_w:e1.off_b] = (r(gY).reshape(-1, e1.cout).T @ c["X0"].reshape(-1, 64)).ravel()
        return G


def head_maps(O, R):
    """Mean and log-variance maps [B, R, R] from head output [B, R/4, R/4, 32]."""
    B, q = O.shape[0], R // 4
    o = O.reshape(B, q, q, 4, 4, 2)              # (Y, X, i, j, {mean, s})
    mu = o[..., 0].transpose(0, 1, 3, 2, 4).reshape(B, R, R)
    ls = o[..., 1].transpose(0, 1, 3, 2, 4).reshape(B, R, R)
    return mu, ls


def to_head(dmu, dls):
    """Inverse of head_maps for gradients."""
    B, R, _ = dmu.shape
    q = R // 4
    o = np.empty((B, q, q, 4, 4, 2))
    o[..., 0] = dmu.reshape(B, q, 4, q, 4).transpose(0, 1, 3, 2, 4)
    o[..., 1] = dls.reshape(B, q, 4, q, 4).transpose(0, 1, 3, 2, 4)
    return o.reshape(B, q, q, 32)


def loss_terms(mu, ls, T, W, alpha):
    """Per-pixel squared-error and likelihood terms of the loss."""
    sc = np.clip(ls, LS_LO, LS_HI)
    r2 = (mu - T) ** 2
    return W * alpha * r2, W * 0.5 * (np.exp(-sc) * r2 + sc)


def loss(O, T, W, alpha, inv_wsum):
    """(sum of l, dO) as fca_nn_loss defines them."""
    R = T.shape[-1]
    mu, ls = head_maps(np.asarray(O, dtype=np.float64), R)
    T = np.asarray(T, dtype=np.float64)
    W = np.asarray(W, dtype=np.float64)
    a, b = loss_terms(mu, ls, T, W, alpha)
    sc = np.clip(ls, LS_LO, LS_HI)
    r = mu - T
    dmu = W * 2.0 * alpha * r * inv_wsum
    dls = np.where((ls >= LS_LO) & (ls <= LS_HI), W * 0.5 * (1 - np.exp(-sc) * r * r), 0.0)
    return float((a + b).sum()), to_head(dmu, dls * inv_wsum)


# ---- terrain pyramid, inputs, targets ----------------------------------------------

def pyramid(z, mask, n, ncls, levels=6):
    """Records per level as a dict of arrays: zmean, zmin, zmax, land, cls."""
    land = (mask == 0).astype(np.int64)
    cls = np.abs(n[..., None].astype(np.float32) - np.asarray(ncls, np.float32)).argmin(-1)
    lev = [dict(zmean=z.astype(np.float64), zmin=z.astype(np.float64),
                zmax=z.astype(np.float64), land=land, cls=np.where(land > 0, cls, 0))]
    for _ in range(1, levels):
        p = lev[-1]
        nr, nc = p["land"].shape

        def blk(a):
            return a.reshape(nr // 2, 2, nc // 2, 2).transpose(0, 2, 1, 3).reshape(
                nr // 2, nc // 2, 4)
        ln = blk(p["land"])
        first = ln.argmax(-1)                     # first child with the most land
        cl = np.take_along_axis(blk(p["cls"]), first[..., None], -1)[..., 0]
        tot = ln.sum(-1)
        lev.append(dict(zmean=blk(p["zmean"]).mean(-1), zmin=blk(p["zmin"]).min(-1),
                        zmax=blk(p["zmax"]).max(-1), land=tot, cls=np.where(tot > 0, cl, 0)))
    return lev


def _span(lo, hi, lim):
    first = math.floor(lo)
    last = min(math.ceil(hi) - 1, lim - 1)
    return [(first + j, min(hi, first + j + 1) - max(lo, first + j))
            for j in range(last - first + 1)]


def zref(lev, lv, cy, cx):
    """Bilinear mean z of level lv at (cy, cx) (cell centres at +0.5),
    clamped at the edges."""
    zm = lev[lv]["zmean"]
    n_r, n_c = zm.shape
    u, v = cy - 0.5, cx - 0.5
    i0, j0 = math.floor(u), math.floor(v)
    fu, fv = u - i0, v - j0
    if i0 < 0:
        i0, fu = 0, 0.0
    if i0 > n_r - 2:
        i0, fu = n_r - 2, 1.0
    if j0 < 0:
        j0, fv = 0, 0.0
    if j0 > n_c - 2:
        j0, fv = n_c - 2, 1.0
    top = zm[i0, j0] * (1 - fv) + zm[i0, j0 + 1] * fv
    bot = zm[i0 + 1, j0] * (1 - fv) + zm[i0 + 1, j0 + 1] * fv
    return top * (1 - fu) + bot * fu


def inputs(lev, samples, ncls, R):
    """(X [B, R/4, R/4, 64] float values before bf16, Lf [B, R, R]) of
    fca_nn_inputs; samples = [(r0, c0, side, level, nscale[8])]."""
    B = len(samples)
    X = np.zeros((B, R // 4, R // 4, 64))
    Lf = np.zeros((B, R, R))
    nlev = len(lev)
    for b, (r0, c0, side, level, nscale) in enumerate(samples):
        rec = lev[level]
        lref = min(level + 2, nlev - 1)
        lnr, lnc = rec["land"].shape
        cell = float(4 ** level)
        f = side / R
        inv = 1.0 / (1 << level)
        for y in range(R):
            sy = _span((r0 + y * f) * inv, (r0 + (y + 1) * f) * inv, lnr)
            for x in range(R):
                sx = _span((c0 + x * f) * inv, (c0 + (x + 1) * f) * inv, lnc)
                Wt = zs = lf = nn = 0.0
                for iy, wy in sy:
                    for ix, wx in sx:
                        w = wy * wx
                        land = rec["land"][iy, ix] / cell
                        k = int(rec["cls"][iy, ix])
                        Wt += w
                        zs += w * rec["zmean"][iy, ix]
                        lf += w * land
                        nn += w * land * ncls[k] * nscale[k]
                zp = zs / Wt
                zr = zref(lev, lref, (r0 + (y + 0.5) * f) / (1 << lref),
                          (c0 + (x + 0.5) * f) / (1 << lref))
                ch = [(zp - 30.0) / 20.0, (zp - zr) / 0.5, lf / Wt,
                      (nn / lf - 0.06) / 0.04 if lf > 0 else 0.0]
                X[b, y // 4, x // 4, ((y % 4) * 4 + x % 4) * 4:((y % 4) * 4 + x % 4) * 4 + 4] = ch
                Lf[b, y, x] = lf / Wt
    return X, Lf


def target(hmax, mask, side, R, scale=1.0, wscale=1.0):
    """(T, W) [R, R] of fca_nn_target for one side x side crop."""
    f = side / R
    T = np.zeros((R, R))
    Wl = np.zeros((R, R))
    land = mask == 0
    for y in range(R):
        sy = _span(y * f, (y + 1) * f, side)
        for x in range(R):
            sx = _span(x * f, (x + 1) * f, side)
            W = L = H = 0.0
            for iy, wy in sy:
                for ix, wx in sx:
                    w = wy * wx
                    W += w
                    if land[iy, ix]:
                        L += w
                        H += w * float(hmax[iy, ix])
            T[y, x] = H / L * scale if L > 0 else 0.0
            Wl[y, x] = L / W * wscale if W > 0 else 0.0
    return T, Wl
