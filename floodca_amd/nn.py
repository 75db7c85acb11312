"""The flood surrogate network on the gfx950 kernels of libfca_nn.so.

ctypes binding of include/fca_nn.h and the network: a convolutional
encoder-decoder over R x R maps (R = 256 by default) whose every forward and
backward operation is one of the library's kernels -- matrix-core BF16 GEMMs
that read 3x3 convolutions in place, and the bias, activation, loss and Adam
kernels around them. PyTorch provides device memory and the stream only; a
missing library raises.

Architecture (widths c1..c4, default 32, 64, 128, 256; about 1.96 M
parameters at the default):
  E1  4x4 stride-4 patch conv, 4 -> c1, at R/4 (the inputs arrive as the
      patch matrix, see fca_nn_inputs)
  E2..E4  3x3 stride-2 convs to c2 at R/8, c3 at R/16, c4 at R/32
  B1..Bn  3x3 convs c4 -> c4 at R/32
  D4  3x3 conv c4 -> c3 at R/32, nearest x2 upsampling, + E3 (skip)
  D3  3x3 conv c3 -> c2 at R/16, x2, + E2
  D2  3x3 conv c2 -> c1 at R/8, x2, + E1
  H   1x1 conv c1 -> 32 at R/4: per output pixel of its 4x4 patch a mean and
      a log-variance (fca_nn_loss's layout)
Every conv but H is followed by ReLU, and every layer's bias is per sample:
b + Wf s for the sample's 4 forcing scalars s (feature-wise modulation), which
is how rain and duration enter the network.
"""
from __future__ import annotations

import ctypes as ct
import math
import os
from dataclasses import dataclass

import numpy as np

from .build import load_hip_library, nn_lib_path


class NnError(RuntimeError):
    pass


class FcaNnGemm(ct.Structure):
    _fields_ = [("M", ct.c_int), ("N", ct.c_int), ("K", ct.c_int),
                ("A", ct.c_void_p), ("lda", ct.c_int), ("a_mode", ct.c_int),
                ("a_lh", ct.c_int), ("a_lw", ct.c_int), ("a_lc", ct.c_int), ("a_sh", ct.c_int),
                ("a_sw", ct.c_int), ("a_s", ct.c_int),
                ("B", ct.c_void_p), ("ldb", ct.c_int), ("b_mode", ct.c_int),
                ("b_lh", ct.c_int), ("b_lw", ct.c_int), ("b_lc", ct.c_int), ("b_sh", ct.c_int),
                ("b_sw", ct.c_int), ("b_s", ct.c_int),
                ("C", ct.c_void_p), ("Cb", ct.c_void_p), ("ldc", ct.c_int),
                ("accumulate", ct.c_int), ("bias", ct.c_void_p), ("bias_rows", ct.c_int),
                ("relu", ct.c_int), ("kchunk", ct.c_int)]


NCLASS = 8
LEVELS = 6


class FcaNnPyrRec(ct.Structure):
    _fields_ = [("zmean", ct.c_float), ("zmin", ct.c_float), ("zmax", ct.c_float),
                ("land", ct.c_uint16), ("cls", ct.c_uint8), ("pad", ct.c_uint8)]


class FcaNnPyr(ct.Structure):
    _fields_ = [("lev", ct.c_void_p * LEVELS), ("nr", ct.c_int), ("nc", ct.c_int)]


class FcaNnSample(ct.Structure):
    _fields_ = [("r0", ct.c_int), ("c0", ct.c_int), ("side", ct.c_int), ("level", ct.c_int),
                ("nscale", ct.c_float * NCLASS)]


_V, _I, _L, _F = ct.c_void_p, ct.c_int, ct.c_longlong, ct.c_float
SIGS = {
    "fca_nn_gemm": [ct.POINTER(FcaNnGemm), _V],
    "fca_nn_splitk_sum": [_V, _I, _L, _V, _I, _V],
    "fca_nn_relu_back": [_V, _V, _L, _I, _V, _V],
    "fca_nn_to_bf16": [_V, _V, _L, _V],
    "fca_nn_add": [_V, _V, _L, _V],
    "fca_nn_colsum_sample": [_V, _I, _I, _I, _V, _V],
    "fca_nn_film": [_V, _V, _V, _I, _I, _V, _V],
    "fca_nn_film_back": [_V, _V, _I, _I, _V, _V, _V],
    "fca_nn_wt": [_V, _I, _I, _V, _V],
    "fca_nn_up2_add": [_V, _V, _I, _I, _I, _I, _V, _V],
    "fca_nn_up2_back": [_V, _I, _I, _I, _I, _V, _V],
    "fca_nn_loss": [_V, _V, _V, _I, _I, _F, _F, _V, _V, _V],
    "fca_nn_adam": [_V, _V, _V, _V, _L, _F, _F, _F, _F, _I, _V],
    "fca_nn_pred_stats": [_V, _V, _I, _I, _V, _F, _V, _V],
    "fca_nn_sq_err": [_V, _V, _V, _I, _I, _F, _V, _V],
    "fca_nn_pyr0": [_V, _V, _V, _L, _V, _I, _V, _V],
    "fca_nn_pyr_down": [_V, _I, _I, _V, _V],
    "fca_nn_inputs": [ct.POINTER(FcaNnPyr), _V, _I, _V, _I, _V, _V, _V],
    "fca_nn_target": [_V, _V, _I, _I, _F, _F, _V, _V, _V],
    "fca_nn_gather": [_V, _V, _I, _L, _V, _V],
    "fca_nn_row_moments": [_V, _V, _I, _L, _V, _V],
}


SAMPLE_DTYPE = np.dtype([("r0", "<i4"), ("c0", "<i4"), ("side", "<i4"), ("level", "<i4"),
                         ("nscale", "<f4", (NCLASS,))])
assert SAMPLE_DTYPE.itemsize == ct.sizeof(FcaNnSample)


class Lib:
    """libfca_nn.so: every call checks its return code."""

    def __init__(self, path: str | None = None):
        path = path or nn_lib_path()
        if not os.path.exists(path):
            raise NnError(f"{path} is not built: run python -m floodca_amd.build")
        self.lib = load_hip_library(path)
        for name, args in SIGS.items():
            f = getattr(self.lib, name)
            f.restype, f.argtypes = ct.c_int, args

    def __getattr__(self, name):
        f = getattr(self.__dict__["lib"], "fca_nn_" + name)

        def call(*args):
            rc = f(*args)
            if rc != 0:
                raise NnError(f"fca_nn_{name} returned {rc}")
            return rc
        return call


_LIB: Lib | None = None


def lib() -> Lib:
    global _LIB
    if _LIB is None:
        _LIB = Lib()
    return _LIB


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise NnError("the surrogate kernels need a GPU")
    return torch


def stream() -> int:
    return _torch().cuda.current_stream().cuda_stream


def ptr(t) -> int:
    return 0 if t is None else t.data_ptr()


def log2i(v: int) -> int:
    lv = int(v).bit_length() - 1
    if v <= 0 or (1 << lv) != v:
        raise NnError(f"{v} is not a power of two")
    return lv


def gemm(M, N, K, A, B, *, a_mode=0, lda=0, b_mode=0, ldb=0, ageo=None, bgeo=None, C=None,
         Cb=None, ldc=None, accumulate=0, bias=None, bias_rows=0, relu=0, kchunk=0):
    """fca_nn_gemm; ageo / bgeo = (H, W, C, sh, sw, S) of an in-place conv mode
    (H, W: the pixel space; C: the source channels; sh, sw: the source)."""
    g = FcaNnGemm(M=M, N=N, K=K, A=ptr(A), lda=lda, a_mode=a_mode, B=ptr(B), ldb=ldb,
                  b_mode=b_mode, C=ptr(C), Cb=ptr(Cb), ldc=N if ldc is None else ldc,
                  accumulate=accumulate, bias=ptr(bias), bias_rows=bias_rows, relu=relu,
                  kchunk=kchunk)
    if ageo:
        h, w, c, sh, sw, s = ageo
        g.a_lh, g.a_lw, g.a_lc, g.a_sh, g.a_sw, g.a_s = log2i(h), log2i(w), log2i(c), sh, sw, s
    if bgeo:
        h, w, c, sh, sw, s = bgeo
        g.b_lh, g.b_lw, g.b_lc, g.b_sh, g.b_sw, g.b_s = log2i(h), log2i(w), log2i(c), sh, sw, s
    lib().gemm(ct.byref(g), stream())


def split_plan(M, N, K, target=512, min_chunk=256):
    """Split-K chunk for a GEMM with few output tiles and a long K: enough
    chunks for ~target workgroups, each at least min_chunk long."""
    tiles = ((M + 63) // 64) * ((N + 63) // 64)
    nz = max(1, min(-(-target // tiles), K // min_chunk))
    chunk = -(-(-(-K // nz)) // 32) * 32
    return chunk, -(-K // chunk)


@dataclass(frozen=True)
class NetCfg:
    R: int = 256
    widths: tuple = (32, 64, 128, 256)
    nbott: int = 2


@dataclass
class Layer:
    name: str
    kind: str      # "patch", "conv" or "head"
    cin: int
    cout: int
    stride: int
    hin: int       # input side
    relu: int
    off_w: int = 0
    off_b: int = 0
    off_f: int = 0

    @property
    def k(self) -> int:
        return {"patch": 64, "conv": 9 * self.cin, "head": self.cin}[self.kind]

    @property
    def hout(self) -> int:
        return self.hin // self.stride

    @property
    def nparam(self) -> int:
        return self.cout * self.k + self.cout * 5


def layers(cfg: NetCfg) -> list[Layer]:
    c1, c2, c3, c4 = cfg.widths
    q = cfg.R // 4
    L = [Layer("E1", "patch", 4, c1, 1, q, 1),
         Layer("E2", "conv", c1, c2, 2, q, 1),
         Layer("E3", "conv", c2, c3, 2, q // 2, 1),
         Layer("E4", "conv", c3, c4, 2, q // 4, 1)]
    L += [Layer(f"B{i + 1}", "conv", c4, c4, 1, q // 8, 1) for i in range(cfg.nbott)]
    L += [Layer("D4", "conv", c4, c3, 1, q // 8, 1),
          Layer("D3", "conv", c3, c2, 1, q // 4, 1),
          Layer("D2", "conv", c2, c1, 1, q // 2, 1),
          Layer("H", "head", c1, 32, 1, q, 0)]
    off = 0
    for ly in L:
        ly.off_w = off
        ly.off_b = off + ly.cout * ly.k
        ly.off_f = ly.off_b + ly.cout
        off += ly.nparam
    return L


def init_params(cfg: NetCfg, seed: int) -> np.ndarray:
    """He-normal conv weights, zero biases and modulation; the head starts
    small."""
    rng = np.random.default_rng(seed)
    L = layers(cfg)
    p = np.zeros(sum(ly.nparam for ly in L), dtype=np.float32)
    for ly in L:
        std = math.sqrt(2.0 / ly.k) if ly.relu else 0.1 * math.sqrt(1.0 / ly.k)
        p[ly.off_w:ly.off_b] = rng.standard_normal(ly.cout * ly.k) * std
    return p


class Work:
    """Device buffers of one batch size."""

    def __init__(self, net: "Net", B: int, train: bool):
        torch = _torch()
        dev = net.dev
        bf, f32 = torch.uint16, torch.float32
        self.B = B
        self.train = train
        q = net.cfg.R // 4
        c1, c2, c3, c4 = net.cfg.widths
        self.s = torch.empty((B, 4), dtype=f32, device=dev)
        self.bs = {ly.name: torch.empty((B, ly.cout), dtype=f32, device=dev) for ly in net.L}
        self.y = {}    # bf16 outputs of each layer (after ReLU)
        for ly in net.L[:-1]:
            self.y[ly.name] = torch.empty((B, ly.hout, ly.hout, ly.cout), dtype=bf, device=dev)
        self.u = {"U4": torch.empty((B, q // 4, q // 4, c3), dtype=bf, device=dev),
                  "U3": torch.empty((B, q // 2, q // 2, c2), dtype=bf, device=dev),
                  "U2": torch.empty((B, q, q, c1), dtype=bf, device=dev)}
        self.O = torch.empty((B, q, q, 32), dtype=f32, device=dev)
        if not train:
            return
        self.dO = torch.empty_like(self.O)
        self.g = {}    # FP32 gradients of layer outputs and of the U maps
        for ly in net.L[:-1]:
            if ly.name in ("E1", "E2", "E3"):
                continue  # the gradients of E1..E3 are those of U2..U4
            self.g[ly.name] = torch.empty((B, ly.hout, ly.hout, ly.cout), dtype=f32, device=dev)
        for k, v in self.u.items():
            self.g[k] = torch.empty(v.shape, dtype=f32, device=dev)
        self.g["E3"], self.g["E2"], self.g["E1"] = self.g["U4"], self.g["U3"], self.g["U2"]
        n = max(t.numel() for t in self.g.values())
        self.gb = torch.empty(max(n, self.O.numel()), dtype=bf, device=dev)
        self.ps = torch.empty(B * max(net.cfg.widths + (32,)), dtype=f32, device=dev)
        parts = 0
        for ly in net.L:
            kk = B * ly.hout * ly.hout
            chunk, nz = split_plan(ly.cout, ly.k, kk)
            parts = max(parts, nz * ly.cout * ly.k)
        self.part = torch.empty(parts, dtype=f32, device=dev)
        self.lpart = torch.empty(-(-B * net.cfg.R * net.cfg.R // 256), dtype=f32, device=dev)


class Net:
    """Parameters (FP32 master, BF16 copy), Adam state and the passes."""

    def __init__(self, cfg: NetCfg = NetCfg(), seed: int = 0, device: str = "cuda"):
        torch = _torch()
        lib()
        self.cfg = cfg
        self.dev = device
        self.L = layers(cfg)
        self.by = {ly.name: ly for ly in self.L}
        p = init_params(cfg, seed)
        self.n = p.size
        self.P = torch.from_numpy(p).to(device)
        self.Pb = torch.empty(self.n, dtype=torch.uint16, device=device)
        self.G = torch.zeros(self.n, dtype=torch.float32, device=device)
        self.m = torch.zeros_like(self.G)
        self.v = torch.zeros_like(self.G)
        self.t = 0
        self.Wt = {ly.name: torch.empty(9 * ly.cin * ly.cout, dtype=torch.uint16, device=device)
                   for ly in self.L if ly.kind == "conv"}
        self._work = {}

    # parameter views
    def _f(self, t, off):
        return t.data_ptr() + 4 * off

    def _h(self, off):
        return self.Pb.data_ptr() + 2 * off

    def work(self, B: int, train: bool) -> Work:
        key = (B, train)
        if key not in self._work:
            self._work[key] = Work(self, B, train)
        return self._work[key]

    def params(self) -> np.ndarray:
        return self.P.cpu().numpy()

    def load(self, p: np.ndarray):
        self.P.copy_(_torch().from_numpy(np.ascontiguousarray(p, dtype=np.float32)))

    # ---- forward -----------------------------------------------------------
    def _conv(self, ly, X, B, Y, bias):
        h = ly.hin
        ho = ly.hout
        gemm(B * ho * ho, ly.cout, ly.k, X, _Addr(self._h(ly.off_w)), a_mode=2,
             ageo=(ho, ho, ly.cin, h, h, ly.stride), b_mode=0, ldb=ly.k,
             Cb=Y, bias=bias, bias_rows=ho * ho, relu=ly.relu)

    def forward(self, X0, s, B: int, train: bool = False) -> Work:
        """X0: [B][R/4][R/4][64] BF16 inputs, s: [B][4] FP32 scalars (device)."""
        L = lib()
        st = stream()
        w = self.work(B, train)
        if s.data_ptr() != w.s.data_ptr():
            w.s.copy_(s)
        L.to_bf16(ptr(self.P), ptr(self.Pb), self.n, st)
        for ly in self.L:
            L.film(self._f(self.P, ly.off_b), self._f(self.P, ly.off_f), ptr(w.s), B, ly.cout,
                   ptr(w.bs[ly.name]), st)
        q = self.cfg.R // 4
        by = self.by
        e1 = by["E1"]
        gemm(B * q * q, e1.cout, 64, X0, _Addr(self._h(e1.off_w)), a_mode=0, lda=64, b_mode=0,
             ldb=64, Cb=w.y["E1"], bias=w.bs["E1"], bias_rows=q * q, relu=1)
        prev = w.y["E1"]
        for name in ["E2", "E3", "E4"] + [f"B{i + 1}" for i in range(self.cfg.nbott)]:
            self._conv(by[name], prev, B, w.y[name], w.bs[name])
            prev = w.y[name]
        for name, skip, u in (("D4", "E3", "U4"), ("D3", "E2", "U3"), ("D2", "E1", "U2")):
            ly = by[name]
            self._conv(ly, prev, B, w.y[name], w.bs[name])
            L.up2_add(ptr(w.y[name]), ptr(w.y[skip]), B, ly.hout, ly.hout, ly.cout, ptr(w.u[u]),
                      st)
            prev = w.u[u]
        hd = by["H"]
        gemm(B * q * q, 32, hd.cin, prev, _Addr(self._h(hd.off_w)), a_mode=0, lda=hd.cin,
             b_mode=0, ldb=hd.cin, C=w.O, bias=w.bs["H"], bias_rows=q * q, relu=0)
        return w

    # ---- loss and backward ----------------------------------------------------
    def loss(self, w: Work, T, Wp, inv_wsum: float, alpha: float = 1.0):
        """Writes dO; returns the device tensor of loss partial sums."""
        lib().loss(ptr(w.O), ptr(T), ptr(Wp), w.B, self.cfg.R, alpha, inv_wsum, ptr(w.dO),
                   ptr(w.lpart), stream())
        return w.lpart

    def _wgrad(self, ly, w, dYb, X, B, xgeo):
        """dW of layer ly from the BF16 output gradient dYb and its input X."""
        L = lib()
        kk = B * ly.hout * ly.hout
        chunk, nz = split_plan(ly.cout, ly.k, kk)
        if xgeo is None:   # plain: X stored [pixels][k]
            gemm(ly.cout, ly.k, kk, dYb, X, a_mode=1, lda=ly.cout, b_mode=1, ldb=ly.k,
                 C=w.part, kchunk=chunk)
        else:
            gemm(ly.cout, ly.k, kk, dYb, X, a_mode=1, lda=ly.cout, b_mode=4, bgeo=xgeo,
                 C=w.part, kchunk=chunk)
        L.splitk_sum(ptr(w.part), nz, ly.cout * ly.k, self._f(self.G, ly.off_w), 0, stream())

    def _bias_grad(self, ly, w, gY, B):
        L = lib()
        st = stream()
        L.colsum_sample(ptr(gY), B, ly.hout * ly.hout, ly.cout, ptr(w.ps), st)
        L.film_back(ptr(w.ps), ptr(w.s), B, ly.cout, self._f(self.G, ly.off_b),
                    self._f(self.G, ly.off_f), st)

    def _conv_back(self, ly, w, X, gX, acc, B):
        L = lib()
        st = stream()
        gY = w.g[ly.name]
        n = gY.numel()
        L.relu_back(ptr(gY), ptr(w.y[ly.name]), n, ly.relu, ptr(w.gb), st)
        self._bias_grad(ly, w, gY, B)
        h, ho = ly.hin, ly.hout
        self._wgrad(ly, w, w.gb, X, B, (ho, ho, ly.cin, h, h, ly.stride))
        if gX is not None:
            L.wt(self._f(self.P, ly.off_w), ly.cout, ly.cin, ptr(self.Wt[ly.name]), st)
            gemm(B * h * h, ly.cin, 9 * ly.cout, w.gb, self.Wt[ly.name], a_mode=3,
                 ageo=(h, h, ly.cout, ho, ho, ly.stride), b_mode=0, ldb=9 * ly.cout, C=gX,
                 accumulate=acc)

    def backward(self, w: Work, X0):
        """Gradients of every parameter into G from w.dO (after loss())."""
        L = lib()
        st = stream()
        B = w.B
        q = self.cfg.R // 4
        by = self.by
        hd = by["H"]
        L.relu_back(ptr(w.dO), 0, w.dO.numel(), 0, ptr(w.gb), st)
        self._bias_grad(hd, w, w.dO, B)
        self._wgrad(hd, w, w.gb, w.u["U2"], B, None)
        gemm(B * q * q, hd.cin, 32, w.gb, _Addr(self._h(hd.off_w)), a_mode=0, lda=32, b_mode=1,
             ldb=hd.cin, C=w.g["U2"])
        for name, u, xin, gin in (("D2", "U2", "U3", "U3"), ("D3", "U3", "U4", "U4"),
                                  ("D4", "U4", None, None)):
            ly = by[name]
            L.up2_back(ptr(w.g[u]), B, ly.hout, ly.hout, ly.cout, ptr(w.g[name]), st)
            if xin is None:
                last = f"B{self.cfg.nbott}" if self.cfg.nbott else "E4"
                self._conv_back(ly, w, w.y[last], w.g[last], 0, B)
            else:
                self._conv_back(ly, w, w.u[xin], w.g[gin], 0, B)
        chain = ["E4"] + [f"B{i + 1}" for i in range(self.cfg.nbott)]
        for i in range(len(chain) - 1, 0, -1):
            self._conv_back(by[chain[i]], w, w.y[chain[i - 1]], w.g[chain[i - 1]], 0, B)
        for name, src in (("E4", "E3"), ("E3", "E2"), ("E2", "E1")):
            self._conv_back(by[name], w, w.y[src], w.g[src], 1, B)
        e1 = by["E1"]
        L.relu_back(ptr(w.g["E1"]), ptr(w.y["E1"]), w.g["E1"].numel(), 1, ptr(w.gb), st)
        self._bias_grad(e1, w, w.g["E1"], B)
        self._wgrad(e1, w, w.gb, X0, B, None)

    def adam(self, lr: float, b1=0.9, b2=0.999, eps=1e-8):
        self.t += 1
        lib().adam(ptr(self.P), ptr(self.G), ptr(self.m), ptr(self.v), self.n, lr, b1, b2, eps,
                   self.t, stream())


class _Addr:
    """A raw device address standing in for a tensor in gemm()."""

    def __init__(self, a: int):
        self.a = a

    def data_ptr(self) -> int:
        return self.a


def forward_flops(cfg: NetCfg) -> int:
    """Multiply-adds x 2 of one sample's forward pass (every GEMM's M*N*K)."""
    return 2 * sum(ly.hout * ly.hout * ly.cout * ly.k for ly in layers(cfg))


def main(argv=None) -> int:
    """Times the forward pass alone: ``python -m floodca_amd.nn --batch 256
    --reps 20`` (random weights, zero inputs; the kernels do the same work
    whatever the values), for profiling the matrix-core GEMMs."""
    import argparse

    ap = argparse.ArgumentParser(description=main.__doc__.split("\n")[0])
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--reps", type=int, default=20)
    args = ap.parse_args(argv)
    torch = _torch()
    net = Net(NetCfg(), seed=0)
    q = net.cfg.R // 4
    X = torch.zeros((args.batch, q, q, 64), dtype=torch.uint16, device="cuda")
    S = torch.zeros((args.batch, 4), dtype=torch.float32, device="cuda")
    for _ in range(3):
        net.forward(X, S, args.batch)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(args.reps):
        net.forward(X, S, args.batch)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.reps
    fl = forward_flops(net.cfg) * args.batch
    print(f"forward: {ms:.3f} ms per batch of {args.batch}, {fl / (ms * 1e-3) / 1e12:.1f} TFLOP/s "
          f"({net.n} parameters, {forward_flops(net.cfg) / 1e9:.3f} GFLOP per sample)", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
