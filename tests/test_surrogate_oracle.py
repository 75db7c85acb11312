"""The surrogate's CPU restatement pinned without a GPU: convolution against
a direct loop, col2im as the adjoint of im2col, the backward pass and the loss
gradient against central finite differences, the inputs and targets against
closed forms; and the C-ABI of libfca_nn.so (struct layouts, exported
symbols)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from floodca_amd import nn
from floodca_amd.build import nn_lib_path
from oracle import surrogate_ref as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TINY = nn.NetCfg(R=32, widths=(8, 8, 16, 16), nbott=1)


def test_conv_im2col_equals_direct_loops():
    rng = np.random.default_rng(1)
    for S_, H in ((1, 5), (2, 6), (2, 4), (1, 1)):
        X = rng.standard_normal((2, H, H, 3))
        W = rng.standard_normal((4, 27))
        cols = S.im2col(X, S_).reshape(2, H // S_, H // S_, 27)
        assert np.allclose(cols @ W.T, S.conv_direct(X, W, S_), rtol=1e-12, atol=1e-12)


def test_col2im_is_the_adjoint_of_im2col():
    rng = np.random.default_rng(2)
    for S_, H in ((1, 6), (2, 8)):
        X = rng.standard_normal((2, H, H, 4))
        D = rng.standard_normal((2, H // S_, H // S_, 9, 4))
        lhs = float((S.im2col(X, S_) * D).sum())
        rhs = float((X * S.col2im(D, H, H, S_)).sum())
        assert lhs == pytest.approx(rhs, rel=1e-12)


def tiny_case(seed, B=2, cfg=TINY):
    rng = np.random.default_rng(seed)
    P = nn.init_params(cfg, seed).astype(np.float64)
    P += rng.standard_normal(P.size) * 0.05      # biases and modulation nonzero
    q = cfg.R // 4
    X0 = rng.standard_normal((B, q, q, 64))
    s = rng.standard_normal((B, 4))
    return P, X0, s


def test_backward_matches_finite_differences():
    """J = sum(O * Gr) for a fixed random Gr: dJ/dP by the restated backward
    pass against central differences, on parameters of every layer (weights,
    biases, modulation)."""
    net = S.RefNet(TINY)
    P, X0, s = tiny_case(3)
    c = net.forward(P, X0, s, emulate=False)
    Gr = np.random.default_rng(4).standard_normal(c["O"].shape)
    G = net.backward(P, c, Gr, emulate=False)
    rng = np.random.default_rng(5)
    picks = []
    for ly in net.L:
        for lo, hi in ((ly.off_w, ly.off_b), (ly.off_b, ly.off_f),
                       (ly.off_f, ly.off_f + 4 * ly.cout)):
            picks += rng.integers(lo, hi, 3).tolist()
    eps = 1e-6
    worst = 0.0
    for i in picks:
        Pp, Pm = P.copy(), P.copy()
        Pp[i] += eps
        Pm[i] -= eps
        fd = ((net.forward(Pp, X0, s, False)["O"] * Gr).sum() -
              (net.forward(Pm, X0, s, False)["O"] * Gr).sum()) / (2 * eps)
        worst = max(worst, abs(fd - G[i]) / max(1.0, abs(fd)))
    assert worst < 1e-5, worst


def test_loss_gradient_matches_finite_differences():
    """The mean's gradient is that of the squared-error term; the
    log-variance's that of the likelihood term (the whole loss in s)."""
    rng = np.random.default_rng(6)
    R, B = 8, 2
    O = rng.standard_normal((B, R // 4, R // 4, 32))
    T = rng.random((B, R, R))
    W = rng.random((B, R, R))
    alpha, inv = 0.7, 1.0 / W.sum()
    _, dO = S.loss(O, T, W, alpha, inv)
    eps = 1e-6
    for idx in [(0, 0, 0, 0), (1, 1, 1, 6), (0, 1, 0, 31), (1, 0, 1, 13), (0, 0, 1, 20)]:
        Op, Om = O.copy(), O.copy()
        Op[idx] += eps
        Om[idx] -= eps
        mp, lp = S.head_maps(Op, R)
        mm, lm = S.head_maps(Om, R)
        if idx[3] % 2 == 0:   # a mean: the squared-error term
            fd = (S.loss_terms(mp, lp, T, W, alpha)[0].sum() -
                  S.loss_terms(mm, lm, T, W, alpha)[0].sum()) / (2 * eps)
        else:
            fd = (S.loss(Op, T, W, alpha, 1.0)[0] - S.loss(Om, T, W, alpha, 1.0)[0]) / (2 * eps)
        assert dO[idx] == pytest.approx(fd * inv, rel=1e-6, abs=1e-12), idx


def test_head_layout_round_trips():
    rng = np.random.default_rng(7)
    O = rng.standard_normal((2, 3, 3, 32))
    mu, ls = S.head_maps(O, 12)
    assert np.array_equal(S.to_head(mu, ls), O)
    # output pixel (4Y+i, 4X+j) is channel (i*4 + j)*2 of head pixel (Y, X)
    assert mu[1, 4 * 2 + 3, 4 * 1 + 1] == O[1, 2, 1, (3 * 4 + 1) * 2]
    assert ls[0, 5, 10] == O[0, 1, 2, ((1 * 4) + 2) * 2 + 1]


def uniform_grid(n, z=35.0, land=True, cls=3):
    ncls = np.array([0.02, 0.025, 0.035, 0.045, 0.06, 0.08, 0.11, 0.15], np.float32)
    zz = np.full((n, n), z, np.float32)
    mask = np.zeros((n, n), np.uint8) if land else np.ones((n, n), np.uint8)
    nn_ = np.full((n, n), ncls[cls], np.float32)
    return zz, mask, nn_, ncls


def test_inputs_closed_form_on_a_uniform_grid():
    zz, mask, nn_, ncls = uniform_grid(64)
    lev = S.pyramid(zz, mask, nn_, ncls)
    sc = np.linspace(0.8, 1.2, 8)
    X, Lf = S.inputs(lev, [(3, 5, 40, 0, sc), (0, 0, 64, 1, sc)], ncls, 8)
    ch = X.reshape(2, 2, 2, 16, 4)
    assert np.allclose(ch[..., 0], (35.0 - 30.0) / 20.0)
    assert np.allclose(ch[..., 1], 0.0)     # no relief against the smooth reference
    assert np.allclose(ch[..., 2], 1.0) and np.allclose(Lf, 1.0)
    assert np.allclose(ch[..., 3], (ncls[3] * sc[3] - 0.06) / 0.04, rtol=1e-6)


def test_inputs_land_fraction_of_a_half_sea_grid():
    zz, mask, nn_, ncls = uniform_grid(64)
    mask[:, :32] = 1
    lev = S.pyramid(zz, mask, nn_, ncls)
    for level in (0, 1, 2, 3):
        _, Lf = S.inputs(lev, [(0, 0, 64, level, np.ones(8))], ncls, 8)
        assert np.allclose(Lf[0, :, :4], 0.0) and np.allclose(Lf[0, :, 4:], 1.0)


def test_relief_channel_of_a_plane_and_a_pit():
    """On a plane the bilinear reference equals the pixel mean away from the
    edges (relief 0); a one-pixel pit reads as negative relief of about its
    depth / 0.5."""
    n = 64
    ncls = np.array([0.02, 0.025, 0.035, 0.045, 0.06, 0.08, 0.11, 0.15], np.float32)
    zz = (20.0 + 0.1 * np.arange(n)[None, :] + 0.05 * np.arange(n)[:, None]).astype(np.float32)
    mask = np.zeros((n, n), np.uint8)
    nn_ = np.full((n, n), ncls[0], np.float32)
    lev = S.pyramid(zz, mask, nn_, ncls)
    X, _ = S.inputs(lev, [(0, 0, 64, 0, np.ones(8))], ncls, 64)
    rel_ = X.reshape(16, 16, 16, 4)[..., 1]
    assert np.allclose(rel_[2:-2, 2:-2], 0.0, atol=1e-5)   # float32 z
    zz[30, 30] -= 1.0
    lev = S.pyramid(zz, mask, nn_, ncls)
    X, _ = S.inputs(lev, [(0, 0, 64, 0, np.ones(8))], ncls, 64)
    pit = X[0, 30 // 4, 30 // 4, ((30 % 4) * 4 + 30 % 4) * 4 + 1]
    assert -2.0 < pit < -1.8      # the 4x4 reference cell averages 1/16 of the pit


def test_target_closed_forms():
    rng = np.random.default_rng(8)
    h = rng.random((16, 16))
    mask = (rng.random((16, 16)) < 0.3).astype(np.uint8)
    T, W = S.target(h, mask, 16, 16, scale=2.0)   # one cell per pixel
    assert np.allclose(T, np.where(mask == 0, 2.0 * h, 0.0))
    assert np.array_equal(W, (mask == 0).astype(float))
    T, W = S.target(np.full((24, 24), 0.3), np.zeros((24, 24), np.uint8), 24, 16)
    assert np.allclose(T, 0.3) and np.allclose(W, 1.0)
    # 2x2 blocks: the LAND mean of each block
    T, W = S.target(h, np.zeros((16, 16), np.uint8), 16, 8)
    assert np.allclose(T, h.reshape(8, 2, 8, 2).mean(axis=(1, 3)))


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, 1.0 + 2 ** -9, -2.5], np.float64)
    y = S.bf16(x)
    assert list(y) == [1.0, 1.0, 1.0 + 4 * 2 ** -8, 1.0, -2.5]
    assert np.array_equal(S.from_bits(S.bf16_bits(x)), y)


# ---- C-ABI ---------------------------------------------------------------------

PROBE = r"""
#include <stddef.h>
#include <stdio.h>
#include "fca_nn.h"
#define F(s, f) printf(#s "." #f " %zu %zu\n", offsetof(s, f), sizeof(((s *)0)->f));
int main(void) {
    printf("FcaNnGemm %zu\nFcaNnPyrRec %zu\nFcaNnPyr %zu\nFcaNnSample %zu\n",
           sizeof(FcaNnGemm), sizeof(FcaNnPyrRec), sizeof(FcaNnPyr), sizeof(FcaNnSample));
    %FIELDS%
    return 0;
}
"""


def test_ctypes_structs_match_the_header(tmp_path):
    structs = {"FcaNnGemm": nn.FcaNnGemm, "FcaNnPyrRec": nn.FcaNnPyrRec,
               "FcaNnPyr": nn.FcaNnPyr, "FcaNnSample": nn.FcaNnSample}
    fields = "".join(f"F({s}, {f})\n" for s, cls in structs.items() for f, _ in cls._fields_)
    src = tmp_path / "probe.c"
    src.write_text(PROBE.replace("%FIELDS%", fields))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    ref = {}
    for ln in out.splitlines():
        parts = ln.split()
        ref[parts[0]] = [int(v) for v in parts[1:]]
    for sname, cls in structs.items():
        assert C.sizeof(cls) == ref[sname][0], sname
        for fname, _ in cls._fields_:
            fld = getattr(cls, fname)
            assert [fld.offset, fld.size] == ref[f"{sname}.{fname}"], (sname, fname)


def test_nn_library_exports_every_declared_symbol():
    path = nn_lib_path()
    assert os.path.exists(path), "run python -m floodca_amd.build"
    txt = open(os.path.join(ROOT, "include", "fca_nn.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = sorted(set(re.findall(r"\b(fca_nn_\w+)\s*\(", txt)))
    assert len(names) >= 20
    syms = subprocess.run(["nm", "-D", "--defined-only", path], capture_output=True, text=True,
                          check=True).stdout.split()
    assert not [n for n in names if n not in syms]
    assert set(names) == set(nn.SIGS)
    lib = C.CDLL(path)  # loads without a GPU
    for n in names:
        getattr(lib, n)


def test_parameter_count_of_the_default_network():
    L = nn.layers(nn.NetCfg())
    total = sum(ly.nparam for ly in L)
    assert 1.9e6 < total < 2.1e6
    assert all(ly.off_w % 8 == 0 for ly in L)   # 16-byte aligned BF16 weight rows
