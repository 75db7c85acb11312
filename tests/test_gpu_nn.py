"""The surrogate's gfx950 kernels against the CPU restatement (GPU).

GEMM: every operand layout, the in-place convolution modes, split K, the
epilogue (per-sample bias, ReLU, BF16 and FP32 outputs, accumulation), against
float64 products of the same BF16 operands -- a wrong matrix-core lane layout
or index gives O(1) errors, summation order only ~K * 2^-24 relative.
Network: forward output and every parameter gradient against
oracle.surrogate_ref with the same BF16 rounding points (norm-wise relative
error; FP32 accumulation in a different order can move a stored activation by
one BF16 step). Inputs, targets, statistics and gather against the
restatement; and a few Adam steps on a fixed batch lower the loss.
"""
import numpy as np
import pytest

from floodca_amd import nn
from oracle import surrogate_ref as S
from test_surrogate_oracle import tiny_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

DEV = "cuda"


def bf_dev(x):
    """float array -> (device BF16 bits, the float64 values they hold)."""
    bits = S.bf16_bits(x)
    return torch.from_numpy(bits.view(np.int16)).to(DEV).view(torch.uint16), S.from_bits(bits)


def to_dev(x, dtype=np.float32):
    return torch.from_numpy(np.ascontiguousarray(x, dtype=dtype)).to(DEV)


def host(t):
    torch.cuda.synchronize()
    if t.dtype == torch.uint16:
        return S.from_bits(t.view(torch.int16).cpu().numpy().view(np.uint16))
    return t.cpu().numpy().astype(np.float64)


def close(got, ref, K, absA):
    """|got - ref| within FP32 accumulation over K terms (plus output rounding)."""
    bound = (K + 2) * 2.0 ** -23 * absA + 1e-30
    return np.all(np.abs(got - ref) <= bound)


@pytest.mark.parametrize("ta,tb", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(100, 70, 72), (64, 64, 32), (5, 130, 8), (257, 33, 296)])
def test_gemm_plain_layouts(ta, tb, M, N, K):
    rng = np.random.default_rng(M * 7 + N + K)
    A = rng.standard_normal((M, K))
    Bm = rng.standard_normal((N, K))
    def stored(X, t):   # [rows][K] or, transposed, [K][ld] with ld padded to 8
        if not t:
            return X, X.shape[1]
        ld = -(-X.shape[0] // 8) * 8
        out = np.full((X.shape[1], ld), 99.0)   # padding must never be read
        out[:, :X.shape[0]] = X.T
        return out, ld
    As, lda = stored(A, ta)
    Bs, ldb = stored(Bm, tb)
    Ad, Av = bf_dev(As)
    Bd, Bv = bf_dev(Bs)
    Av = Av[:, :M].T if ta else Av
    Bv = Bv[:, :N].T if tb else Bv
    Cd = torch.full((M, N), 7.0, device=DEV)
    nn.gemm(M, N, K, Ad, Bd, a_mode=ta, lda=lda, b_mode=tb, ldb=ldb, C=Cd)
    ref = Av @ Bv.T
    assert close(host(Cd), ref, K, np.abs(Av) @ np.abs(Bv).T)


def test_gemm_epilogue_bias_relu_bf16_accumulate_and_split_k():
    rng = np.random.default_rng(11)
    M, N, K, rows = 96, 48, 1024, 32
    A = rng.standard_normal((M, K))
    Bm = rng.standard_normal((N, K))
    Ad, Av = bf_dev(A)
    Bd, Bv = bf_dev(Bm)
    bias = rng.standard_normal((M // rows, N))
    ref = Av @ Bv.T
    absA = np.abs(Av) @ np.abs(Bv).T
    # per-sample bias + ReLU, both outputs
    Cd = torch.empty((M, N), device=DEV)
    Cb = torch.empty((M, N), dtype=torch.uint16, device=DEV)
    nn.gemm(M, N, K, Ad, Bd, lda=K, ldb=K, C=Cd, Cb=Cb, bias=to_dev(bias), bias_rows=rows,
            relu=1)
    want = np.maximum(ref + np.repeat(bias, rows, axis=0), 0.0)
    got = host(Cd)
    assert close(got, want, K, absA + 10)
    assert np.array_equal(host(Cb), S.bf16(got))
    # accumulate
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    Cd = to_dev(C0)
    nn.gemm(M, N, K, Ad, Bd, lda=K, ldb=K, C=Cd, accumulate=1)
    assert close(host(Cd), C0 + ref, K, absA + np.abs(C0))
    # split K: partials then the ordered sum
    chunk, nz = nn.split_plan(M, N, K, target=64, min_chunk=64)
    assert nz > 1
    part = torch.empty(nz * M * N, device=DEV)
    nn.gemm(M, N, K, Ad, Bd, lda=K, ldb=K, C=part, kchunk=chunk)
    Cd = torch.empty((M, N), device=DEV)
    nn.lib().splitk_sum(part.data_ptr(), nz, M * N, Cd.data_ptr(), 0, nn.stream())
    assert close(host(Cd), ref, K, absA)


@pytest.mark.parametrize("S_,H,Ci,Co,B", [(1, 8, 8, 16, 2), (2, 16, 16, 8, 3), (2, 2, 32, 64, 2),
                                          (1, 1, 16, 16, 4), (1, 32, 8, 8, 1)])
def test_conv_modes_against_im2col(S_, H, Ci, Co, B):
    rng = np.random.default_rng(S_ * 100 + H + Ci + Co)
    Ho = H // S_
    X = rng.standard_normal((B, H, H, Ci))
    W = rng.standard_normal((Co, 9 * Ci)) * 0.3
    dY = rng.standard_normal((B, Ho, Ho, Co))
    Xd, Xv = bf_dev(X)
    Wd, Wv = bf_dev(W)
    dYd, dYv = bf_dev(dY)
    cols = S.im2col(Xv, S_).reshape(B * Ho * Ho, 9 * Ci)
    # forward: mode 2
    Y = torch.empty((B * Ho * Ho, Co), device=DEV)
    nn.gemm(B * Ho * Ho, Co, 9 * Ci, Xd, Wd, a_mode=2, ageo=(Ho, Ho, Ci, H, H, S_), b_mode=0,
            ldb=9 * Ci, C=Y)
    assert close(host(Y), cols @ Wv.T, 9 * Ci, np.abs(cols) @ np.abs(Wv).T)
    # input gradient: mode 3 with the transposed weights
    Wt = torch.empty(9 * Ci * Co, dtype=torch.uint16, device=DEV)
    Wf = to_dev(Wv)
    nn.lib().wt(Wf.data_ptr(), Co, Ci, Wt.data_ptr(), nn.stream())
    dX = torch.empty((B * H * H, Ci), device=DEV)
    nn.gemm(B * H * H, Ci, 9 * Co, dYd, Wt, a_mode=3, ageo=(H, H, Co, Ho, Ho, S_), b_mode=0,
            ldb=9 * Co, C=dX)
    dcol = (dYv.reshape(-1, Co) @ Wv).reshape(B, Ho, Ho, 9, Ci)
    ref = S.col2im(dcol, H, H, S_).reshape(-1, Ci)
    bound = S.col2im(((np.abs(dYv).reshape(-1, Co) @ np.abs(Wv)).reshape(B, Ho, Ho, 9, Ci)), H, H,
                     S_).reshape(-1, Ci)
    assert close(host(dX), ref, 9 * Co, bound)
    # weight gradient: mode 1 x mode 4, split over pixels
    K = B * Ho * Ho
    chunk, nz = nn.split_plan(Co, 9 * Ci, K, target=64, min_chunk=32)
    part = torch.empty(nz * Co * 9 * Ci, device=DEV)
    nn.gemm(Co, 9 * Ci, K, dYd, Xd, a_mode=1, lda=Co, b_mode=4, bgeo=(Ho, Ho, Ci, H, H, S_),
            C=part, kchunk=chunk)
    dW = torch.empty((Co, 9 * Ci), device=DEV)
    nn.lib().splitk_sum(part.data_ptr(), nz, Co * 9 * Ci, dW.data_ptr(), 0, nn.stream())
    dYm = dYv.reshape(-1, Co)
    assert close(host(dW), dYm.T @ cols, K, np.abs(dYm).T @ np.abs(cols))


def net_case(cfg, B, seed):
    P, X0, s = tiny_case(seed, B=B, cfg=cfg)
    P = P.astype(np.float32).astype(np.float64)
    X0d, X0v = bf_dev(X0)
    return P, X0d, X0v, s.astype(np.float32).astype(np.float64)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("cfg,B,tol_e2e", [
    (nn.NetCfg(R=32, widths=(8, 8, 16, 16), nbott=1), 2, 3e-2),
    (nn.NetCfg(R=64, widths=(8, 16, 16, 32), nbott=2), 3, 3e-2),
    # ten random ReLU layers deep: one-step BF16 differences of the stored
    # activations compound end to end; the backward-alone check stays tight
    (nn.NetCfg(), 2, 0.15)])
def test_network_forward_and_gradients_match_the_oracle(cfg, B, tol_e2e):
    P, X0d, X0v, s = net_case(cfg, B, 21)
    net = nn.Net(cfg)
    net.load(P)
    w = net.forward(X0d, to_dev(s), B, train=True)
    ref = S.RefNet(cfg)
    c = ref.forward(P, X0v, s, emulate=True)
    O = host(w.O)
    assert rel(O, c["O"]) < 1e-2, (rel(O[..., 0::2], c["O"][..., 0::2]),
                                   rel(O[..., 1::2], c["O"][..., 1::2]))
    rng = np.random.default_rng(22)
    R = cfg.R
    T = rng.random((B, R, R)) * 2.0
    Wp = rng.random((B, R, R))
    inv = 1.0 / float(Wp.astype(np.float32).sum())
    Td, Wd = to_dev(T), to_dev(Wp)
    part = net.loss(w, Td, Wd, inv, alpha=1.0)
    # the loss kernel on the GPU's own output isolates it from the forward pass
    l_ref, dO_ref = S.loss(O, T.astype(np.float32), Wp.astype(np.float32), 1.0, inv)
    assert float(host(part).sum()) == pytest.approx(l_ref, rel=1e-4)
    dO = host(w.dO)
    assert rel(dO, dO_ref) < 1e-5, (rel(dO[..., 0::2], dO_ref[..., 0::2]),
                                    rel(dO[..., 1::2], dO_ref[..., 1::2]))
    net.backward(w, X0d)
    G = host(net.G)
    # backward alone: the restated backward over the GPU's own activations
    # (same ReLU masks and operands) and its own dO
    cg = dict(c)
    for k, t in list(w.y.items()) + list(w.u.items()):
        cg[k] = host(t)
    cg["O"] = O
    Gg = ref.backward(P, cg, dO, emulate=True)
    # end to end: the restated forward and backward from the same inputs
    Gref = ref.backward(P, c, dO, emulate=True)
    for ly in ref.L:
        for a, b in ((ly.off_w, ly.off_b), (ly.off_b, ly.off_f),
                     (ly.off_f, ly.off_f + 4 * ly.cout)):
            assert rel(G[a:b], Gg[a:b]) < 5e-3, (ly.name, a, rel(G[a:b], Gg[a:b]))
            assert rel(G[a:b], Gref[a:b]) < tol_e2e, (ly.name, a, rel(G[a:b], Gref[a:b]))


def test_adam_steps_lower_the_loss_on_a_fixed_batch():
    """The target is a function of the input at the same pixel (its first
    channel, clipped at 0): 80 Adam steps must cut the squared error steadily."""
    cfg = nn.NetCfg(R=64, widths=(32, 16, 16, 32), nbott=1)
    B = 4
    _, X0d, X0v, s = net_case(cfg, B, 31)
    net = nn.Net(cfg, seed=31)
    ch0 = X0v.reshape(B, 16, 16, 4, 4, 4)[..., 0].transpose(0, 1, 3, 2, 4).reshape(B, 64, 64)
    T = to_dev(np.clip(ch0, 0.0, None))
    Wp = to_dev(np.ones((B, 64, 64)))
    inv = 1.0 / (B * 64 * 64)
    se = torch.empty((B, 2), dtype=torch.float64, device=DEV)
    mse = []
    for it in range(81):
        w = net.forward(X0d, to_dev(s), B, train=True)
        nn.lib().sq_err(w.O.data_ptr(), T.data_ptr(), Wp.data_ptr(), B, 64, 1.0, se.data_ptr(),
                        nn.stream())
        e = host(se)
        mse.append(e[:, 0].sum() / e[:, 1].sum())
        if it == 80:
            break
        net.loss(w, T, Wp, inv)
        net.backward(w, X0d)
        net.adam(3e-3)
    assert np.isfinite(mse).all()
    assert mse[-1] < 0.6 * mse[0] and mse[80] < mse[40] < mse[0], mse[::10]


def test_pyramid_inputs_and_targets_match_the_restatement():
    rng = np.random.default_rng(41)
    n = 128
    z = (30 + 10 * rng.standard_normal((n, n))).astype(np.float32)
    mask = (rng.random((n, n)) < 0.3).astype(np.uint8)
    ncls = np.array([0.02, 0.025, 0.035, 0.045, 0.06, 0.08, 0.11, 0.15], np.float32)
    n_ = ncls[rng.integers(0, 8, (n, n))]
    lev = S.pyramid(z, mask, n_, ncls)
    NL = nn.LEVELS
    recs = []
    for k in range(NL):
        recs.append(torch.empty(((n >> k) * (n >> k), 16), dtype=torch.uint8, device=DEV))
    L = nn.lib()
    st = nn.stream()
    dn = to_dev(ncls)
    zd, md, nd = to_dev(z), to_dev(mask, np.uint8), to_dev(n_)
    L.pyr0(zd.data_ptr(), md.data_ptr(), nd.data_ptr(), n * n, dn.data_ptr(), 8,
           recs[0].data_ptr(), st)
    for k in range(1, NL):
        m = n >> (k - 1)
        L.pyr_down(recs[k - 1].data_ptr(), m, m, recs[k].data_ptr(), st)
    for k in range(NL):
        r = recs[k].cpu().numpy().view(np.dtype([("zmean", "<f4"), ("zmin", "<f4"),
                                                  ("zmax", "<f4"), ("land", "<u2"), ("cls", "u1"),
                                                  ("pad", "u1")]))[:, 0]
        m = n >> k
        assert np.allclose(r["zmean"].reshape(m, m), lev[k]["zmean"], rtol=1e-6)
        assert np.array_equal(r["zmin"].reshape(m, m), lev[k]["zmin"].astype(np.float32))
        assert np.array_equal(r["land"].reshape(m, m), lev[k]["land"])
        assert np.array_equal(r["cls"].reshape(m, m), lev[k]["cls"])
    pyr = nn.FcaNnPyr()
    for k in range(NL):
        pyr.lev[k] = recs[k].data_ptr()
    pyr.nr = pyr.nc = n
    R = 16
    samples = [(0, 0, 128, 3, np.ones(8)), (5, 17, 37, 1, np.linspace(0.8, 1.2, 8)),
               (90, 3, 16, 0, np.full(8, 1.1)), (33, 70, 51, 1, np.ones(8) * 0.9)]
    arr = (nn.FcaNnSample * len(samples))()
    for i, (r0, c0, side, level, sc) in enumerate(samples):
        arr[i].r0, arr[i].c0, arr[i].side, arr[i].level = r0, c0, side, level
        arr[i].nscale[:] = [float(v) for v in sc]
    smp = torch.frombuffer(bytearray(bytes(arr)), dtype=torch.uint8).to(DEV)
    X = torch.empty((len(samples), R // 4, R // 4, 64), dtype=torch.uint16, device=DEV)
    Lf = torch.empty((len(samples), R, R), device=DEV)
    L.inputs(nn.ct.byref(pyr), smp.data_ptr(), len(samples), dn.data_ptr(), R, X.data_ptr(),
             Lf.data_ptr(), st)
    Xr, Lr = S.inputs(lev, samples, ncls.astype(np.float64), R)
    got = host(X)
    assert np.all(np.abs(got - Xr) <= 2.0 ** -7 * np.abs(Xr) + 1e-5)
    assert np.allclose(host(Lf), Lr, rtol=1e-5, atol=1e-6)
    # targets: crops of a random hmax field
    for side in (16, 23, 64, 128):
        hm = rng.random((side, side)).astype(np.float32)
        mk = (rng.random((side, side)) < 0.4).astype(np.uint8)
        T = torch.empty((R, R), device=DEV)
        W = torch.empty((R, R), device=DEV)
        hd, kd = to_dev(hm), to_dev(mk, np.uint8)
        L.target(hd.data_ptr(), kd.data_ptr(), side, R, 4.0, 0.5, T.data_ptr(), W.data_ptr(),
                 st)
        Tr, Wr = S.target(hm, mk, side, R, 4.0, 0.5)
        assert np.allclose(host(T), Tr, rtol=1e-6, atol=1e-7), side
        assert np.allclose(host(W), Wr, rtol=1e-6, atol=1e-7), side


def test_prediction_statistics_squared_error_and_gather():
    rng = np.random.default_rng(51)
    B, R = 3, 16
    O = rng.standard_normal((B, R // 4, R // 4, 32)).astype(np.float32)
    Lf = rng.random((B, R, R)).astype(np.float32)
    T = rng.random((B, R, R)).astype(np.float32)
    L = nn.lib()
    st = nn.stream()
    Od, Lfd = to_dev(O), to_dev(Lf)
    out = torch.empty((B, 4), device=DEV)
    ys = np.array([0.5, 0.25, 2.0], np.float32)
    ysd = to_dev(ys)
    L.pred_stats(Od.data_ptr(), Lfd.data_ptr(), B, R, ysd.data_ptr(), 0.1, out.data_ptr(), st)
    mu, ls = S.head_maps(O.astype(np.float64), R)
    sig = np.exp(0.5 * np.clip(ls, S.LS_LO, S.LS_HI)) * ys[:, None, None]
    mu = mu * ys[:, None, None]
    want = np.stack([(Lf * (mu >= 0.1)).sum((1, 2)), (Lf * sig).sum((1, 2)),
                     (Lf * np.maximum(mu, 0)).sum((1, 2)), Lf.sum((1, 2))], 1)
    assert np.allclose(host(out), want, rtol=1e-5)
    se = torch.empty((B, 2), dtype=torch.float64, device=DEV)
    Td = to_dev(T)
    L.sq_err(Od.data_ptr(), Td.data_ptr(), Lfd.data_ptr(), B, R, 0.5, se.data_ptr(), st)
    mu = S.head_maps(O.astype(np.float64), R)[0] * 0.5
    Lf64 = Lf.astype(np.float64)
    want = np.stack([(Lf64 * (np.maximum(mu, 0) - T) ** 2).sum((1, 2)), Lf64.sum((1, 2))], 1)
    assert np.allclose(host(se), want, rtol=1e-9)
    src = rng.standard_normal((10, 12)).astype(np.float32)
    idx = np.array([3, 3, 9, 0, 5], np.int32)
    dst = torch.empty((5, 12), device=DEV)
    sd, idd = to_dev(src), to_dev(idx, np.int32)
    L.gather(sd.data_ptr(), idd.data_ptr(), 5, 48, dst.data_ptr(), st)
    assert np.array_equal(host(dst), src[idx].astype(np.float64))
