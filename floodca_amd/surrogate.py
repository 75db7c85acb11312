"""Surrogate-screened scenario ensembles on one GPU.

The pipeline (no counterpart in the reference; the plain CA is its judge and
its fallback):

1. Dataset: scenarios from the ensemble generator (floodca_amd.ensemble.draw:
   crops of one large DEM, storm intensity, duration and roughness drawn per
   scenario) simulated by the batched GPU CA; each scenario's hmax is reduced
   on the device, without a copy to the host, to an R x R target map (the
   LAND-mean hmax over each pixel's footprint) with LAND-fraction weights.
2. Inputs: an R x R map per scenario sampled from a terrain pyramid of the
   DEM (mean z, z against a smooth local reference, LAND fraction, Manning n
   of the scenario's classes), plus four forcing scalars that modulate every
   layer's bias.
3. Training: floodca_amd.nn's network with the library's own forward,
   backward and Adam kernels; the mean learns from the squared error, the
   log-variance from the Gaussian likelihood. The network predicts hmax in
   units of the scenario's storm depth D (hmax / D, about 1 where rain ponds
   where it falls), and each scenario's pixel weights carry (D / D0)^2, so the
   loss is the squared error in metres up to the constant D0^2.
4. Screening: the network predicts the hmax map and its per-pixel spread for
   every candidate; the predicted flooded area (LAND area with predicted
   hmax >= 0.10 m) and the mean predicted spread rank the candidates.
5. The CA reruns the top fraction by predicted flooded area plus the top
   fraction by uncertainty among the rest; their CA results replace the
   predictions, and the error of the predictions on these CA-judged members
   is reported.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import ensemble as E
from . import nn
from .grid import Grid
from .hip import Batch
from .synth import MANNING_CLASSES

D0 = 0.35          # reference storm depth of the loss weights (m)
THR_FLOOD = 0.10   # flooded threshold of the screening statistic (m)
PHI = 0.5          # storm peak position of the ensemble generator


def _torch():
    return nn._torch()


def depth(sc: E.Scenario) -> float:
    """Storm depth (m): 50 mm/h x intensity factor x duration
    (floodca_amd.ensemble.build)."""
    return 0.05 * sc.factor * sc.dur_h


def scalars(sc: E.Scenario, phi: float = PHI) -> list:
    """The four forcing scalars: log intensity factor, duration, peak
    position, log storm depth (each roughly unit scale)."""
    return [math.log(sc.factor) / 0.5, (sc.dur_h - 7.0) / 2.9, (phi - 0.5) / 0.2,
            math.log(depth(sc) / D0) / 0.6]


def level_for(side: int, R: int) -> int:
    """Pyramid level whose cells are at most one pixel footprint across
    (at most 3: a footprint of up to 8 cells)."""
    f = side / R
    return int(max(0, min(3, math.floor(math.log2(f)) if f >= 1 else 0)))


class Terrain:
    """The DEM's record pyramid on the device (fca_nn_pyr0 / fca_nn_pyr_down)."""

    def __init__(self, base: Grid):
        torch = _torch()
        nr, nc = base.shape
        step = 1 << nn.LEVELS
        if nr % step or nc % step:
            raise ValueError(f"grid sides must be multiples of {step}")
        L = nn.lib()
        st = nn.stream()
        self.base = base
        self.ncls = torch.from_numpy(MANNING_CLASSES.astype(np.float32)).cuda()
        z = torch.from_numpy(np.ascontiguousarray(base.z, dtype=np.float32)).cuda()
        m = torch.from_numpy(np.ascontiguousarray(base.mask)).cuda()
        n = torch.from_numpy(np.ascontiguousarray(base.n, dtype=np.float32)).cuda()
        self.lev = []
        for k in range(nn.LEVELS):
            self.lev.append(torch.empty(((nr >> k) * (nc >> k), 16), dtype=torch.uint8,
                                        device="cuda"))
        L.pyr0(z.data_ptr(), m.data_ptr(), n.data_ptr(), nr * nc, self.ncls.data_ptr(),
               MANNING_CLASSES.size, self.lev[0].data_ptr(), st)
        for k in range(1, nn.LEVELS):
            L.pyr_down(self.lev[k - 1].data_ptr(), nr >> (k - 1), nc >> (k - 1),
                       self.lev[k].data_ptr(), st)
        torch.cuda.synchronize()
        del z, m, n
        self.pyr = nn.FcaNnPyr()
        for k in range(nn.LEVELS):
            self.pyr.lev[k] = self.lev[k].data_ptr()
        self.pyr.nr, self.pyr.nc = nr, nc

    def upload(self, scens, R, pad_to=1):
        """Device sample table (FcaNnSample rows), scalars [n][4] and storm
        depths [n] of scens, padded with copies of the last row to a multiple
        of pad_to rows."""
        torch = _torch()
        n = len(scens)
        m = -(-n // pad_to) * pad_to
        tab = np.zeros(m, nn.SAMPLE_DTYPE)
        sv = np.zeros((m, 4), np.float32)
        dv = np.zeros(m, np.float32)
        for i, sc in enumerate(scens):
            tab[i] = (sc.r0, sc.c0, sc.side, level_for(sc.side, R), sc.nscale)
            sv[i] = scalars(sc)
            dv[i] = depth(sc)
        tab[n:] = tab[n - 1]
        sv[n:] = sv[n - 1]
        dv[n:] = dv[n - 1]
        smp = torch.from_numpy(tab.view(np.uint8)).cuda()
        return smp, torch.from_numpy(sv).cuda(), torch.from_numpy(dv).cuda()

    def build(self, smp, row0, k, R, X, Lf):
        """Inputs of sample-table rows row0..row0+k into X [k][R/4][R/4][64]
        (BF16) and Lf [k][R][R]."""
        nn.lib().inputs(nn.ct.byref(self.pyr), smp.data_ptr() + row0 * nn.SAMPLE_DTYPE.itemsize,
                        k, self.ncls.data_ptr(), R, X.data_ptr(), Lf.data_ptr(), nn.stream())


def row_moments(T, W, rows=None):
    """Per row (sum W, sum W*T, sum W*T^2) of [N][R][R] device maps (host
    float64 [N][3])."""
    torch = _torch()
    N = T.shape[0] if rows is None else rows
    mo = torch.empty((N, 3), dtype=torch.float64, device="cuda")
    nn.lib().row_moments(T.data_ptr(), W.data_ptr(), N, T[0].numel(), mo.data_ptr(),
                         nn.stream())
    return mo.cpu().numpy()


@dataclass
class Dataset:
    R: int
    X: object           # [N][R/4][R/4][64] BF16 inputs
    S: object           # [N][4] scalars
    T: object           # [N][R][R] target hmax / storm depth
    W: object           # [N][R][R] LAND fraction x (depth / D0)^2
    wsum: np.ndarray    # [N] sum of W per sample
    scens: list
    ca: list            # per scenario (flooded counts at 0.10/0.25/0.50 m, max depth, steps)
    t_ca: float = 0.0   # seconds in batched CA sessions (upload, run, reductions, release)
    t_target: float = 0.0
    t_inputs: float = 0.0
    mo: np.ndarray = None  # [N][3] sum W, sum W*T, sum W*T^2


def ca_targets(base: Grid, scens, R, T, W, row0=0, batch=512, prec="f32", strip_rows=0,
               timing=None, log=None, dbase=None):
    """Runs scens through the batched CA (chunks of `batch`) and writes their
    targets into rows row0.. of T and W (device): hmax / storm depth, LAND
    fraction x (depth / D0)^2. Returns the per-scenario CA results
    [(flooded, max depth, steps)]. log, if given, receives one line per
    chunk. With dbase (ensemble.device_base of base) the scenarios are built
    on the device; otherwise on the host and uploaded."""
    torch = _torch()
    L = nn.lib()
    out = []
    t_ca = t_tg = 0.0
    for c0 in range(0, len(scens), batch):
        chunk = scens[c0:c0 + batch]
        t0 = time.perf_counter()
        if dbase is not None:
            cases = None
            b = Batch.from_crops(dbase, [E.crop_of(sc, prec) for sc in chunk],
                                 [E.config_of(sc, base.dx, prec) for sc in chunk], prec=prec,
                                 strip_rows=strip_rows)
        else:
            cases = [E.build(base, sc, prec) for sc in chunk]
            accs = [E.acc_start(prec, g, c) for g, c in cases]
            b = Batch([g for g, _ in cases], [c for _, c in cases], accs, prec=prec,
                      strip_rows=strip_rows)
        try:
            b.run()
            t1 = time.perf_counter()
            st = nn.stream()
            for i in range(len(chunk)):
                hm, mk, nr, nc = b.device_fields(i)
                if nr != nc:
                    raise ValueError("scenario crops are square")
                row = row0 + c0 + i
                d = depth(chunk[i])
                L.target(hm, mk, nr, R, 1.0 / d, (d / D0) ** 2, T[row].data_ptr(),
                         W[row].data_ptr(), st)
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            for i in range(len(chunk)):
                fl, md = b.flooded(i, E.THRESHOLDS)
                out.append((fl, md, b.steps(i)))
        finally:
            b.close()
        t3 = time.perf_counter()
        t_ca += (t1 - t0) + (t3 - t2)
        t_tg += t2 - t1
        del cases
        if log is not None:
            log(f"CA chunk {c0 // batch + 1}/{-(-len(scens) // batch)}: {len(chunk)} "
                f"scenarios in {t3 - t0:.1f}s")
    if timing is not None:
        timing["ca"] = timing.get("ca", 0.0) + t_ca
        timing["target"] = timing.get("target", 0.0) + t_tg
    return out


def make_dataset(base: Grid, terrain: Terrain, scens, R=256, batch=512, prec="f32",
                 strip_rows=0, log=None, dbase=None) -> Dataset:
    torch = _torch()
    N = len(scens)
    q = R // 4
    X = torch.empty((N, q, q, 64), dtype=torch.uint16, device="cuda")
    S = torch.empty((N, 4), dtype=torch.float32, device="cuda")
    T = torch.empty((N, R, R), dtype=torch.float32, device="cuda")
    W = torch.empty((N, R, R), dtype=torch.float32, device="cuda")
    timing = {}
    ca = ca_targets(base, scens, R, T, W, 0, batch, prec, strip_rows, timing, log, dbase)
    t0 = time.perf_counter()
    smp, S, _ = terrain.upload(scens, R)
    Lf = torch.empty((min(N, 512), R, R), dtype=torch.float32, device="cuda")
    for c0 in range(0, N, 512):
        terrain.build(smp, c0, min(512, N - c0), R, X[c0], Lf)
    torch.cuda.synchronize()
    t_in = time.perf_counter() - t0
    mo = row_moments(T, W)
    return Dataset(R=R, X=X, S=S, T=T, W=W, wsum=mo[:, 0], scens=list(scens), ca=ca,
                   t_ca=timing["ca"], t_target=timing["target"], t_inputs=t_in, mo=mo)


@dataclass
class TrainLog:
    steps: int = 0
    seconds: float = 0.0
    loss: list = field(default_factory=list)     # (step, mean loss per weight)
    val_rmse: list = field(default_factory=list)  # (epoch, metres)
    train_rmse: list = field(default_factory=list)  # (epoch, metres) on a fixed subset


class Trainer:
    """Mini-batch Adam over a Dataset's training rows."""

    def __init__(self, net: nn.Net, data: Dataset, train_idx, B=64, seed=0):
        torch = _torch()
        self.net, self.data, self.B = net, data, B
        self.idx = np.asarray(train_idx, dtype=np.int32)
        self.rng = np.random.default_rng(seed)
        R, q = data.R, data.R // 4
        self.bX = torch.empty((B, q, q, 64), dtype=torch.uint16, device="cuda")
        self.bS = torch.empty((B, 4), dtype=torch.float32, device="cuda")
        self.bT = torch.empty((B, R, R), dtype=torch.float32, device="cuda")
        self.bW = torch.empty((B, R, R), dtype=torch.float32, device="cuda")
        self.bI = torch.empty(B, dtype=torch.int32, device="cuda")
        self.log = TrainLog()

    def _gather(self, rows):
        L = nn.lib()
        st = nn.stream()
        d = self.data
        self.bI.copy_(_torch().from_numpy(rows))
        n = len(rows)
        for src, dst in ((d.X, self.bX), (d.S, self.bS), (d.T, self.bT), (d.W, self.bW)):
            row_bytes = src[0].numel() * src.element_size()
            L.gather(src.data_ptr(), self.bI.data_ptr(), n, row_bytes, dst.data_ptr(), st)

    def epoch(self, lr: float, alpha: float = 1.0, log_every: int = 0):
        torch = _torch()
        t0 = time.perf_counter()
        perm = self.rng.permutation(self.idx)
        nb = len(perm) // self.B
        for j in range(nb):
            rows = perm[j * self.B:(j + 1) * self.B]
            self._gather(rows)
            ws = float(self.data.wsum[rows].sum())
            w = self.net.forward(self.bX, self.bS, self.B, train=True)
            part = self.net.loss(w, self.bT, self.bW, 1.0 / max(ws, 1e-9), alpha)
            self.net.backward(w, self.bX)
            self.net.adam(lr)
            self.log.steps += 1
            if log_every and self.log.steps % log_every == 0:
                lp = float(part.cpu().numpy().astype(np.float64).sum())
                self.log.loss.append((self.log.steps, lp / max(ws, 1e-9)))
        torch.cuda.synchronize()
        self.log.seconds += time.perf_counter() - t0


def to_metres(se, scens):
    """Per-row (sum W*(pred - T)^2, sum W) in the depth units of the dataset
    -> (sum LAND fraction x (pred - hmax)^2 in m^2, sum LAND fraction)."""
    se = np.array(se, dtype=np.float64)
    r = np.array([(depth(sc) / D0) ** 2 for sc in scens])
    se[:, 0] *= D0 * D0
    se[:, 1] /= r
    return se


def predict_rmse(net: nn.Net, data: Dataset, rows, B=64):
    """Per-row (sum Lf*(max(mu,0) - hmax)^2 in m^2, sum Lf) for dataset rows."""
    torch = _torch()
    L = nn.lib()
    rows = np.asarray(rows, dtype=np.int32)
    R, q = data.R, data.R // 4
    bX = torch.empty((B, q, q, 64), dtype=torch.uint16, device="cuda")
    bS = torch.empty((B, 4), dtype=torch.float32, device="cuda")
    bT = torch.empty((B, R, R), dtype=torch.float32, device="cuda")
    bW = torch.empty((B, R, R), dtype=torch.float32, device="cuda")
    bI = torch.empty(B, dtype=torch.int32, device="cuda")
    se = torch.empty((B, 2), dtype=torch.float64, device="cuda")
    out = np.zeros((len(rows), 2))
    for c0 in range(0, len(rows), B):
        r = rows[c0:c0 + B]
        n = len(r)
        bI[:n].copy_(torch.from_numpy(r))
        st = nn.stream()
        for src, dst in ((data.X, bX), (data.S, bS), (data.T, bT), (data.W, bW)):
            L.gather(src.data_ptr(), bI.data_ptr(), n, src[0].numel() * src.element_size(),
                     dst.data_ptr(), st)
        w = net.forward(bX, bS, B)
        L.sq_err(w.O.data_ptr(), bT.data_ptr(), bW.data_ptr(), n, R, 1.0, se.data_ptr(), st)
        out[c0:c0 + n] = se[:n].cpu().numpy()
    return to_metres(out, [data.scens[i] for i in rows])


def rmse(se) -> float:
    se = np.asarray(se)
    return float(math.sqrt(se[:, 0].sum() / max(se[:, 1].sum(), 1e-30)))


@dataclass
class Screen:
    area: np.ndarray      # predicted flooded LAND area (m^2)
    sigma: np.ndarray     # LAND-mean predicted spread of hmax (m)
    volume: np.ndarray    # predicted LAND-mean hmax x LAND area (m^3)
    seconds: float


def screen(net: nn.Net, terrain: Terrain, scens, R=256, B=256, dx=None) -> Screen:
    """Predicted statistics of every candidate: the sample table goes up
    once, then per chunk of B the inputs, the forward pass and the
    reductions run on the device; four floats per candidate come back at the
    end. The time covers all of it."""
    torch = _torch()
    L = nn.lib()
    dx = dx or terrain.base.dx
    q = R // 4
    n = len(scens)
    X = torch.empty((B, q, q, 64), dtype=torch.uint16, device="cuda")
    Lf = torch.empty((B, R, R), dtype=torch.float32, device="cuda")
    st_all = torch.empty((-(-n // B) * B, 4), dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    smp, S, D = terrain.upload(scens, R, pad_to=B)
    for c0 in range(0, n, B):
        k = min(B, n - c0)
        terrain.build(smp, c0, k, R, X, Lf)
        w = net.forward(X, S[c0:c0 + B], B)
        L.pred_stats(w.O.data_ptr(), Lf.data_ptr(), k, R, D[c0].data_ptr(), THR_FLOOD,
                     st_all[c0].data_ptr(), nn.stream())
    raw = st_all[:n].cpu().numpy().astype(np.float64)
    secs = time.perf_counter() - t0
    pix = np.array([(sc.side / R * dx) ** 2 for sc in scens])
    area = raw[:, 0] * pix
    sigma = raw[:, 1] / np.maximum(raw[:, 3], 1e-30)
    volume = raw[:, 2] * pix
    return Screen(area=area, sigma=sigma, volume=volume, seconds=secs)


def select(sc: Screen, frac_area=0.02, frac_unc=0.02):
    """Indices of the top frac_area by predicted flooded area, then the top
    frac_unc by mean predicted spread among the rest (each count rounded up)."""
    n = len(sc.area)
    ka, ku = int(math.ceil(frac_area * n)), int(math.ceil(frac_unc * n))
    order = np.lexsort((np.arange(n), -sc.area))
    top_a = order[:ka]
    rest = np.setdiff1d(np.arange(n), top_a)
    order_u = rest[np.lexsort((rest, -sc.sigma[rest]))]
    return top_a, order_u[:ku]


def audit_sample(n: int, flagged, frac: float, seed: int) -> np.ndarray:
    """A seeded uniform sample (count rounded up, sorted) of the members not
    in flagged: the CA runs them so that the error of the predictions the
    pipeline delivers without a rerun is measured, not inferred from the
    validation rows."""
    rest = np.setdiff1d(np.arange(n), np.asarray(flagged))
    k = min(len(rest), int(math.ceil(frac * n)))
    if k == 0:
        return np.zeros(0, dtype=np.int64)
    rng = np.random.default_rng(seed)
    return np.sort(rng.choice(rest, size=k, replace=False))


def judged_error(net: nn.Net, terrain: Terrain, base: Grid, scens, R=256, batch=512,
                 strip_rows=0, timing=None, log=None, dbase=None):
    """CA reruns of scens: returns (per-scenario CA results, per-scenario
    (sum Lf*(pred - hmax)^2 in m^2, sum Lf) of the surrogate's prediction
    against the CA's target, per-scenario target moments as row_moments in
    the dataset's depth units)."""
    torch = _torch()
    N = len(scens)
    q = R // 4
    T = torch.empty((N, R, R), dtype=torch.float32, device="cuda")
    W = torch.empty((N, R, R), dtype=torch.float32, device="cuda")
    ca = ca_targets(base, scens, R, T, W, 0, batch, "f32", strip_rows, timing, log, dbase)
    B = 64
    X = torch.empty((B, q, q, 64), dtype=torch.uint16, device="cuda")
    Lf = torch.empty((B, R, R), dtype=torch.float32, device="cuda")
    se = torch.empty((-(-N // B) * B, 2), dtype=torch.float64, device="cuda")
    smp, S, _ = terrain.upload(scens, R, pad_to=B)
    for c0 in range(0, N, B):
        k = min(B, N - c0)
        terrain.build(smp, c0, k, R, X, Lf)
        w = net.forward(X, S[c0:c0 + B], B)
        nn.lib().sq_err(w.O.data_ptr(), T[c0].data_ptr(), W[c0].data_ptr(), k, R, 1.0,
                        se[c0].data_ptr(), nn.stream())
    out = to_metres(se[:N].cpu().numpy(), scens)
    mo = row_moments(T, W)
    return ca, out, mo


def depth_fit(data: Dataset, rows) -> float:
    """The method without the network for hmax maps: hmax = c x storm depth
    on every LAND pixel, c fitted by least squares (in metres) over rows."""
    mo = data.mo[np.asarray(rows)]
    return float(mo[:, 1].sum() / max(mo[:, 0].sum(), 1e-30))


def depth_error(c: float, scens, mo) -> np.ndarray:
    """Per-row (sum Lf*(c*depth - hmax)^2 in m^2, sum Lf) from target moments
    in the dataset's depth units."""
    se = mo[:, 2] - 2 * c * mo[:, 1] + c * c * mo[:, 0]
    return to_metres(np.stack([np.maximum(se, 0.0), mo[:, 0]], 1), scens)


@dataclass
class PipelineCfg:
    n_train: int = 4096        # CA runs of the dataset, validation rows included
    n_val: int = 256
    n_cand: int = 100_000
    frac_area: float = 0.02
    frac_unc: float = 0.02
    frac_audit: float = 0.01   # unflagged members the CA also runs, to judge the rest
    n_alone: int = 64          # candidates timed one session each (the all-CA baseline)
    epochs: int = 30
    batch: int = 64
    lr: float = 2e-3
    R: int = 256
    widths: tuple = (32, 64, 128, 256)
    nbott: int = 2
    side: tuple = (256, 2048)
    dur: tuple = (2.0, 12.0)
    seed: int = 20261206
    ca_batch: int = 512
    screen_batch: int = 256
    # spread of the training storms' log intensity (the candidates keep the
    # generator's 0.5): a wider training draw covers the deep-storm tail the
    # spread ranking flags
    train_sigma: float = 0.5


def rmse_interval(se, n_boot: int = 2000, seed: int = 0, q=(0.025, 0.975)):
    """Percentile bootstrap interval of rmse(se) over its rows (scenarios
    resampled with replacement, seeded)."""
    se = np.asarray(se, dtype=np.float64)
    if len(se) == 0:
        return (float("nan"), float("nan"))
    rng = np.random.default_rng(seed)
    idx = rng.integers(0, len(se), size=(n_boot, len(se)))
    num = se[idx, 0].sum(axis=1)
    den = np.maximum(se[idx, 1].sum(axis=1), 1e-30)
    r = np.sqrt(num / den)
    return (float(np.quantile(r, q[0])), float(np.quantile(r, q[1])))


def _spearman(a, b) -> float:
    ra = np.argsort(np.argsort(a)).astype(float)
    rb = np.argsort(np.argsort(b)).astype(float)
    ra -= ra.mean()
    rb -= rb.mean()
    return float((ra * rb).sum() / math.sqrt((ra * ra).sum() * (rb * rb).sum()))


def _topk_overlap(score, truth, k) -> float:
    a = set(np.argsort(-np.asarray(score), kind="stable")[:k].tolist())
    b = set(np.argsort(-np.asarray(truth), kind="stable")[:k].tolist())
    return len(a & b) / k


def run_pipeline(base: Grid, pc: PipelineCfg = PipelineCfg(), log=print) -> dict:
    """Dataset, training, screening, CA reruns, and the baselines: the
    per-scenario all-CA rate (a seeded sample of candidates run one session
    each), the batched all-CA rate (the dataset's CA phase), and the method
    without the network (hmax = c x storm depth; ranking by storm depth x
    LAND area)."""
    torch = _torch()
    res = {}
    dx = base.dx
    t0 = time.perf_counter()
    terrain = Terrain(base)
    dbase = E.device_base(base)
    res["t_terrain_s"] = time.perf_counter() - t0
    scens = E.draw(pc.n_train, pc.seed, base.shape, side=pc.side, dur=pc.dur,
                   sigma=pc.train_sigma)
    t0 = time.perf_counter()
    data = make_dataset(base, terrain, scens, pc.R, pc.ca_batch,
                        log=lambda m: log(f"surrogate: dataset {m}"), dbase=dbase)
    res["t_dataset_s"] = time.perf_counter() - t0
    res["dataset_ca_s"] = data.t_ca
    log(f"surrogate: dataset of {pc.n_train} CA runs in {res['t_dataset_s']:.1f}s "
        f"(CA {data.t_ca:.1f}s, targets {data.t_target:.2f}s, inputs {data.t_inputs:.2f}s)")
    tr = np.arange(pc.n_train - pc.n_val)
    va = np.arange(pc.n_train - pc.n_val, pc.n_train)
    net = nn.Net(nn.NetCfg(R=pc.R, widths=tuple(pc.widths), nbott=pc.nbott), seed=pc.seed)
    trainer = Trainer(net, data, tr, B=pc.batch, seed=pc.seed)
    t0 = time.perf_counter()
    for ep in range(pc.epochs):
        lr = pc.lr * (0.05 + 0.95 * 0.5 * (1 + math.cos(math.pi * ep / max(pc.epochs, 1))))
        trainer.epoch(lr, log_every=50)
        if ep == 0 or (ep + 1) % max(1, pc.epochs // 5) == 0 or ep + 1 == pc.epochs:
            v = rmse(predict_rmse(net, data, va))
            t = rmse(predict_rmse(net, data, tr[:len(va)]))
            trainer.log.val_rmse.append((ep + 1, v))
            trainer.log.train_rmse.append((ep + 1, t))
            log(f"surrogate: epoch {ep + 1}/{pc.epochs} lr {lr:.2e} train hmax RMSE "
                f"{t * 100:.2f} cm, validation {v * 100:.2f} cm ({trainer.log.seconds:.1f}s)")
    torch.cuda.synchronize()
    res["t_train_s"] = time.perf_counter() - t0
    res["train_steps"] = trainer.log.steps
    res["params"] = net.n
    val_se = predict_rmse(net, data, va)
    c_dep = depth_fit(data, tr)
    dep_se = depth_error(c_dep, [data.scens[i] for i in va], data.mo[va])
    res["val_rmse_m"] = rmse(val_se)
    res["val_rmse_depth_baseline_m"] = rmse(dep_se)
    # ranking on validation, where the CA truth is known
    true_area = np.array([data.ca[i][0][0] for i in va], float) * dx * dx
    vs = screen(net, terrain, [data.scens[i] for i in va], pc.R, pc.screen_batch)
    land = np.array([data.mo[i, 0] / (depth(data.scens[i]) / D0) ** 2 for i in va]) / pc.R ** 2
    heur = np.array([depth(data.scens[i]) * data.scens[i].side ** 2 for i in va]) * land
    k = max(1, int(round(0.1 * len(va))))
    res["val_rank_spearman"] = _spearman(vs.area, true_area)
    res["val_rank_spearman_depth_baseline"] = _spearman(heur, true_area)
    res["val_top10pct_overlap"] = _topk_overlap(vs.area, true_area, k)
    res["val_top10pct_overlap_depth_baseline"] = _topk_overlap(heur, true_area, k)
    log(f"surrogate: validation RMSE {res['val_rmse_m'] * 100:.2f} cm (depth baseline "
        f"{res['val_rmse_depth_baseline_m'] * 100:.2f} cm), area rank Spearman "
        f"{res['val_rank_spearman']:.3f} (baseline {res['val_rank_spearman_depth_baseline']:.3f})")
    # screening
    cands = E.draw(pc.n_cand, pc.seed + 1, base.shape, side=pc.side, dur=pc.dur)
    sc = screen(net, terrain, cands, pc.R, pc.screen_batch)
    res["t_screen_s"] = sc.seconds
    top_a, top_u = select(sc, pc.frac_area, pc.frac_unc)
    audit = audit_sample(pc.n_cand, np.concatenate([top_a, top_u]), pc.frac_audit, pc.seed + 3)
    judged = np.concatenate([top_a, top_u, audit])
    log(f"surrogate: screened {pc.n_cand} candidates in {sc.seconds:.2f}s; CA reruns "
        f"{len(top_a)} by area + {len(top_u)} by uncertainty + {len(audit)} unflagged (audit)")
    timing = {}
    t0 = time.perf_counter()
    ca, se, mo = judged_error(net, terrain, base, [cands[i] for i in judged], pc.R,
                              pc.ca_batch, timing=timing,
                              log=lambda m: log(f"surrogate: rerun {m}"), dbase=dbase)
    res["t_rerun_s"] = time.perf_counter() - t0
    na, nf = len(top_a), len(top_a) + len(top_u)
    dep = depth_error(c_dep, [cands[i] for i in judged], mo)
    res["judged"] = int(len(judged))
    res["judged_audit"] = int(len(audit))
    res["judged_rmse_m"] = rmse(se)
    res["judged_rmse_area_m"] = rmse(se[:na])
    res["judged_rmse_uncertainty_m"] = rmse(se[na:nf])
    res["judged_rmse_audit_m"] = rmse(se[nf:]) if len(audit) else float("nan")
    res["judged_rmse_audit_95ci_m"] = list(rmse_interval(se[nf:], seed=pc.seed))
    # the error of what the pipeline hands back, judged by the CA: the flagged
    # members return the CA's own maps (no error), the unflagged ones the
    # network's, whose error the audit measures. Over the judged members:
    se_arr = np.asarray(se, dtype=np.float64)
    w_all = max(float(se_arr[:, 1].sum()), 1e-30)
    res["judged_delivered_rmse_m"] = (float(math.sqrt(se_arr[nf:, 0].sum() / w_all))
                                      if len(audit) else float("nan"))
    # and over all candidates, the audit standing in for the unflagged ones
    if len(audit):
        n_unfl = pc.n_cand - nf
        w_aud = max(float(se_arr[nf:, 1].sum()), 1e-30)
        w_unfl = w_aud * n_unfl / len(audit)
        mse_aud = float(se_arr[nf:, 0].sum()) / w_aud
        res["delivered_rmse_all_est_m"] = float(
            math.sqrt(mse_aud * w_unfl / (w_unfl + float(se_arr[:nf, 1].sum()))))
    res["judged_rmse_depth_baseline_m"] = rmse(dep)
    res["judged_rmse_audit_depth_baseline_m"] = rmse(dep[nf:]) if len(audit) else float("nan")
    res["judged_mean_ca_max_depth_m"] = float(np.mean([r[1] for r in ca]))
    # what the two flagged sets look like against all candidates
    for key, rows in (("area", top_a), ("uncertainty", top_u), ("audit", audit),
                      ("candidates", None)):
        sel = cands if rows is None else [cands[i] for i in rows]
        if not len(sel):
            continue
        res[f"{key}_mean_side"] = float(np.mean([c.side for c in sel]))
        res[f"{key}_mean_storm_depth_m"] = float(np.mean([depth(c) for c in sel]))
    # the per-scenario all-CA baseline on a seeded sample of candidates
    rng = np.random.default_rng(pc.seed + 2)
    samp = sorted(rng.choice(pc.n_cand, size=min(pc.n_alone, pc.n_cand), replace=False).tolist())
    t_alone, _ = E.run_alone_crops(dbase, [cands[i] for i in samp], base.dx)
    dbase.close()
    res["alone_s_per_scenario"] = t_alone / len(samp)
    res["all_ca_alone_scenarios_per_h"] = 3600.0 / res["alone_s_per_scenario"]
    res["all_ca_batched_scenarios_per_h"] = pc.n_train / data.t_ca * 3600.0
    t_sur = res["t_screen_s"] + res["t_rerun_s"]
    t_once = res["t_terrain_s"] + res["t_dataset_s"] + res["t_train_s"]
    res["scenarios_per_h"] = pc.n_cand / t_sur * 3600.0
    res["scenarios_per_h_with_training"] = pc.n_cand / (t_sur + t_once) * 3600.0
    res["speedup_vs_alone"] = res["scenarios_per_h"] / res["all_ca_alone_scenarios_per_h"]
    res["speedup_vs_alone_with_training"] = (res["scenarios_per_h_with_training"] /
                                             res["all_ca_alone_scenarios_per_h"])
    res["speedup_vs_batched"] = res["scenarios_per_h"] / res["all_ca_batched_scenarios_per_h"]
    res["train_rmse_by_epoch_m"] = trainer.log.train_rmse
    res["val_rmse_by_epoch_m"] = trainer.log.val_rmse
    log(f"surrogate: judged RMSE {res['judged_rmse_m'] * 100:.2f} cm (depth baseline "
        f"{res['judged_rmse_depth_baseline_m'] * 100:.2f} cm): by area "
        f"{res['judged_rmse_area_m'] * 100:.2f}, by uncertainty "
        f"{res['judged_rmse_uncertainty_m'] * 100:.2f}, audit of unflagged "
        f"{res['judged_rmse_audit_m'] * 100:.2f} cm (95% interval "
        f"{res['judged_rmse_audit_95ci_m'][0] * 100:.2f}-"
        f"{res['judged_rmse_audit_95ci_m'][1] * 100:.2f}; depth baseline "
        f"{res['judged_rmse_audit_depth_baseline_m'] * 100:.2f} cm); delivered (CA maps on "
        f"the flagged, the network's on the rest) {res['judged_delivered_rmse_m'] * 100:.2f} cm "
        f"on the judged, {res.get('delivered_rmse_all_est_m', float('nan')) * 100:.2f} cm "
        f"estimated over all candidates")
    log(f"surrogate: {res['scenarios_per_h']:.0f} "
        f"scenarios/h vs all-CA alone {res['all_ca_alone_scenarios_per_h']:.0f} "
        f"({res['speedup_vs_alone']:.1f}x; {res['speedup_vs_alone_with_training']:.1f}x "
        f"counting dataset and training)")
    res["_objects"] = dict(net=net, data=data, terrain=terrain, screen=sc, cands=cands,
                           top_a=top_a, top_u=top_u, audit=audit)
    return res
