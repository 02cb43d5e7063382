"""Learned log-delta proposal: spectral features of a rotated column -> small FP32 MLP -> log delta, finished by
a short Newton refinement and checked against the full grid (which stays as fallback and arbiter).

Everything numeric runs in the gfx950 kernels of libfasterlmm_amd.so: features (flmm_mlp_features), the dense
layers forward and backward (flmm_mlp_dense_fwd / _bwd), the MSE gradient and Adam (flmm_adam), and the checked
refinement (flmm_delta_refine_learned).  Training data are simulated directly in a kinship eigenbasis: for
eigenvalues s and heritability h2, the rotated phenotype has independent entries y_k ~ N(0, sg2 s_k + se2), which
is exactly the LMM's distribution; labels are the grid + Newton optimum of the same likelihood.
"""
from __future__ import annotations

from dataclasses import dataclass
from pathlib import Path

import numpy as np
import torch

from . import _lib, engine
from .engine import TILE, PTILE, KCHUNK, rup

N_IN, H = 34, 128
NBIN = 32
EDGES = np.logspace(-2.0, 3.0, NBIN + 1)  # absolute eigenvalue bins (a kinship Z Z^T / M has mean eigenvalue 1)
SIZES = [(H, N_IN), (H,), (H, H), (H,), (1, H), (1,)]
OFFS = np.cumsum([0] + [int(np.prod(s)) for s in SIZES])
N_PARAMS = int(OFFS[-1])


def bins_of(s: np.ndarray) -> np.ndarray:
    return np.clip(np.searchsorted(EDGES, np.asarray(s), side="right") - 1, 0, NBIN - 1).astype(np.int32)


def features(uy_plane: torch.Tensor, P: int, p_pad: int, s: torch.Tensor, n: int, c: int) -> torch.Tensor:
    """(P, 34) fp32 features of the columns of a fp32 UY plane over the ascending spectrum s."""
    dev = uy_plane.device
    b = torch.from_numpy(bins_of(s.cpu().numpy())).to(dev)
    feat = torch.empty(max(P, 1) * N_IN, dtype=torch.float32, device=dev)
    _lib.call("flmm_mlp_features", uy_plane.data_ptr(), P, p_pad, s.numel(), b.data_ptr(), float(n), float(c),
              feat.data_ptr(), _lib.stream())
    return feat[:P * N_IN].view(P, N_IN)


class DeltaMLP:
    """34 -> 128 -> 128 -> 1 tanh MLP with parameters in one flat fp32 GPU buffer."""

    def __init__(self, seed: int = 0, device=None):
        dev = device or _lib.device()
        rng = np.random.default_rng(seed)
        theta = np.zeros(N_PARAMS, dtype=np.float32)
        for (a, b), shape in zip(zip(OFFS[:-1], OFFS[1:]), SIZES):
            if len(shape) == 2:
                fan_in, fan_out = shape[1], shape[0]
                lim = np.sqrt(6.0 / (fan_in + fan_out))
                theta[a:b] = rng.uniform(-lim, lim, size=b - a)
        self.theta = torch.from_numpy(theta).to(dev)
        self.x_mean = torch.zeros(N_IN, dtype=torch.float32, device=dev)
        self.x_std = torch.ones(N_IN, dtype=torch.float32, device=dev)
        self.y_mean = 0.0
        self.y_std = 1.0

    def part(self, i: int, t: torch.Tensor | None = None) -> torch.Tensor:
        return (self.theta if t is None else t)[int(OFFS[i]):int(OFFS[i + 1])]

    # forward on normalised inputs; returns (pred, activations)
    def _forward(self, x: torch.Tensor):
        B = x.shape[0]
        dev = x.device
        h1 = torch.empty(B, H, dtype=torch.float32, device=dev)
        h2 = torch.empty(B, H, dtype=torch.float32, device=dev)
        out = torch.empty(B, 1, dtype=torch.float32, device=dev)
        st = _lib.stream()
        _lib.call("flmm_mlp_dense_fwd", x.data_ptr(), B, N_IN, self.part(0).data_ptr(), self.part(1).data_ptr(), H, 1,
                  h1.data_ptr(), st)
        _lib.call("flmm_mlp_dense_fwd", h1.data_ptr(), B, H, self.part(2).data_ptr(), self.part(3).data_ptr(), H, 1,
                  h2.data_ptr(), st)
        _lib.call("flmm_mlp_dense_fwd", h2.data_ptr(), B, H, self.part(4).data_ptr(), self.part(5).data_ptr(), 1, 0,
                  out.data_ptr(), st)
        return out[:, 0], (h1, h2)

    def normalise(self, feat: torch.Tensor) -> torch.Tensor:
        return ((feat - self.x_mean) / self.x_std).contiguous()

    def predict(self, feat: torch.Tensor) -> torch.Tensor:
        """Proposed log delta (FP64) for (P, 34) features."""
        if feat.shape[0] == 0:
            return torch.empty(0, dtype=torch.float64, device=feat.device)
        pred, _ = self._forward(self.normalise(feat))
        return pred.double() * self.y_std + self.y_mean

    def grad(self, x: torch.Tensor, target: torch.Tensor) -> tuple[float, torch.Tensor]:
        """MSE loss and its gradient (flat, like theta) on normalised inputs / targets, by the GPU kernels."""
        B = x.shape[0]
        dev = x.device
        pred, (h1, h2) = self._forward(x)
        g = torch.empty(N_PARAMS, dtype=torch.float32, device=dev)
        dpred = torch.empty(B, 1, dtype=torch.float32, device=dev)
        dh2 = torch.empty(B, H, dtype=torch.float32, device=dev)
        dh1 = torch.empty(B, H, dtype=torch.float32, device=dev)
        st = _lib.stream()
        _lib.call("flmm_mlp_mse_grad", pred.data_ptr(), target.data_ptr(), B, dpred.data_ptr(), st)
        out = pred.view(B, 1)
        _lib.call("flmm_mlp_dense_bwd", dpred.data_ptr(), dpred.data_ptr(), out.data_ptr(), h2.data_ptr(), B, H, 1, 0,
                  self.part(4).data_ptr(), dh2.data_ptr(), self.part(4, g).data_ptr(), self.part(5, g).data_ptr(), st)
        _lib.call("flmm_mlp_dense_bwd", dh2.data_ptr(), dh2.data_ptr(), h2.data_ptr(), h1.data_ptr(), B, H, H, 1,
                  self.part(2).data_ptr(), dh1.data_ptr(), self.part(2, g).data_ptr(), self.part(3, g).data_ptr(), st)
        _lib.call("flmm_mlp_dense_bwd", dh1.data_ptr(), dh1.data_ptr(), h1.data_ptr(), x.data_ptr(), B, N_IN, H, 1,
                  self.part(0).data_ptr(), None, self.part(0, g).data_ptr(), self.part(1, g).data_ptr(), st)
        loss = float(((pred - target) ** 2).mean())
        return loss, g

    def fit(self, feat: torch.Tensor, target: torch.Tensor, *, epochs: int = 30, batch: int = 4096,
            lr: float = 2e-3, seed: int = 0, verbose: bool = False) -> list[float]:
        """Adam on the MSE of standardised targets; returns the per-epoch mean loss."""
        dev = feat.device
        self.x_mean = feat.mean(dim=0)
        self.x_std = feat.std(dim=0).clamp_min(1e-3)
        self.y_mean = float(target.double().mean())
        self.y_std = float(target.double().std().clamp_min(1e-6))
        x_all = self.normalise(feat)
        y_all = ((target.double() - self.y_mean) / self.y_std).float().contiguous()
        m = torch.zeros_like(self.theta)
        v = torch.zeros_like(self.theta)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        n = feat.shape[0]
        step = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device=dev)
            tot = 0.0
            nb = 0
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                loss, g = self.grad(x_all[idx].contiguous(), y_all[idx].contiguous())
                step += 1
                lr_t = lr * (0.5 * (1 + np.cos(np.pi * ep / epochs)))
                _lib.call("flmm_adam", self.theta.data_ptr(), g.data_ptr(), m.data_ptr(), v.data_ptr(), N_PARAMS,
                          max(lr_t, 1e-5), 0.9, 0.999, 1e-8, step, _lib.stream())
                tot += loss
                nb += 1
            hist.append(tot / max(nb, 1))
            if verbose:
                print(f"[mlp] epoch {ep} mse {hist[-1]:.4e}", flush=True)
        return hist

    def save(self, path: str | Path) -> None:
        np.savez(path, theta=self.theta.cpu().numpy(), x_mean=self.x_mean.cpu().numpy(),
                 x_std=self.x_std.cpu().numpy(), y_mean=self.y_mean, y_std=self.y_std)

    @classmethod
    def load(cls, path: str | Path) -> "DeltaMLP":
        z = np.load(path)
        net = cls()
        dev = net.theta.device
        net.theta = torch.from_numpy(z["theta"]).to(dev)
        net.x_mean = torch.from_numpy(z["x_mean"]).to(dev)
        net.x_std = torch.from_numpy(z["x_std"]).to(dev)
        net.y_mean = float(z["y_mean"])
        net.y_std = float(z["y_std"])
        return net


@dataclass
class LearnedFit:
    log_delta: torch.Tensor
    fallback: torch.Tensor  # int32 per column, 1 = fell back to the grid bracket
    delta: engine.DeltaResult


def refine_learned(basis, net: DeltaMLP, n_grid: int, lo: float, hi: float, *, n: int, c: int,
                   radius: float = 1.0, judge_tol: float = 1e-6, projected_loss: bool = True,
                   hi_from: int | None = None) -> LearnedFit:
    """Learned delta for every column of a core._Basis: grid (arbiter) + MLP proposal + checked Newton; hi_from:
    the certified screen's first column (engine.screen_start), whose planes get the hi halves only."""
    s_lam, nco = basis._loss(projected_loss)
    idx, lam64 = engine.delta_grid(basis.uy, basis.P, basis.p_pad, basis.kpad, basis.s, s_lam, nco, n_grid, lo, hi,
                                   col_scale=basis.col_scale, s_min=basis.s_range[0])
    feat = features(basis.uy, basis.P, basis.p_pad, basis.s, n, c)
    t0 = net.predict(feat)
    dev = basis.uy.device
    P, p_pad, kpad = basis.P, basis.p_pad, basis.kpad
    log_delta = torch.empty(max(P, 1), dtype=torch.float64, device=dev)
    log_delta[:P].copy_(t0)
    fb = torch.zeros(max(P, 1), dtype=torch.int32, device=dev)
    rWr = torch.empty(max(P, 1), dtype=torch.float64, device=dev)
    h2 = torch.empty(max(P, 1), dtype=torch.float64, device=dev)
    layout = engine.screen_layout(hi_from, P, p_pad, (float(lo), float(hi)), dev)  # the fits stay in [lo, hi]
    la = engine.layout_args(layout, p_pad)
    q1 = engine.Planes(basis.prec, la[0], kpad, dev)
    q2 = engine.Planes(basis.prec, la[0], kpad, dev)
    q1h, q1l = q1.ptrs
    q2h, q2l = q2.ptrs
    _lib.call("flmm_delta_refine_learned", basis.prec, basis.uy.data_ptr(), P, p_pad, kpad, basis.s.data_ptr(),
              basis.s.numel(), s_lam.data_ptr(), s_lam.numel(), float(nco), n_grid, float(lo), float(hi),
              idx.data_ptr(), lam64.data_ptr(), float(radius), float(judge_tol), log_delta.data_ptr(), fb.data_ptr(),
              rWr.data_ptr(),
              h2.data_ptr(), q1h, q1l, q2h, q2l, q1.scale_ptr, q2.scale_ptr, p_pad if hi_from is None else hi_from,
              *la, _lib.stream())
    d = engine.DeltaResult(log_delta[:P], rWr[:P], h2[:P], q1, q2, idx, int(basis.s.numel()), hi_from=hi_from,
                           layout=layout, p_pad=p_pad)
    basis.delta = d
    return LearnedFit(log_delta[:P], fb[:P], d)


# ---------------------------------------------------------------- training data

def kinship_spectra(n: int, *, seed: int = 0, m_values=(1500, 4000, 12000, 40000),
                    fst_values=(0.005, 0.03, 0.1, 0.25), n_pop_values=(2, 6), device=None) -> list[torch.Tensor]:
    """Eigenvalues (ascending, FP64, on the GPU) of intercept-projected Balding-Nichols kinships over a grid of SNP
    counts (M / N from below 1 to 20: rank-deficient to nearly full-rank bulks), drift levels and numbers of
    subpopulations (the spikes), so the spectral features cover the shapes real panels take."""
    from . import sim
    from .core import _projected_eigh, _projector
    from .io import standardise_columns
    dev = device or _lib.device()
    out = []
    k = 0
    for m in m_values:
        for fst in fst_values:
            for n_pop in n_pop_values:
                rng = np.random.default_rng(seed + 1000 * k)
                k += 1
                G = sim.balding_nichols(rng, n, m, n_pop=n_pop, fst=fst, maf=(0.05, 0.5))
                Z = standardise_columns(torch.from_numpy(G).to(dev))
                K = (Z @ Z.T) / m
                X = torch.ones(n, 1, dtype=torch.float64, device=dev)
                _, Iperp = _projector(X)
                s, _ = _projected_eigh(K, Iperp, 1)
                out.append(s.contiguous())
    return out


def simulate_columns(s: torch.Tensor, h2: torch.Tensor, gen: torch.Generator) -> torch.Tensor:
    """(neff, P) rotated phenotypes y_k ~ N(0, sg2 s_k + se2), sg2 = h2 / mean(s), se2 = 1 - h2."""
    sg2 = h2 / s.mean()
    var = sg2.unsqueeze(0) * s.unsqueeze(1) + (1.0 - h2).unsqueeze(0)
    return torch.randn(s.numel(), h2.numel(), generator=gen, device=s.device, dtype=torch.float64) * var.sqrt()


def basis_from_columns(s: torch.Tensor, uy: torch.Tensor, prec: int):
    """A core._Basis-like holder for columns already in the eigenbasis (no rotation needed)."""
    class _B:
        pass
    b = _B()
    neff, P = uy.shape
    b.s = s.contiguous()
    b.neff = neff
    b.kpad = rup(neff, TILE)
    b.P, b.p_pad = P, rup(max(P, 1), PTILE)
    b.prec = prec
    pl = engine.pack(_lib.PREC_FP32, uy.contiguous(), P, neff, 1, P, b.p_pad, b.kpad)
    b.uy = pl.hi
    b.delta = None
    b.col_scale, b.s_range = None, (0.0, 0.0)  # simulated columns: the FP32 grid
    b._loss = lambda projected=True: (b.s, float(neff))
    return b


# ---------------------------------------------------------------- training and evaluation (the learned-delta workload)

def make_training_set(spectra: list[torch.Tensor], n_pheno: int, *, n_grid: int = 100, seed: int = 0,
                      h2_max: float = 0.95, prec: int | None = None):
    """Simulated columns on the given spectra (equal shares), their features and grid + Newton labels."""
    dev = spectra[0].device
    prec = engine.prec_code() if prec is None else prec
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    feats, labels = [], []
    per = -(-n_pheno // len(spectra))
    for s in spectra:
        h2 = torch.rand(per, generator=gen, device=dev, dtype=torch.float64) * h2_max
        for c0 in range(0, per, 65536):
            hh = h2[c0:c0 + 65536]
            uy = simulate_columns(s, hh, gen)
            b = basis_from_columns(s, uy, prec)
            del uy
            idx, lam64 = engine.delta_grid(b.uy, b.P, b.p_pad, b.kpad, s, s, float(b.neff), n_grid, -10.0, 10.0)
            d = engine.delta_refine(prec, _lib.DELTA_NEWTON, b.uy, b.P, b.p_pad, b.kpad, s, s, float(b.neff), n_grid,
                                    -10.0, 10.0, idx, lam64=lam64)
            feats.append(features(b.uy, b.P, b.p_pad, s, b.neff + 1, 1))
            labels.append(d.log_delta.float())
    return torch.cat(feats)[:n_pheno], torch.cat(labels)[:n_pheno]


def default_model_path() -> Path:
    return Path(__file__).resolve().parent / "delta_mlp.npz"


def train_default(*, n: int = 2000, n_pheno: int = 200_000, epochs: int = 30, seed: int = 0,
                  verbose: bool = False) -> tuple[DeltaMLP, dict]:
    """Train the proposal on n_pheno simulated phenotypes over N=n Balding-Nichols kinship spectra."""
    import time
    t0 = time.perf_counter()
    spectra = kinship_spectra(n)
    t1 = time.perf_counter()
    feat, lab = make_training_set(spectra, n_pheno, seed=seed)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    net = DeltaMLP(seed=seed)
    hist = net.fit(feat, lab, epochs=epochs, seed=seed, verbose=verbose)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    info = {"spectra": len(spectra), "n": n, "n_pheno": n_pheno, "epochs": epochs, "mse_first": hist[0],
            "mse_last": hist[-1], "label_std": net.y_std, "s_spectra": t1 - t0, "s_labels": t2 - t1,
            "s_train": t3 - t2}
    return net, info


LOG2PI = float(np.log(2.0 * np.pi))


def evaluate(net: DeltaMLP, *, n: int = 2000, n_pheno: int = 10_000, n_perm: int = 100, m_kin: int = 20_000,
             m_test: int = 1_000, n_grid: int = 100, seed: int = 20250101, verbose: bool = False) -> dict:
    """The learned-delta evaluation: n_pheno held-out phenotypes (h2 ~ U(0, 0.95), a fresh Balding-Nichols panel of
    n strains) x (1 + n_perm) permutation columns on one projected kinship.  Delta is fitted by the plain grid +
    Newton (the grid's exact bracket optimum, every column), by the learned proposal + checked Newton (grid as
    arbiter and fallback, every column), and by the reference's grid + golden section (the real columns).  m_test
    SNPs are scanned against the columns with each.  Reports the fallback rate, the log-likelihood gap of the
    learned optimum, |delta log10 p| on the real columns' per-test p and on every column's genome-wide max F, and
    the delta-search times."""
    import sys
    import time
    from . import sim
    from .core import _Basis, _new_outputs, _projected_eigh, _projector
    from .io import standardise_columns
    from .pvalues import neglog10p

    def say(msg):
        if verbose:
            print(f"[learned-eval] {msg}", file=sys.stderr, flush=True)

    dev = _lib.device()
    rng = np.random.default_rng(seed)
    say(f"panel N={n}, {m_kin} kinship + {m_test} test SNPs; {n_pheno} phenotypes x (1+{n_perm})")
    G = sim.balding_nichols(rng, n, m_kin + m_test, fst=0.05, maf=(0.05, 0.5))
    Z = standardise_columns(torch.from_numpy(G).to(dev))
    del G
    Zk, Zt = Z[:, :m_kin], Z[:, m_kin:].contiguous()
    K = (Zk @ Zk.T) / m_kin
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    lam, V = torch.linalg.eigh(K)
    h2 = torch.rand(n_pheno, generator=gen, device=dev, dtype=torch.float64) * 0.95
    g = V @ (lam.clamp_min(0).sqrt().unsqueeze(1) * torch.randn(n, n_pheno, generator=gen, device=dev,
                                                                  dtype=torch.float64))
    g = (g - g.mean(0)) / g.std(0)
    e = torch.randn(n, n_pheno, generator=gen, device=dev, dtype=torch.float64)
    y = h2.sqrt() * g + (1.0 - h2).sqrt() * e
    del g, e, V, lam
    P = n_pheno * (1 + n_perm)
    Y = torch.empty(n, P, dtype=torch.float64, device=dev)
    Y[:, :n_pheno] = y
    for b0 in range(0, n_pheno, 256):  # pheno-major permutation columns, as perm_threshold lays them out
        b1 = min(b0 + 256, n_pheno)
        orders = torch.rand(b1 - b0, n_perm, n, generator=gen, device=dev).argsort(dim=2)
        yp = torch.gather(y[:, b0:b1].T.unsqueeze(1).expand(b1 - b0, n_perm, n), 2, orders)
        Y[:, n_pheno + b0 * n_perm:n_pheno + b1 * n_perm] = yp.reshape((b1 - b0) * n_perm, n).T
        del orders, yp
    X = torch.ones(n, 1, dtype=torch.float64, device=dev)
    Xpinv, Iperp = _projector(X)
    s, U = _projected_eigh(K, Iperp, 1)
    prec = engine.prec_code()
    Y -= X @ (Xpinv @ Y)
    st = _Basis(s, U, 1, n - 1, prec, nu=n - 2)
    st.rotate_columns(Y, in_span=True)
    st_real = _Basis(s, U, 1, n - 1, prec, nu=n - 2)
    st_real.rotate_columns(Y[:, :n_pheno].contiguous(), in_span=True)
    del Y
    neff = n - 1
    say("rotated; fitting delta")

    def timed(fn):
        torch.cuda.synchronize()
        t = time.perf_counter()
        r = fn()
        torch.cuda.synchronize()
        return r, time.perf_counter() - t

    def fit_mode(basis, mode):
        with engine.configure(delta=mode):
            return basis.fit(n_grid, -10.0, 10.0, True)

    (ld_n, d_n), t_newton = timed(lambda: fit_mode(st, "newton"))
    fit, t_learned = timed(lambda: refine_learned(st, net, n_grid, -10.0, 10.0, n=n, c=1))
    ld_l, d_l, fb = fit.log_delta, fit.delta, fit.fallback
    (ld_g, d_g), t_golden = timed(lambda: fit_mode(st_real, "golden"))

    def loss(ld, rWr):
        dl = ld.exp()
        logdet = torch.empty_like(ld)
        for c0 in range(0, ld.numel(), 65536):
            logdet[c0:c0 + 65536] = torch.log(s.unsqueeze(1) + dl[c0:c0 + 65536].unsqueeze(0)).sum(0)
        return neff * (LOG2PI + torch.log(rWr / neff) + 1.0) + logdet

    say(f"delta: grid+Newton {t_newton:.2f}s, learned {t_learned:.2f}s, golden (real columns) {t_golden:.2f}s; "
        f"scanning")
    L_n, L_l, L_g = loss(ld_n, d_n.rWr), loss(ld_l, d_l.rWr), loss(ld_g, d_g.rWr)
    gap = L_l - L_n
    rows = torch.arange(m_test, device=dev)
    df = n - 2
    lp = {}
    for tag, basis, d, p_cols in (("newton", st, d_n, P), ("learned", st, d_l, P), ("golden", st_real, d_g, n_pheno)):
        out = _new_outputs(m_test, n_pheno, p_cols, dev)
        basis.scan(Zt, rows, n_pheno, out, d)
        lp[tag] = (neglog10p(out.f, df), neglog10p(out.max_F, df))
        del out
    fbf = fb.double()
    acc = fb == 0
    t0 = net.predict(features(st.uy, st.P, st.p_pad, s, n, 1))
    err = (t0 - ld_n).abs().float()
    q = torch.tensor([0.5, 0.9, 0.99], device=dev)
    diag = {"proposal_abs_err_real_q50_q90_q99": torch.quantile(err[:n_pheno], q).tolist(),
            "proposal_abs_err_perm_q50_q90_q99": torch.quantile(err[n_pheno:][:1 << 24], q).tolist(),
            "grid_newton_log_delta_real_q10_q50_q90": torch.quantile(ld_n[:n_pheno].float(),
                                                                    torch.tensor([0.1, 0.5, 0.9], device=dev)).tolist(),
            "grid_newton_at_upper_edge_frac": float((ld_n >= 10.0 - 1e-9).double().mean())}
    for r in (0.5, 2.0, 4.0):
        fr = refine_learned(st, net, n_grid, -10.0, 10.0, n=n, c=1, radius=r).fallback.double()
        diag[f"fallback_rate_radius_{r:g}"] = float(fr.mean())
        diag[f"fallback_rate_real_radius_{r:g}"] = float(fr[:n_pheno].mean())
    info = {**diag,
        "n": n, "n_pheno": n_pheno, "n_perm": n_perm, "columns": P, "m_test": m_test, "n_grid": n_grid,
        "fallback_rate": float(fbf.mean()), "fallback_rate_real": float(fbf[:n_pheno].mean()),
        "ll_gap_learned_minus_grid_newton_max": float(gap.max()),
        "ll_gap_learned_minus_grid_newton_min": float(gap.min()),
        "ll_gap_learned_minus_golden_real_max": float((L_l[:n_pheno] - L_g).max()),
        "ll_gap_grid_newton_minus_golden_real_max": float((L_n[:n_pheno] - L_g).max()),
        "max_abs_dlog10p_real_learned_vs_grid_newton": float((lp["learned"][0] - lp["newton"][0]).abs().max()),
        "max_abs_dlog10p_maxF_learned_vs_grid_newton": float((lp["learned"][1] - lp["newton"][1]).abs().max()),
        "max_abs_dlog10p_real_learned_vs_golden": float((lp["learned"][0] - lp["golden"][0]).abs().max()),
        "max_abs_dlog10p_maxF_real_learned_vs_golden": float(
            (lp["learned"][1][:n_pheno] - lp["golden"][1]).abs().max()),
        "max_abs_dlogdelta_accepted_vs_grid_newton": float((ld_l - ld_n).abs()[acc].max()) if bool(acc.any()) else None,
        "seconds_delta_grid_newton": t_newton, "seconds_delta_learned": t_learned,
        "seconds_delta_golden_real_columns_only": t_golden,
    }
    if verbose:
        print(info, flush=True)
    return info


_MODEL: DeltaMLP | None = None


def model() -> DeltaMLP:
    """The proposal network used by engine.settings.delta == "learned" (loaded once from delta_mlp.npz, written by
    ``python -m fasterlmm_amd.learned`` on a GPU)."""
    global _MODEL
    if _MODEL is None:
        p = default_model_path()
        if not p.exists():
            raise RuntimeError(f"{p} is missing; train it with `python -m fasterlmm_amd.learned` on the GPU")
        _MODEL = DeltaMLP.load(p)
    return _MODEL


def set_model(net: DeltaMLP | None) -> None:
    global _MODEL
    _MODEL = net


if __name__ == "__main__":
    import argparse
    import json
    ap = argparse.ArgumentParser(description="train (default) or evaluate the learned delta proposal on the GPU")
    ap.add_argument("--eval", action="store_true", help="evaluate the saved model instead of training")
    ap.add_argument("--model", default=str(default_model_path()))
    ap.add_argument("--n-pheno", type=int, default=None)
    ap.add_argument("--n-perm", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=30)
    a = ap.parse_args()
    if a.eval:
        info = evaluate(DeltaMLP.load(a.model), n_pheno=a.n_pheno or 10_000, n_perm=a.n_perm, verbose=True)
    else:
        net, info = train_default(n_pheno=a.n_pheno or 200_000, epochs=a.epochs, verbose=True)
        net.save(a.model)
    print(json.dumps(info), flush=True)
