"""Grouped Wald scans: many (SNP block x column set) scans in one persistent launch, captured as a HIP graph.

The reference scans every missingness group of phenotypes on its own (fasterlmm/cli.py:337-375: one
``perm_threshold`` per NA pattern, each with its own LOCO loop, fasterlmm/perms.py:62).  Here the scans of all
groups for one SNP block run in ONE launch of ``flmm_wald_scan_grouped``: the groups' operand planes, sizes and
outputs are listed in a device table, and a persistent grid walks the tiles of every group.  Groups differ in
retained N (so in K = Neff), in column count and in output arrays; nothing in the launch depends on them except
through the table, so the launch (plus the per-group maximum fold ``flmm_max_reduce_grouped``) is captured once
as a HIP graph and replayed for every SNP block after the table is rewritten in place.

The tile variant (128 x 128 square tiles or 128 x 256 wide tiles) and the persistent grid size are chosen by
``ScanTuner``: each candidate is timed with HIP events on the first table of a shape and the fastest is kept
(optionally persisted as JSON).
"""
from __future__ import annotations

import ctypes as C
import functools
import json
import os
from dataclasses import dataclass
from typing import Callable

import numpy as np
import torch
from torch import Tensor

from . import _lib, engine
from .engine import TILE, Planes, DeltaResult, ScanOutputs

W_TP = 256  # column tile of the wide variant


@dataclass
class ScanJob:
    """One scan of a grouped launch: SNP planes s1 / s2 (mc rows) against the column planes of d."""
    s1: Planes
    s2: Planes
    d: DeltaResult
    mc: int
    P: int
    n_real: int
    nu: float
    row_index: Tensor  # (mc,) int64: output row of each SNP
    out: ScanOutputs
    pmax: Tensor | None = None  # per-tile maxima scratch, allocated by GroupedScan
    finish: Callable[[], None] | None = None  # run after the scan (the FP64 re-scoring of large-F real tests)

    @property
    def factored(self) -> bool:
        return self.d.fac is not None

    def tiles(self, variant: int) -> int:
        tp = TILE if variant == _lib.SCAN_SQUARE else W_TP
        return -(-self.mc // TILE) * -(-self.P // tp)


def _factored(jobs: list[ScanJob]) -> bool:
    """Whether the jobs use factored weights (all of them) or exact ones (none); a mix cannot share a launch."""
    f = {j.factored for j in jobs}
    if len(f) > 1:
        raise ValueError("a grouped launch needs every job with factored weights or none")
    return f == {True}


def _descriptor(j: ScanJob) -> _lib.ScanGroup:
    g = _lib.ScanGroup()
    g.s1_hi, g.s1_lo = j.s1.ptrs
    g.s2_hi, g.s2_lo = j.s2.ptrs
    g.s1_scale, g.s2_scale = j.s1.scale_ptr, j.s2.scale_ptr
    g.mc_pad = j.s1.rows_pad
    g.q1_hi, g.q1_lo = j.d.q1.ptrs
    g.q2_hi, g.q2_lo = j.d.q2.ptrs
    g.q1_scale, g.q2_scale = j.d.q1.scale_ptr, j.d.q2.scale_ptr
    g.p_pad, g.kpad, g.mc, g.P, g.n_real = j.d.q1.rows_pad, j.s1.kpad, j.mc, j.P, j.n_real
    g.kpad2 = j.d.fac.rpad if j.factored else 0
    g.rWr, g.h2, g.nu = j.d.rWr.data_ptr(), j.d.h2.data_ptr(), float(j.nu)
    g.row_index = j.row_index.data_ptr()
    if j.n_real > 0:
        o = j.out
        g.f, g.beta, g.se, g.sfve, g.nullh2 = (t.data_ptr() for t in (o.f, o.beta, o.se, o.sfve, o.nullh2))
    g.pmax = j.pmax.data_ptr()
    g.max_F = j.out.max_F.data_ptr()
    return g


class ScanTuner:
    """Project-written autotuner of the grouped scan: (tile variant, persistent grid size) per table shape.

    The shape key is (precision, group count, the tile counts of both variants, the largest K, factored or exact
    weights).  Candidates are the square and wide variants (factored weights: the factored wide variant only) on
    1, 2 or 4 workgroups per CU (256 CUs) and one workgroup per tile; each is run once to warm up and ``reps``
    times under HIP events on the launch stream, and the fastest mean wins."""

    CUS = 256

    def __init__(self, path: str | None = None, reps: int = 2):
        self.path = path if path is not None else os.environ.get("FLMM_SCAN_TUNE_CACHE")
        self.reps = reps
        self.cache: dict[str, tuple[int, int]] = {}
        self.trials: dict[str, list] = {}
        if self.path and os.path.exists(self.path):
            with open(self.path) as fh:
                self.cache = {k: tuple(v) for k, v in json.load(fh).items()}

    @staticmethod
    def key(prec: int, jobs: list[ScanJob]) -> str:
        kmax = max(j.s1.kpad for j in jobs)
        return f"p{prec}-g{len(jobs)}-sq{sum(j.tiles(_lib.SCAN_SQUARE) for j in jobs)}-" \
               f"w{sum(j.tiles(_lib.SCAN_WIDE) for j in jobs)}-k{kmax}-f{int(_factored(jobs))}"

    def candidates(self, jobs: list[ScanJob]) -> list[tuple[int, int]]:
        out = []
        variants = (_lib.SCAN_WIDE_FACTORED,) if _factored(jobs) else (_lib.SCAN_WIDE, _lib.SCAN_SQUARE)
        for variant in variants:
            total = sum(j.tiles(variant) for j in jobs)
            per_tile = engine.rup(max(total, 8), 8)
            for blocks in sorted({min(self.CUS * f, per_tile) for f in (1, 2, 4)} | {per_tile}):
                out.append((variant, blocks))
        return out

    def choose(self, runner: "GroupedScan", jobs: list[ScanJob]) -> tuple[int, int]:
        k = self.key(runner.prec, jobs)
        if k in self.cache:
            return self.cache[k]
        st = torch.cuda.current_stream()
        trials = []
        for variant, blocks in self.candidates(jobs):
            table = runner._write_table(variant, jobs)
            runner._launch_scan(variant, blocks, table)  # warm-up (and module load)
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(st)
            for _ in range(self.reps):
                runner._launch_scan(variant, blocks, table)
            b.record(st)
            b.synchronize()
            trials.append((a.elapsed_time(b) / self.reps, variant, blocks))
        best = min(trials)
        self.trials[k] = trials
        self.cache[k] = (best[1], best[2])
        if self.path:
            os.makedirs(os.path.dirname(os.path.abspath(self.path)), exist_ok=True)
            with open(self.path, "w") as fh:
                json.dump({kk: list(v) for kk, v in self.cache.items()}, fh, indent=1)
        return self.cache[k]


@dataclass
class _Graph:
    graph: torch.cuda.CUDAGraph
    variant: int
    blocks: int
    n_cap: int
    p_cap: int


class GroupedScan:
    """Runs lists of ScanJobs (one SNP block of several column sets) as one grouped launch.

    ``run(jobs)`` writes the table (pinned host buffer -> device buffer of fixed address), asks the tuner for
    the variant and grid size, and launches the persistent scan and the per-group maximum fold -- by replaying a
    captured HIP graph when ``graph`` is set (captured on first use for each variant / grid size, with room for
    ``max_groups`` groups of at most ``max_P`` columns) or eagerly otherwise.  Jobs with factored weights and
    jobs with exact ones (a group whose Neff <= 128 keeps the exact weights) run as two such launches, each with
    its own table buffer.  Each job's pmax scratch is allocated here.  The outputs of a job are complete once the
    launch stream reaches the end of ``run``."""

    def __init__(self, prec: int, max_groups: int, max_P: int, *, graph: bool = True, tuner: ScanTuner | None = None,
                 variant: int | None = None, blocks: int | None = None):
        if max_groups < 1 or max_groups > 65535:
            raise ValueError(f"max_groups must be in [1, 65535], got {max_groups}")
        self.dev = _lib.device()
        self.prec = prec
        self.max_groups = int(max_groups)
        self.max_P = int(max_P)
        self.use_graph = graph
        self.tuner = tuner if tuner is not None else ScanTuner()
        self.fixed = (variant, blocks) if variant is not None else None
        if self.fixed is not None and (blocks is None or blocks <= 0 or blocks % 8):
            raise ValueError("a fixed grid size must be a positive multiple of 8")
        self._nbytes = int(_lib.load().flmm_scan_table_bytes(self.max_groups))
        self._bufs: dict[bool, list] = {}  # factored? -> [pinned host table, device table, copy event]
        self._graphs: dict[tuple[int, int], _Graph] = {}
        self.last_choice: tuple[int, int] | None = None
        self.last_choices: list[tuple[int, int]] = []

    # table ---------------------------------------------------------------------------------------------------
    def _buf(self, variant: int) -> list:
        kind = variant == _lib.SCAN_WIDE_FACTORED
        b = self._bufs.get(kind)
        if b is None:
            b = [torch.empty(self._nbytes, dtype=torch.uint8, pin_memory=True),
                 torch.empty(self._nbytes, dtype=torch.uint8, device=self.dev), None]
            self._bufs[kind] = b
        return b

    def _write_table(self, variant: int, jobs: list[ScanJob]) -> Tensor:
        if len(jobs) > self.max_groups:
            raise ValueError(f"{len(jobs)} groups exceed the runner's max_groups={self.max_groups}")
        b = self._buf(variant)
        host, table, copied = b
        if copied is not None:
            copied.synchronize()  # the previous copy has read the pinned buffer
        arr = (_lib.ScanGroup * len(jobs))(*(_descriptor(j) for j in jobs))
        _lib.call("flmm_scan_table_fill", self.prec, variant, C.addressof(arr) if jobs else None, len(jobs),
                  host.data_ptr())
        table.copy_(host, non_blocking=True)
        b[2] = torch.cuda.Event()
        b[2].record()
        return table

    def _launch_scan(self, variant: int, blocks: int, table: Tensor) -> None:
        _lib.call("flmm_wald_scan_grouped", self.prec, variant, table.data_ptr(), int(blocks), _lib.stream())

    def _launch_reduce(self, table: Tensor, n_groups: int, max_P: int) -> None:
        _lib.call("flmm_max_reduce_grouped", table.data_ptr(), int(n_groups), int(max_P), _lib.stream())

    def _graph_for(self, variant: int, blocks: int) -> _Graph:
        key = (variant, blocks)
        g = self._graphs.get(key)
        if g is None:
            table = self._buf(variant)[1]
            # eager launches first (code-object load happens outside the capture); both are idempotent on the
            # table already written: the scan rewrites the same outputs and the fold is a maximum
            self._launch_scan(variant, blocks, table)
            self._launch_reduce(table, self.max_groups, self.max_P)
            graph = torch.cuda.CUDAGraph()
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                with torch.cuda.graph(graph, stream=side):
                    self._launch_scan(variant, blocks, table)
                    self._launch_reduce(table, self.max_groups, self.max_P)
            torch.cuda.current_stream().wait_stream(side)
            g = _Graph(graph, variant, blocks, self.max_groups, self.max_P)
            self._graphs[key] = g
        return g

    # run -----------------------------------------------------------------------------------------------------
    def prepare(self, jobs: list[ScanJob]) -> None:
        for j in jobs:
            if j.s1.prec != self.prec or j.d.q1.prec != self.prec:
                raise ValueError("every job of a grouped launch must use the runner's precision")
            if j.P > self.max_P:
                raise ValueError(f"a job has P={j.P} columns, above the runner's max_P={self.max_P}")
            n_mt = -(-j.mc // TILE)
            need = max(n_mt * j.d.q1.rows_pad, 1)
            if j.pmax is None or j.pmax.numel() < need:
                j.pmax = torch.empty(need, dtype=torch.float64, device=self.dev)

    def choose(self, jobs: list[ScanJob]) -> tuple[int, int]:
        if self.fixed is not None:
            variant, blocks = self.fixed
            if _factored(jobs):  # the wide tile with its second product over kpad2
                if variant == _lib.SCAN_SQUARE:
                    raise ValueError("factored weights run on the wide tile only")
                variant = _lib.SCAN_WIDE_FACTORED
            return variant, blocks
        return self.tuner.choose(self, jobs)

    def run(self, jobs: list[ScanJob]) -> None:
        jobs = [j for j in jobs if j.mc > 0 and j.P > 0]
        if not jobs:
            return
        self.prepare(jobs)
        self.last_choices = []
        for part in ([j for j in jobs if j.factored], [j for j in jobs if not j.factored]):
            if not part:
                continue
            variant, blocks = self.choose(part)
            self.last_choices.append((variant, blocks))
            table = self._write_table(variant, part)
            if self.use_graph:
                self._graph_for(variant, blocks).graph.replay()
            else:
                self._launch_scan(variant, blocks, table)
                self._launch_reduce(table, len(part), max(j.P for j in part))
        self.last_choice = self.last_choices[0]


def per_group_loop(jobs: list[ScanJob]) -> None:
    """The same scans as one flmm_wald_scan + flmm_max_reduce launch pair per job (the per-group loop the
    grouped launch replaces)."""
    for j in jobs:
        engine.wald_scan(j.s1, j.s2, j.d, j.mc, j.P, j.n_real, j.nu, j.row_index, j.out, pmax=j.pmax)


# GROUPED LOCO DRIVER -------

class GroupedLoco:
    """Leave-one-chromosome-out compat scans of several strain subsets (the NA groups of fasterlmm/cli.py:337-375),
    prepared so that the scans of every group can run together.

    Group g keeps the rows ``keeps[g]`` of the raw genotypes Z (N, M; NaN = missing) and covariates X, and scans
    the columns of Ys[g] (N_g, P_g) -- reals first, n_reals[g] of them.  Per group the result equals
    ``core.loco_scan_compat(standardise_columns(Z[keep]), X[keep], Ys[g], chrom, n_real=n_reals[g])`` (the
    reference's per-group ``perm_threshold`` on ``subset_dataset``, perms.py:40-62): genotypes are standardised
    over the group's rows, K_loco comes from the group's standardised panel (or from Z_kin's rows), and every
    group has its own eigendecomposition, delta fit and dof N_g - C - 1.

    Construction computes each group's full kinship (one chromosome's columns at a time, so no group holds a
    standardised copy of Z), covariate projector and packed scan columns.  ``chromosome_jobs(c)`` then builds every
    group's basis, delta fit and SNP planes for chromosome c and yields one list of ScanJobs per block of up to
    ``snp_block`` SNPs; the caller runs each list (GroupedScan.run or per_group_loop) before asking for the next."""

    def __init__(self, Z: Tensor, X: Tensor, Ys: list, keeps: list, chrom, *, n_reals: list[int] | None = None,
                 Z_kin: Tensor | None = None, chrom_kin=None, n_grid: int = 64, snp_block: int = 16384,
                 progress=None, eigh_batch: int = 8, loco: bool = True):
        from . import core
        self.loco = bool(loco)  # False: one kinship of the whole panel per group, every SNP scanned against it
        self.progress = progress  # progress(chromosome, k) after the k-th basis of a chromosome, if given
        self.eigh_batch = max(1, int(eigh_batch))
        if (Z_kin is None) != (chrom_kin is None):
            raise ValueError("Z_kin and chrom_kin go together")
        if len(Ys) != len(keeps):
            raise ValueError("one column matrix per group is needed")
        self.home = Z.device if isinstance(Z, Tensor) else torch.device("cpu")
        dev = self.dev = _lib.device()
        self.Z = torch.as_tensor(Z).to(device=dev, dtype=torch.float64)
        Zkd = None if Z_kin is None else torch.as_tensor(Z_kin).to(device=dev, dtype=torch.float64)
        X = torch.as_tensor(X).to(device=dev, dtype=torch.float64)
        self.chrom = np.asarray(chrom)
        ck = self.chrom if chrom_kin is None else np.asarray(chrom_kin)
        self.Zk = self.Z if Zkd is None else Zkd
        self.Mk = self.Zk.shape[1]
        self.D = D = X.shape[1]
        self.prec = prec = engine.prec_code()
        self.n_grid, self.snp_block = int(n_grid), int(snp_block)
        self.G = G = len(Ys)
        self.rows = [torch.as_tensor(np.asarray(k, dtype=np.int64)).to(dev) for k in keeps]
        self.Ns = [int(r.numel()) for r in self.rows]
        for g, n in enumerate(self.Ns):
            if n - D - 1 < 1:
                raise ValueError(f"group {g} keeps {n} rows, too few for {D} covariates")
        Yg = [torch.as_tensor(y).to(device=dev, dtype=torch.float64) for y in Ys]
        self.Ps = [int(y.shape[1]) for y in Yg]
        self.n_reals = [P if n_reals is None else min(int(n_reals[g]), P) for g, P in enumerate(self.Ps)]
        self.outs = [core._new_outputs(self.Z.shape[1], self.n_reals[g], self.Ps[g], dev) for g in range(G)]
        self.chroms = sorted(np.unique(self.chrom).tolist())
        self.kin_cols = {c: torch.from_numpy(np.flatnonzero(ck == c)).to(dev) for c in sorted(np.unique(ck).tolist())}
        # per group: K_all of its standardised panel, projector, packed columns
        self.K_all = [torch.zeros(n, n, dtype=torch.float64, device=dev) for n in self.Ns]
        with engine.stage("kinship"):
            for c, kc in self.kin_cols.items():
                for g in range(G):
                    zs = self._std_rows(self.Zk, kc, g)
                    self.K_all[g] += engine.gram(zs, prec)
                    del zs
        self.Iperps, self.ypls, self.y_reals = [], [], []
        with engine.stage("pack_columns"):
            for g in range(G):
                Xg = X.index_select(0, self.rows[g])
                Xpinv, Iperp = core._projector(Xg)
                self.Iperps.append(Iperp)
                Y_r = (Yg[g] - Xg @ (Xpinv @ Yg[g])).contiguous()
                self.y_reals.append(Y_r[:, : self.n_reals[g]].contiguous())
                self.ypls.append(engine.pack(prec, Y_r, self.Ps[g], self.Ns[g], 1, self.Ps[g],
                                             engine.rup(max(self.Ps[g], 1), engine.PTILE),
                                             engine.rup(self.Ns[g], engine.KCHUNK)))
                del Y_r

    def _std_rows(self, mat: Tensor, cols: Tensor, g: int) -> Tensor:
        from .io import standardise_columns
        return standardise_columns(mat.index_select(1, cols).index_select(0, self.rows[g]))

    def _projected_kernel(self, c, g: int) -> Tensor:
        """Group g's projected kinship (I - X X^+)(K_loco + I)(I - X X^+) for chromosome c (core.py:503-510)."""
        N, dev = self.Ns[g], self.dev
        kc = self.kin_cols.get(c) if self.loco else None
        eye = torch.eye(N, dtype=torch.float64, device=dev)
        if kc is not None and kc.numel():
            zs = self._std_rows(self.Zk, kc, g)
            K = (self.K_all[g] - engine.gram(zs, self.prec)) / float(self.Mk - kc.numel())
            del zs
        else:
            K = self.K_all[g] / float(self.Mk)
        return (self.Iperps[g] @ (K + eye)) @ self.Iperps[g]

    def _bases(self, c) -> list:
        """Every group's basis, rotated columns and delta fit for chromosome c.  The FP64 eigendecompositions run
        batched: groups sorted by retained N, `eigh_batch` at a time, each projected kernel zero-extended to the
        batch's largest N with -1 on the padding diagonal (a decoupled block whose eigenvalues sort before the
        kernel's D zeros and are dropped with them); one batched rocSOLVER call runs several times faster than the
        same eighs one by one, whose tridiagonalisations are chains of tiny launches."""
        from . import core
        D, dev = self.D, self.dev
        out = [None] * self.G
        order = sorted(range(self.G), key=lambda g: self.Ns[g])
        for b0 in range(0, self.G, self.eigh_batch):
            gs = order[b0:b0 + self.eigh_batch]
            nb = max(self.Ns[g] for g in gs)
            with engine.stage("kinship"):
                stack = -torch.eye(nb, dtype=torch.float64, device=dev).repeat(len(gs), 1, 1)
                for i, g in enumerate(gs):
                    N = self.Ns[g]
                    stack[i, :N, :N] = self._projected_kernel(c, g)
                    stack[i, :N, N:] = 0.0
                    stack[i, N:, :N] = 0.0
            with engine.stage("eigh"):
                ev, U = core._eigh(stack)
                del stack
            for i, g in enumerate(gs):
                N = self.Ns[g]
                lo = nb - N + D  # padding eigenvalues (-1) first, then the D zeros of the projection
                s = torch.clamp(ev[i, lo:] - 1.0, min=0.0)
                with engine.stage("pack_basis"):
                    st = core._Basis(s, U[i, :N, :], lo, N - D, self.prec, nu=N - D - 1)
                with engine.stage("rotate_columns"):
                    st.use_column_planes(self.ypls[g], self.Ps[g], self.y_reals[g], in_span=True)
                with engine.stage("delta"):
                    _, d = st.fit(self.n_grid, -10.0, 10.0, True, projected_loss=True)
                out[g] = (st, d)
                if self.progress is not None:
                    self.progress(c, b0 + i)
            del ev, U
        return out

    def chromosome_jobs(self, c):
        """Jobs for the SNPs of chromosome c (LOCO), or for every SNP when the scan is not LOCO (c is ignored)."""
        cols = np.flatnonzero(self.chrom == c) if self.loco else np.arange(self.chrom.shape[0])
        tc = torch.from_numpy(cols).to(self.dev)
        bases = self._bases(c)
        for b0 in range(0, tc.numel(), self.snp_block):
            ri = tc[b0:b0 + self.snp_block]
            local = torch.arange(ri.numel(), device=self.dev)
            jobs = []
            with engine.stage("snps"):
                for g, (st, d) in enumerate(bases):
                    zs = self._std_rows(self.Z, ri, g)
                    s1, s2 = st.snp_planes(zs, local, d)
                    del zs
                    jobs.append(ScanJob(s1, s2, d, ri.numel(), self.Ps[g], self.n_reals[g], st.nu, ri, self.outs[g],
                                        finish=functools.partial(self._rescore, st, d, ri, local, g)))
            yield jobs
            del jobs

    def _rescore(self, st, d, ri: Tensor, local: Tensor, g: int) -> None:
        """FP64 re-scoring of the block's large-F real tests of group g (core._Basis.rescore); only the selected
        SNPs are standardised again."""
        st.rescore(lambda idx: self._std_rows(self.Z, ri.index_select(0, idx), g), local, ri, self.n_reals[g],
                   self.outs[g], d)

    def results(self) -> list:
        from . import core
        for o, n in zip(self.outs, self.n_reals):
            engine.refresh_real_max(o, n)
        res = [core.ScanResult(o.f, o.beta, o.se, o.sfve, o.nullh2, o.max_F) for o in self.outs]
        return [core._home(r, self.home) for r in res]


def loco_scan_compat_groups(Z: Tensor, X: Tensor, Ys: list, keeps: list, chrom, *, n_reals: list[int] | None = None,
                            Z_kin: Tensor | None = None, chrom_kin=None, n_grid: int = 64, mode: str = "grouped",
                            runner: GroupedScan | None = None, snp_block: int = 16384, test_chroms=None,
                            on_chrom=None, loco: bool = True) -> list:
    """LOCO compat scans of several strain subsets (see GroupedLoco for the per-group contract), chromosome by
    chromosome, with the scans of all groups for each SNP block run as one grouped launch (mode="grouped",
    through ``runner``: a GroupedScan, by default graph-captured and autotuned) or as one launch pair per group
    (mode="loop").  ``test_chroms`` limits the scanned chromosomes (each still leaves itself out of K); rows of
    other chromosomes keep zeros and do not enter max_F.  Returns one ScanResult per group, on the device of Z."""
    if mode not in ("grouped", "loop"):
        raise ValueError(f"mode must be 'grouped' or 'loop', got {mode!r}")
    gl = GroupedLoco(Z, X, Ys, keeps, chrom, n_reals=n_reals, Z_kin=Z_kin, chrom_kin=chrom_kin, n_grid=n_grid,
                     snp_block=snp_block, loco=loco)
    if mode == "grouped" and runner is None:
        runner = GroupedScan(gl.prec, gl.G, max(gl.Ps))
    chroms = gl.chroms if test_chroms is None else [c for c in gl.chroms if c in set(test_chroms)]
    if not loco:
        chroms = chroms[:1]  # one pass: the single kinship against every SNP
    for ci, c in enumerate(chroms):
        for jobs in gl.chromosome_jobs(c):
            with engine.stage("scan"):
                if mode == "grouped":
                    runner.run(jobs)
                else:
                    per_group_loop(jobs)
            with engine.stage("rescore"):
                for j in jobs:
                    if j.finish is not None:
                        j.finish()
        if on_chrom is not None:
            on_chrom(ci + 1, len(chroms))
    return gl.results()
