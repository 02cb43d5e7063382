"""The allocation-free scan entry point (flmm_wald_scan_lowrank_noalloc) against flmm_wald_scan_lowrank +
flmm_max_reduce with a caller scratch buffer, through the C ABI: bit-identical outputs and max_F on two calls in a
row, the second with a smaller SNP block, and the size limit reported as an argument error."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest
import torch

from oracle import lmm, perms as operms

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _args(s1, s2, d, mc, P, n_real, nu, row_index, outs):
    s1h, s1l = s1.ptrs
    s2h, s2l = s2.ptrs
    q1h, q1l = d.q1.ptrs
    q2h, q2l = d.q2.ptrs
    head = (s1.prec, s1h, s1l, s2h, s2l, s1.scale_ptr, s2.scale_ptr, s1.rows_pad, q1h, q1l, q2h, q2l,
            d.q1.scale_ptr, d.q2.scale_ptr, d.q1.rows_pad, s1.kpad, d.fac.rpad, mc, P, n_real, d.rWr.data_ptr(),
            d.h2.data_ptr(), float(nu), row_index.data_ptr())
    return head + tuple(t.data_ptr() for t in outs)


@pytest.mark.parametrize("precision", ["split_fp16", "fp32"])
def test_noalloc_scan_matches_scratch_scan(example, precision):
    from fasterlmm_amd import _lib, core, engine
    Zs = lmm.standardise_columns(example["Z_raw"].astype(np.float64))
    chrom, X = example["chrom"], example["X_cov2"]
    Y = example["Y"][:, :8]
    orders = operms.perm_orders(19930909, list(range(8)), Zs.shape[0], 40)
    Yp = np.concatenate([Y, operms.pack_perm_columns(Y, orders)], axis=1)  # 8 + 320 columns: two column tiles
    n_real, P = 8, Yp.shape[1]
    T = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV)  # noqa: E731
    with engine.configure(precision=precision, weights="factored"):
        spec = core.fastlmm_compat_rotate(T(lmm.grm(Zs[:, chrom != 1])), T(X), T(Yp))
        st = spec._gpu
        _, d = st.fit(64, -10.0, 10.0, True)
        assert d.fac is not None
        Zd = T(Zs)
        lib = _lib.load()
        st_ptr = _lib.stream()
        M = Zs.shape[1]
        mf_ref = torch.full((P,), float("-inf"), dtype=torch.float64, device=DEV)
        mf_new = mf_ref.clone()
        for mc in (1400, 300):  # 11 SNP tiles, then 3: the second call leaves stale scratch past its own extent
            print(f"noalloc {precision}: block of {mc} SNPs x {P} columns", flush=True)
            row_index = torch.arange(mc, device=DEV, dtype=torch.int64) + (M - mc) // 2
            s1, s2 = st.snp_planes(Zd, row_index, d)
            ref = [torch.zeros(M, n_real, dtype=torch.float64, device=DEV) for _ in range(5)]
            new = [torch.zeros(M, n_real, dtype=torch.float64, device=DEV) for _ in range(5)]
            n_mt = (mc + 127) // 128
            pmax = torch.empty(n_mt * d.q1.rows_pad, dtype=torch.float64, device=DEV)
            _lib.call("flmm_wald_scan_lowrank", *_args(s1, s2, d, mc, P, n_real, st.nu, row_index, ref),
                      pmax.data_ptr(), st_ptr)
            _lib.call("flmm_max_reduce", pmax.data_ptr(), n_mt, d.q1.rows_pad, P, mf_ref.data_ptr(), st_ptr)
            rc = lib.flmm_wald_scan_lowrank_noalloc(*_args(s1, s2, d, mc, P, n_real, st.nu, row_index, new),
                                                    mf_new.data_ptr(), ctypes.c_void_p(st_ptr))
            assert rc == 0, lib.flmm_last_error()
            torch.cuda.synchronize()
            for a, b in zip(ref, new):
                assert torch.equal(a, b)
            assert torch.equal(mf_ref, mf_new)
            assert torch.isfinite(mf_new).all()
            assert float(ref[0][row_index].abs().sum()) > 0.0
    # the size limit: ceil(mc / 128) x p_pad x 8 bytes must fit the scratch; checked before any device work
    big = 1 + _lib.NOALLOC_SCRATCH_BYTES // (8 * 1024) * 128
    rc = lib.flmm_wald_scan_lowrank_noalloc(2, *([None] * 6), big, *([None] * 6), 1024, 1024, 128, big, 1000, 0,
                                            None, None, 1.0, None, None, None, None, None, None, mf_new.data_ptr(),
                                            None)
    assert rc == 1 and b"scratch" in lib.flmm_last_error()
