"""The FP64 eigendecomposition retries on the host when the device runs out of memory and lets every other error
through (the reference's eigh fallback, fasterlmm/core.py:54-77; its test tests/test_core_fallback.py:25-55).
Host-only: the GPU call is replaced by a stub that raises."""
from __future__ import annotations

import pytest
import torch


def test_out_of_memory_retries_on_host(monkeypatch):
    from fasterlmm_amd import core
    real = torch.linalg.eigh
    calls = []

    def flaky(K):
        calls.append(K.device.type)
        if len(calls) == 1:
            raise RuntimeError("HIP out of memory. Tried to allocate 2.00 GiB")
        return real(K)

    monkeypatch.setattr(torch.linalg, "eigh", flaky)
    A = torch.randn(20, 20, dtype=torch.float64)
    K = A @ A.T
    s, U = core._eigh(K)
    assert len(calls) == 2
    ws, wU = real(K)
    torch.testing.assert_close(s, ws)
    torch.testing.assert_close(U @ torch.diag(s) @ U.T, K)


def test_other_errors_propagate(monkeypatch):
    from fasterlmm_amd import core

    def broken(K):
        raise RuntimeError("rocsolver: invalid argument")

    monkeypatch.setattr(torch.linalg, "eigh", broken)
    with pytest.raises(RuntimeError, match="invalid argument"):
        core._eigh(torch.eye(4, dtype=torch.float64))
