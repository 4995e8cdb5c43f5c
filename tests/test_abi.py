"""C ABI checks that need no GPU: the library loads, exports every symbol that
``include/aknn.h`` declares, and rejects bad arguments on the host before any launch."""
from __future__ import annotations

import ctypes
import os
import re

import pytest
import torch

from cuda_knn_aggregate import _C, _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "aknn.h")).read()
    return sorted(set(re.findall(r"AKNN_API\s+[\w\s\*]*?\b(aknn_\w+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    syms = header_symbols()
    assert set(_lib.EXPORTS) == set(syms), syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    for name in header_symbols():
        assert hasattr(lib, name), name
    assert lib.aknn_abi_version() == _lib.ABI_VERSION


def _dims(**kw):
    base = dict(B=2, Q=10, P=8, F=4, K=3, dtype=torch.float32)
    base.update(kw)
    return _C.make_dims(base["B"], base["Q"], base["P"], base["F"], base["K"], base["dtype"])


@pytest.mark.parametrize("kw,code,msg", [
    (dict(K=0), _lib.AKNN_E_INVALID, "K must be >= 1"),
    (dict(K=9), _lib.AKNN_E_INVALID, "must not exceed"),
    (dict(K=65, P=100), _lib.AKNN_E_UNSUPPORTED, "K must be <= 64"),
    (dict(F=0), _lib.AKNN_E_INVALID, "bad sizes"),
    (dict(F=516), _lib.AKNN_E_UNSUPPORTED, "F must be <= 512"),
    (dict(F=129), _lib.AKNN_E_UNSUPPORTED, "F must be <= 128"),
    (dict(F=258, dtype=torch.float64), _lib.AKNN_E_UNSUPPORTED, "F must be <= 256"),
])
def test_forward_rejects_bad_dims(kw, code, msg):
    lib = _lib.load()
    d = _dims(**kw)
    rc = lib.aknn_forward(ctypes.byref(d), None, *([None] * 9), None, 0, None, None)
    assert rc == code
    assert msg in lib.aknn_last_error().decode()


def test_forward_rejects_small_workspace():
    lib = _lib.load()
    d = _dims()
    need = lib.aknn_forward_workspace_size(ctypes.byref(d))
    if need == 0:  # the sort-size query needs a device; the error names the query
        assert "workspace_size" in lib.aknn_last_error().decode()
        pytest.skip("no device for the workspace size query")
    assert need >= 2 * 8 * 16 * 4
    fake = ctypes.c_void_p(0x1000)
    rc = lib.aknn_forward(ctypes.byref(d), None, *([fake] * 9), fake, need - 1, None, None)
    assert rc == _lib.AKNN_E_WORKSPACE


def test_backward_rejects_missing_outputs_and_csr():
    lib = _lib.load()
    d = _dims()
    fake = ctypes.c_void_p(0x1000)
    rc = lib.aknn_backward(ctypes.byref(d), _lib.NEED_DQ, *([fake] * 7), None, None, None,
                           None, 0, None, None, None, None, None, None)
    assert rc == _lib.AKNN_E_INVALID and "dq requested" in lib.aknn_last_error().decode()
    rc = lib.aknn_backward(ctypes.byref(d), _lib.NEED_DP, *([fake] * 7), None, None, None,
                           None, 0, None, fake, None, None, None, None)
    assert rc == _lib.AKNN_E_INVALID and "CSR" in lib.aknn_last_error().decode()
    rc = lib.aknn_backward(ctypes.byref(d), _lib.NEED_DP | _lib.NEED_DQ, *([fake] * 7), fake, fake, None,
                           None, 0, fake, fake, None, None, None, None)
    assert rc == _lib.AKNN_E_WORKSPACE
    assert lib.aknn_backward_workspace_size(ctypes.byref(d), _lib.NEED_DP | _lib.NEED_DQ) >= 2 * 10 * 3 * 3 * 4
    assert lib.aknn_backward_workspace_size(ctypes.byref(d), _lib.NEED_DQ) == 0
    rc = lib.aknn_backward2(ctypes.byref(d), _lib.NEED_DQ, None, *([fake] * 7), None, None, None,
                            None, 0, None, fake, None, None, None, None, None)
    assert rc == _lib.AKNN_E_INVALID and "grad_grad_q" in lib.aknn_last_error().decode()


def test_ragged_offsets_validated():
    lib = _lib.load()
    qh = torch.tensor([0, 5, 3], dtype=torch.int64)      # decreasing
    ph = torch.tensor([0, 4, 8], dtype=torch.int64)
    d = _C.make_dims(2, 3, 8, 4, 2, torch.float32, offsets=(qh, ph, qh, ph))
    rc = lib.aknn_forward(ctypes.byref(d), None, *([ctypes.c_void_p(0x1000)] * 9), None, 0, None, None)
    assert rc == _lib.AKNN_E_INVALID and "offsets" in lib.aknn_last_error().decode()


def test_python_layer_rejects_cpu_and_noncontiguous():
    q = torch.rand(1, 4, 3)
    with pytest.raises(RuntimeError, match="CUDA tensor"):
        _C.knn_aggregate_aniso_forward(q, q, q, q, torch.rand(1, 4, 3, 3), 2)


@pytest.mark.parametrize("F,dt", [(512, torch.float32), (513, torch.float32), (1030, torch.float32),
                                  (1031, torch.float32), (129, torch.float32), (257, torch.float64),
                                  (600, torch.float64), (64, torch.float64)])
def test_feature_slabs_cover_the_row(F, dt):
    """Wide feature rows are split into contiguous column ranges, each within the
    library's per-call width (the same rule its argument check applies)."""
    slabs = _C._feature_slabs(F, dt)

    def cap(n):
        return 128 * (4 if dt == torch.float32 and n % 4 == 0 else 2 if n % 2 == 0 else 1)

    if F <= cap(F):
        assert slabs is None
        return
    assert slabs[0][0] == 0 and slabs[-1][1] == F
    for (a0, a1), (b0, _) in zip(slabs, slabs[1:]):
        assert a1 == b0
    for a0, a1 in slabs:
        assert 0 < a1 - a0 <= cap(a1 - a0)
