"""CPU checks of the C ABI library: it loads, exports every declared entry
point, and fails loudly (no CPU fallback) where no gfx950 device is visible."""
import ctypes as C
import os
import re

import pytest

import tasep_amd as T


def _declared():
    with open(T.HEADER) as f:
        src = f.read()
    return sorted(set(re.findall(r"\b(tasep_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = T.load_library()
    names = _declared()
    assert len(names) >= 18
    missing = [nm for nm in names if not hasattr(L, nm)]
    assert not missing, missing
    assert set(T.EXPORTS) == set(names)


def test_csr_nnz_closed_form():
    for n in (2, 10, 28, 30):
        assert T.csr_nnz(n) == 2 * (1 << n) + (n - 1) * (1 << (n - 2))
    assert T.csr_nnz(1) == 0


def _gpu_visible():
    return os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK)


@pytest.mark.skipif(_gpu_visible(), reason="a GPU is visible here")
def test_no_device_raises():
    with pytest.raises(T.TasepError) as ei:
        T.Engine(8, "f64", "ref", "mf")
    assert ei.value.code in (-5, -3)


def test_strerror_codes():
    L = T.load_library()
    for code in (0, -1, -2, -3, -4, -5, -99):
        assert isinstance(L.tasep_strerror(code), bytes)
