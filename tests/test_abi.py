"""The C-ABI library loads, exports every symbol include/mifft.h declares,
and reports the reference's error cases from host-side checks (no GPU)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "mifft.h")


def declared_symbols():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"MIFFT_API\s+[^;(]*?\b(mifft_\w+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "mifft_plan_create" in names and "mifft_execute_device" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    from metalfft_amd import _lib
    L = _lib.load()
    for name in declared_symbols():
        assert hasattr(L, name), f"{name} declared in mifft.h but not exported"


def test_binding_table_matches_header():
    from metalfft_amd import _lib
    assert sorted(_lib.exported_names()) == declared_symbols()


def test_unsupported_size_is_reported_before_device_use():
    import metalfft_amd as m
    for bad in (0, 32, 100, 4095, 3 * 16384, 1 << 21, -4096):
        with pytest.raises(m.FFTError) as ei:
            m.MetalFFT(bad)
        assert ei.value.case == "unsupportedFFTSize"
        assert ei.value.got == bad or bad < 0


def test_status_strings():
    from metalfft_amd import _lib
    L = _lib.load()
    assert b"Unsupported FFT size" in L.mifft_status_string(8)
    for code in range(12):
        assert L.mifft_status_string(code)


def test_variant_listing():
    import metalfft_amd as m
    for n in m.SUPPORTED_SIZES:
        names = m.kernel_variants(n)
        assert names and all(str(n) in s for s in names)
    assert m.kernel_variants(100) == []


def test_null_arguments_rejected():
    from metalfft_amd import _lib
    L = _lib.load()
    assert L.mifft_plan_create(4096, None) == 11
    assert L.mifft_execute_device(None, None, None, 1, -1, None) == 11
    assert L.mifft_forward_host(None, None, 0, None) == 11
