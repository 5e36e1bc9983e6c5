"""The C-ABI library loads and exports every entry point include/kmsd.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import re
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "kmsd.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(kmsd_\w+)\s*\(", text, re.M)))


def test_header_declares_the_boundary():
    names = declared()
    for must in ("kmsd_moments", "kmsd_chain", "kmsd_series", "kmsd_h_action", "kmsd_v_cov",
                 "kmsd_metric_solve", "kmsd_metric_sample", "kmsd_matrix_matmul", "kmsd_mu_table"):
        assert must in names


def test_library_exports_every_declared_symbol():
    from kubomsd_amd.backends import hip
    lib = ctypes.CDLL(str(hip.LIB_PATH))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(declared()) <= set(hip.SIGNATURES)


def test_build_info_and_missing_device_is_loud():
    import torch
    from kubomsd_amd.backends import hip
    info = hip.build_info()
    assert info["target"] == "gfx950"
    if not torch.cuda.is_available():
        import pytest
        from kubomsd_amd.errors import BackendUnavailable
        with pytest.raises(BackendUnavailable):
            hip.HipContext(0)
