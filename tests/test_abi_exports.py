"""CPU-side checks of the drop-in boundary: the C-ABI library loads and exports every entry point that
include/fitsne_abi.h declares (no compute calls: there is no GPU here)."""
import os
import re

import pytest

import fitsne_gpu as fg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "fitsne_abi.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fitsne_[A-Za-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_mirror_and_resident_tiers():
    syms = declared_symbols()
    for s in ["fitsne_compute_fft_gradient", "fitsne_precompute_2d", "fitsne_n_body_fft_2d", "fitsne_precompute_1d",
              "fitsne_n_body_fft_1d", "fitsne_interpolate", "fitsne_evaluate_error_fft",
              "fitsne_distances2similarities_batch", "fitsne_ctx_create", "fitsne_step", "fitsne_kl"]:
        assert s in syms


@pytest.mark.skipif(not os.path.exists(fg.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    L = fg.lib()
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, missing
    # and the Python binding binds exactly the declared set
    assert sorted(fg.SIGNATURES) == declared_symbols()


@pytest.mark.skipif(not os.path.exists(fg.LIB_PATH), reason="library not built")
def test_no_device_is_an_error_not_a_fallback():
    if fg.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(fg.FitsneError) as e:
        fg.Context()
    assert e.value.code == fg.E_NODEVICE
