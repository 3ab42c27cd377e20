"""The C-ABI library builds for gfx950, loads, and exports every symbol the
public header declares (no GPU needed; no compute call is made)."""
import ctypes
import os

import pytest

from molfdir_amd import relccsd


def test_header_declares_the_reference_entry_points():
    syms = relccsd.declared_symbols()
    for name in ("rcc_create", "rcc_set_fock", "rcc_set_ints", "rcc_mp2", "rcc_hointm",
                 "rcc_hvintm", "rcc_hovintm", "rcc_t1eqns", "rcc_gointm", "rcc_gvintm",
                 "rcc_aintm", "rcc_hintm", "rcc_t2eqns", "rcc_bintm", "rcc_denom", "rcc_zcore",
                 "rcc_diis", "rcc_energy", "rcc_t1diag", "rcc_iterate", "rcc_cceqn"):
        assert name in syms


@pytest.mark.skipif(not os.path.exists(relccsd.LIB_PATH), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(relccsd.LIB_PATH)
    missing = [s for s in relccsd.declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(relccsd, "_lib", None)
    monkeypatch.setattr(relccsd, "LIB_PATH", str(tmp_path / "absent.so"))
    with pytest.raises(relccsd.RccError):
        relccsd.load()
