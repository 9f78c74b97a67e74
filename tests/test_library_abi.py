"""The C ABI library loads and exports every symbol include/tenes_amd.h
declares (no GPU needed: no compute call is made)."""
import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "tenes_amd.h"


def declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tenes_[a-z0-9_]+)\s*\(", txt)))


def test_header_parses():
    names = declared()
    assert "tenes_calc_ctm_environment" in names
    assert len(names) >= 30


def test_library_exports_all_declared_symbols():
    from tenes_amd import _lib

    lib = _lib.load()
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", str(_lib.LIB_PATH)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (tenes_[a-z0-9_]+)", out))
    assert set(declared()) <= exported


def test_ctypes_signatures_cover_header():
    from tenes_amd import _lib

    assert set(_lib.exported_symbols()) == set(declared())


def test_parameter_defaults_match_reference():
    """PEPS_Parameters.cpp:29-77 and transfer_matrix.hpp:38-45 defaults,
    through host-only entry points."""
    import tenes_amd as ta

    p = ta.peps_parameters()
    assert p.CHI == 2 and p.Max_CTM_Iteration == 100 and p.CTM_Projector_corner == 1
    assert p.Inverse_projector_cut == 1e-12 and p.CTM_Convergence_Epsilon == 1e-6
    assert p.Use_RSVD == 0 and p.RSVD_Oversampling_factor == 2.0
    t = ta.tm_parameters()
    assert (t.num_eigvals, t.maxdim_dense_eigensolver, t.arnoldi_maxdim, t.arnoldi_maxiter) == (4, 200, 50, 1)
    assert t.arnoldi_rtol == 1e-10


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    from tenes_amd import _lib

    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", tmp_path / "nope.so")
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        _lib.load()
