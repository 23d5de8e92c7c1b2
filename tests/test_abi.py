"""The C-ABI library loads and exports every entry point that include/hpq.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "hpq.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hpq_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("hpq_deposit_current", "hpq_explicit_deposition", "hpq_deposit_current_slice",
              "hpq_advance_plasma_particles", "hpq_poisson_dirichlet", "hpq_solve_poisson_psi_exmby_eypbx_ez_bz",
              "hpq_solve_poisson_bxby", "hpq_mg_solve1", "hpq_solve_one_slice"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    from hipace2_amd import _abi
    if not os.path.exists(_abi.LIB_PATH):
        pytest.skip("libhpq.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_abi.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(_abi.EXPORTS) == declared_symbols()


def test_product_does_not_reference_the_oracle():
    """The product package never imports, links or loads the oracle."""
    pkg = os.path.join(ROOT, "hipace2_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower(), os.path.join(dirpath, f)


def test_layout_names_match_header():
    """hpq_layout() writes HPQ_LAYOUT_N entries; the binding names every one of them."""
    from hipace2_amd import _abi
    text = open(os.path.join(ROOT, "include", "hpq.h")).read()
    n = int(re.search(r"#define HPQ_LAYOUT_N (\d+)", text).group(1))
    assert len(_abi.LAYOUT_NAMES) == n and len(set(_abi.LAYOUT_NAMES)) == n
