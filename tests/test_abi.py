"""CPU tests of the C-ABI library: it loads without a GPU and exports every
function that include/plutomx.h declares (no compute call is made)."""
import ctypes
import pathlib
import re

import pytest

import plutomx

ROOT = pathlib.Path(__file__).resolve().parents[1]


def declared_functions():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(pmx_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_headers_declare_functions():
    assert "pmx_hydro_advance_step" in declared_functions()


@pytest.mark.parametrize("name", declared_functions())
def test_library_exports(name):
    lib = ctypes.CDLL(str(plutomx.LIB_PATH))
    assert hasattr(lib, name), f"{name} declared but not exported"


def test_python_binding_covers_exports():
    assert set(plutomx.EXPORTS) <= set(declared_functions())


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(plutomx.PmxError):
        plutomx.load_library(tmp_path / "nope.so")
