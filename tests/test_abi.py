"""CPU tests of the C-ABI library: it loads without a GPU and exports every
function that include/plutomx.h declares (no compute call is made)."""
import ctypes
import pathlib
import re

import pytest

import plutomx
import setups

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


def test_header_compiles_as_c():
    """include/plutomx.h is a plain C header (C callers such as PLUTO's
    driver include it directly)."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    hdr = pathlib.Path(__file__).resolve().parents[1] / "include" / "plutomx.h"
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Werror", "-fsyntax-only", "-x", "c", str(hdr)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("struct,cname", [("WindDesc", "pmx_wind_desc"),
                                          ("HydroDesc", "pmx_hydro_desc")])
def test_descriptor_layout_matches_header(tmp_path, struct, cname):
    """The ctypes mirrors of the descriptors have the C structs' size and
    the offset of every field."""
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    py = getattr(plutomx, struct)
    names = [f[0] for f in py._fields_]
    src = tmp_path / "sz.c"
    body = "".join(f' printf("%zu\\n", offsetof({cname}, {n}));' for n in names)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "plutomx.h"\n'
                   f'int main(void){{printf("%zu\\n", sizeof({cname}));{body} return 0;}}\n')
    exe = tmp_path / "sz"
    subprocess.run([cc, "-I", str(ROOT / "include"), str(src), "-o", str(exe)], check=True)
    vals = list(map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                        check=True).stdout.split()))
    assert ctypes.sizeof(py) == vals[0]
    for n, off in zip(names, vals[1:]):
        assert getattr(py, n).offset == off, n


def test_wind_descriptor_validation_without_gpu():
    """pmx_wind_create checks the descriptor before it touches the device,
    and the Python wrapper passes every field (dfloor, boundary types)."""
    import dataclasses
    import setups
    bad = [dataclasses.replace(setups.cvidl_wind(16, 8), dfloor=-1.0),
           dataclasses.replace(setups.cvidl_wind(16, 8),
                               bc=(setups.BC_WIND_POLE, 1, setups.BC_WIND_INNER, 2)),
           dataclasses.replace(setups.cvidl_wind(16, 8), bc=(9, 1, 2, 2)),
           dataclasses.replace(setups.cvidl_wind(16, 8), krad=999.0, alpharad=999.0)]
    msgs = ["dfloor must be >= 0", "is an r-beg", "unknown boundary type",
            "needs the force multiplier fits"]
    for cfg, msg in zip(bad, msgs):
        with pytest.raises(plutomx.PmxError, match=msg):
            plutomx.Wind(cfg)


def test_noalloc_step_refuses_grids_over_its_limit():
    """pmx_hydro_advance_step_noalloc checks the grid against the size of
    its fixed device array before any device work."""
    lib = plutomx.load_library()
    cfg = setups.c2_sedov(128)            # 2.1 M cells, over the 524288-cell limit
    keep = []
    d = plutomx.hydro_desc(cfg, keep)
    inv = ctypes.c_double(0.0)
    rc = lib.pmx_hydro_advance_step_noalloc(ctypes.byref(d), ctypes.c_void_p(16),
                                            ctypes.c_void_p(16), 1.e-3, ctypes.byref(inv))
    assert rc == -1
    assert b"exceeds the limit of 524288 cells" in lib.pmx_last_error()


def test_zero_copy_snapshot_view_lifetime():
    """A zero-copy snapshot view is read-only, is tracked weakly by its
    context, and a context closed while a view is referenced is destroyed
    when the last view goes (host logic of Wind.snapshot_wait(copy=False)
    and Wind.close, with a stand-in for the library)."""
    import gc
    import weakref
    import numpy as np
    buf = (ctypes.c_double * 12)(*range(12))
    holder = plutomx._PinnedView(ctypes.cast(buf, ctypes.POINTER(ctypes.c_double)), (3, 4))
    views = weakref.WeakSet([holder])
    a = np.asarray(holder)
    del holder
    assert not a.flags.writeable and a[2, 3] == 11.0
    part = a[1:]
    del a
    assert len(views) == 1          # a slice keeps the view's owner alive

    class Lib:
        destroyed = []

        def pmx_wind_destroy(self, h):
            self.destroyed.append(h)

    lib = Lib()
    plutomx._DeferredDestroy(lib, "ctx", list(views))
    gc.collect()
    assert lib.destroyed == []
    del part
    gc.collect()
    assert lib.destroyed == ["ctx"] and len(views) == 0
