"""pytest configuration: import paths and the ``gpu`` marker.

Tests marked ``gpu`` need a visible AMD GPU and the built libplutomx.so;
everything else runs on the CPU (oracle, golden fixtures, host logic,
library exports)."""
import pathlib
import sys

ROOT = pathlib.Path(__file__).resolve().parents[1]
for p in (ROOT / "pluto-sirocco_amd", ROOT):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and libplutomx.so")
