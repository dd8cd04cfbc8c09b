"""The drop-in glue (pluto-sirocco_amd/glue/pmx_glue.c) compiles against
each drop-in problem's definitions.h and PLUTO's headers, and refuses at
compile time a configuration the device step does not implement: FAILSAFE
YES (Src/failsafe.c:5-158, Src/main.c:262-266), whose retry needs the step
to stop at the first failing stage (Src/Time_Stepping/rk_step.c:186-188).

Compile-only (gcc -fsyntax-only through oracle/ref/Makefile's glue-syntax
target); needs PLUTO's sources, so it is skipped where they are absent."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]
REF_SRC = pathlib.Path("/root/reference/Src")
PROBLEMS = ["cart2d_plm_vl_rk2", "cart3d_ppm_rk3", "wind_plain", "wind_blondin",
            "wind_cvidl_prod"]

pytestmark = pytest.mark.skipif(not (REF_SRC / "pluto.h").exists(),
                                reason="PLUTO's sources are not present")


def glue_syntax(problem, defs=""):
    return subprocess.run(["make", "-s", "-C", str(ROOT / "oracle" / "ref"), f"PROB={problem}",
                           "GLUE=1", f"GLUE_DEFS={defs}", "glue-syntax"],
                          capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("problem", PROBLEMS)
def test_glue_compiles_for_drop_in_problems(problem):
    r = glue_syntax(problem)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.parametrize("problem", ["cart2d_plm_vl_rk2", "wind_cvidl_prod"])
def test_glue_refuses_failsafe(problem):
    r = glue_syntax(problem, "-DFAILSAFE=YES")
    assert r.returncode != 0
    assert "FAILSAFE YES is not supported" in r.stderr
    # FAILSAFE NO (PLUTO's default, Src/pluto.h:410-412) spelled out builds
    assert glue_syntax(problem, "-DFAILSAFE=NO").returncode == 0
