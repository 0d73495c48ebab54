"""The C++ Simulation drop-in and its headless frame loop on the GPU."""
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
RUN = ROOT / "position-based-fluids_amd" / "bin" / "pbf_run"

pytestmark = pytest.mark.gpu


def test_pbf_run_frame_loop():
    assert RUN.exists(), "build first"
    r = subprocess.run([str(RUN), str(ROOT / "tests" / "golden" / "dam4k.par"), "30", "10"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("frame")]
    assert [int(ln[1]) for ln in lines] == [10, 20, 30]
    for ln in lines:
        assert ln[3] == "4096" and ln[-1] == "1"
        ymin, ymax = float(ln[7]), float(ln[9])
        assert 0.0 <= ymin and ymax <= 2.0
    # the block falls: kinetic energy grows from zero
    assert float(lines[-1][5]) > 0.0


def test_pbf_run_with_a_ramp():
    # LoadForceMasks with one solid ramp quad: the run stays healthy and differs
    # from the plain run
    out = []
    for extra in ([], ["ramp"]):
        r = subprocess.run([str(RUN), str(ROOT / "tests" / "golden" / "dam4k.par"), "20", "10", *extra],
                           capture_output=True, text=True, timeout=120)
        assert r.returncode == 0, r.stderr
        lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("frame")]
        assert [ln[-1] for ln in lines] == ["1", "1"]
        out.append(lines)
    assert out[0][-1][5] != out[1][-1][5]
