"""The C++ Simulation drop-in and its headless frame loop on the GPU."""
import os
import subprocess
import time
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


def test_pbf_run_render_staging():
    # CopyRenderBuffers into vectors and into caller-owned memory: the VBO is
    # the state's positions (w = speed), the texture has ceil(4096 / 2048) rows
    r = subprocess.run([str(RUN), str(ROOT / "tests" / "golden" / "dam4k.par"), "20", "10", "--render"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("render")]
    assert [ln[2] for ln in lines] == ["10", "20"]
    for ln in lines:
        assert ln[4] == "1" and ln[6] == "1" and ln[8] == "2", ln


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


def _frames(lines):
    return [ln.split() for ln in lines if ln.startswith("frame")]


def test_pbf_run_hot_reload(tmp_path):
    # the frame loop watches the scenario: a change of gravity is applied to
    # the running simulation (no reset), a change of particle count re-creates
    # it (src/Runner.cpp:43-77)
    par = tmp_path / "scene.par"
    text = (ROOT / "tests" / "golden" / "dam4k.par").read_text().replace("ResetSimOnChange        1",
                                                                          "ResetSimOnChange        0")
    par.write_text(text)
    proc = subprocess.Popen([str(RUN), str(par), "600", "5", "--watch", "--fps", "100"], stdout=subprocess.PIPE,
                            stderr=subprocess.PIPE, text=True)
    lines = []
    try:
        def read_until(pred, limit=60.0):
            t0 = time.time()
            while time.time() - t0 < limit:
                ln = proc.stdout.readline()
                if not ln:
                    break
                lines.append(ln.strip())
                if pred(ln):
                    return ln
            raise AssertionError("\n".join(lines[-20:]))

        def touch(new_text, step):
            par.write_text(new_text)
            t = time.time() + 10 * step
            os.utime(par, (t, t))

        read_until(lambda ln: ln.startswith("frame 20 "))
        touch(text.replace("Garvity                 9.8", "Garvity                 -30.0"), 1)
        rl = read_until(lambda ln: ln.startswith("reload")).split()
        assert rl[4] == "0" and rl[6] == "1"       # applied live, no reset
        f_reload = int(rl[2])
        read_until(lambda ln: ln.startswith("frame") and int(ln.split()[1]) >= f_reload + 40)
        touch(text.replace("ParticleCount           4096", "ParticleCount           3000"), 2)
        rl2 = read_until(lambda ln: ln.startswith("reload")).split()
        assert rl2[4] == "1" and rl2[6] == "1"     # new particle count: re-created
        read_until(lambda ln: ln.startswith("frame") and int(ln.split()[1]) >= int(rl2[2]) + 5)
    finally:
        proc.kill()
        proc.wait()
    fr = _frames(lines)
    before = [f for f in fr if int(f[1]) < f_reload]
    after = [f for f in fr if f_reload < int(f[1]) < int(rl2[2])]
    assert all(f[3] == "4096" for f in before + after)
    # upward gravity: the top of the block rises after the reload
    assert float(after[-1][9]) > float(after[0][9]) + 0.05
    assert [f for f in fr if int(f[1]) > int(rl2[2])][0][3] == "3000"
