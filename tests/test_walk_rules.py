"""CPU check of the voxel walk's loop rules (tests/cprobe/walk_rules.c): the packed mirrored-coordinate
walk of csrc/vf_device.h against the one-step walk that tracks each voxel's entry time and face, over
random bricks and rays with exact tMax ties; and a planted mistake (tMax advanced on the exit trip) is caught."""
import os
import shutil
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "cprobe", "walk_rules.c")


def _build(tmp, defs=()):
    exe = os.path.join(tmp, "walk_rules")
    src = SRC
    if defs:
        text = open(SRC).read()
        for a, b in defs:
            assert a in text
            text = text.replace(a, b)
        src = os.path.join(tmp, "walk_rules_planted.c")
        open(src, "w").write(text)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_packed_walk_matches_one_step_walk():
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([_build(tmp), "2000000"], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout
        assert "mismatches 0" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_planted_exit_trip_mistake_is_caught():
    # tMax advanced on the exit trip too: a hit's entry time is then one tDelta late
    planted = [("volatile float nx = tMx + ((a0 && !stop) ? tDx : -0.0f);", "volatile float nx = tMx + (a0 ? tDx : -0.0f);"),
               ("volatile float ny = tMy + ((a1 && !stop) ? tDy : -0.0f);", "volatile float ny = tMy + (a1 ? tDy : -0.0f);"),
               ("volatile float nz = tMz + ((a2 && !stop) ? tDz : -0.0f);", "volatile float nz = tMz + (a2 ? tDz : -0.0f);")]
    with tempfile.TemporaryDirectory() as tmp:
        r = subprocess.run([_build(tmp, planted), "200000"], capture_output=True, text=True)
        assert r.returncode != 0 and "mismatches 0" not in r.stdout
