"""The pins of the CPU oracle catch plausible mistakes (CPU only).

oracle/vrt_oracle.cpp carries MUT(k) switches: built with -DORC_MUTANT=k it
makes one deliberate mistake of the kind a transcription of fshader.glsl,
render.cpp or level.cpp could make (a dropped term, a flipped sign or
comparison, a transposed operand, a wrong rounding).  Each mutant is built
into a temporary directory and tests/test_oracle_pins.py is run against it in
a child process together with tests/test_oracle_lazy.py (VRT_ORACLE_LIB
points the loader at the mutant); the pins must fail on every one of them.  A mutant that survives would be a mistake
the oracle could carry unnoticed, and through it the kernels."""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "vrt_oracle.cpp")
FLAGS = ["-std=c++17", "-O1", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fwrapv", "-fPIC", "-shared"]

MUTANTS = {
    1: "DDA ties go to x (<= instead of <)",
    2: "step bias sign flipped (-1e-6)",
    3: "start cell by floor instead of truncation",
    4: "depth-field jump does not advance the start",
    5: "step 0 counted as a forward step",
    6: "normalize as v / |v| instead of v * (1 / |v|)",
    7: "overbright test > 1.25 instead of >= 1.25",
    8: "local light range < instead of <=",
    9: "depth table offsets not shrunk toward 0",
    10: "UNORM8 conversion truncates (no +0.5)",
    11: "primary ray not normalized",
    12: "camera matrix read transposed",
    13: "a jump counted as two steps",
    14: "removeSphere carves r^2 <= instead of <",
    15: "fixDepthField stores distances above -2",
    16: "falloff d / r instead of (r - d) / r",
    17: "sun lambert term not clamped at 0",
    18: "light with a zero coordinate treated as unused",
    19: "red and blue channels swapped",
    20: "local-light shadow ray without its distance budget",
    21: "removeSphere repairs only the carved ball",
    22: "depth view scaled by 1/255 instead of 1/100",
    23: "aspect ratio applied to y",
    24: "overbright clamp applied after the last light too",
    25: "on-demand world: repair ball of radius r + 2",
    26: "on-demand world: raw-empty cells keep the first build's word after edits",
}


def test_every_switch_is_listed():
    text = open(SRC).read()
    ids = {int(t.split(")")[0]) for t in text.split("MUT(")[1:] if t.split(")")[0].isdigit()}
    assert ids == set(MUTANTS)


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    d = tmp_path_factory.mktemp("mutants")

    def build(k):
        out = str(d / ("liboracle_m%d.so" % k))
        subprocess.run(["g++"] + FLAGS + ["-DORC_MUTANT=%d" % k, "-o", out, SRC], check=True)
        return k, out

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        return dict(ex.map(build, sorted(MUTANTS)))


def run_pins(lib):
    env = dict(os.environ, VRT_ORACLE_LIB=lib, OMP_NUM_THREADS="2")
    cmd = [sys.executable, "-m", "pytest", os.path.join(ROOT, "tests", "test_oracle_pins.py"),
           os.path.join(ROOT, "tests", "test_oracle_lazy.py"), "-x", "-q", "-p", "no:cacheprovider"]
    return subprocess.run(cmd, env=env, cwd=ROOT, capture_output=True, text=True, timeout=600)


def test_pins_catch_every_mutant(mutant_libs):
    with ThreadPoolExecutor(max_workers=4) as ex:
        results = dict(zip(sorted(mutant_libs), ex.map(run_pins, [mutant_libs[k] for k in sorted(mutant_libs)])))
    survivors = [k for k, r in results.items() if r.returncode == 0]
    broken = [k for k, r in results.items() if r.returncode not in (0, 1)]
    assert not broken, {k: results[k].stdout[-2000:] + results[k].stderr[-2000:] for k in broken}
    assert not survivors, ["%d: %s" % (k, MUTANTS[k]) for k in survivors]
