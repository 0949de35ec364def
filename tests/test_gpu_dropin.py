"""The drop-in itself: the reference's own executables (eval/keygen.cpp,
encrypt.cpp, cloud.cpp, decrypt.cpp), built against this build's idash.h and
libraries by oracle/dropin.mk, run beside this build's executables on the same
reference-format input files. Every file both write (params.bin, keys.bin,
encrypted_data.bin, encrypted_prediction.bin, result.csv, result_bypos.csv)
must be identical byte for byte, and so must their printouts apart from times
and memory use."""
import filecmp
import os
import re
import subprocess

import pytest

from idash2019_2_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MINE = os.path.join(ROOT, "idash2019_2_amd", "host")
THEIRS = os.path.join(ROOT, "oracle", "_ref", "dropin")
STEPS = [("keygen", ["targets.txt", "tags.txt", "1"]), ("encrypt", ["tags.txt"]),
         ("cloud", ["model"]), ("decrypt", []), ("decrypt", ["bypos"])]
FILES = ["params.bin", "keys.bin", "encrypted_data.bin", "encrypted_prediction.bin", "result.csv",
         "result_bypos.csv"]
VOLATILE = re.compile(r"time|RAM usage|Number of threads")


def run_all(exe_dir, cwd):
    printed = []
    for exe, args in STEPS:
        r = subprocess.run([os.path.join(exe_dir, exe), *args], cwd=cwd, capture_output=True,
                           text=True, timeout=300, env=dict(os.environ, IHE_SEED="9"))
        assert r.returncode == 0, f"{exe_dir}/{exe}: {r.stdout}{r.stderr}"
        printed += [ln for ln in r.stdout.splitlines() if not VOLATILE.search(ln)]
    return printed


@pytest.mark.skipif(not all(os.path.exists(os.path.join(THEIRS, e)) for e, _ in STEPS),
                    reason="reference executables not built (oracle/dropin.mk needs the "
                           "reference sources)")
@pytest.mark.parametrize("ns", [128, 1024, 100, "small"])
def test_reference_mains_write_the_same_files(tmp_path, ns):
    # "small": the 500-tag / 2,000-target / 128-sample workload itself
    c = synth.workload("small") if ns == "small" else synth.generate(60, 150, ns, 5, seed=ns + 1)
    a, b = tmp_path / "theirs", tmp_path / "mine"
    for d in (a, b):
        d.mkdir()
        synth.write_reference_files(c, str(d), nan_fraction=0.02, seed=2)
    out_a = run_all(THEIRS, str(a))
    out_b = run_all(MINE, str(b))
    assert out_a == out_b
    for f in FILES:
        assert filecmp.cmp(a / f, b / f, shallow=False), f
