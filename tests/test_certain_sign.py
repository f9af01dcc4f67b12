"""The PIP test's certain-sign rule, checked on the CPU (tests/cpp/certain_sign_check.c).

The HIP test decides a (point, edge) from the scaled line form only when
|t| > 2^-48 s + 2^-1000; the checker draws adversarial (point, edge) pairs (points
on, beside and along segments, nearly vertical and horizontal edges, shared
coordinates, subnormal to 2^120 scales) and compares every certain decision with
the reference's rounded lhs < rhs and its 4-ULP collinearity test. The same file
built with a smaller tolerance, or with the lower limit on the edge's size
removed, finds counterexamples: the check has the power to fail.
"""

import os
import shutil
import subprocess

import pytest

SRC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "certain_sign_check.c")


def build(tmp_path, name, *defines):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = str(tmp_path / name)
    subprocess.run([cc, "-O2", "-ffp-contract=off", *defines, SRC, "-o", exe, "-lm"],
                   check=True, capture_output=True)
    return exe


def run(exe, iters, seed):
    r = subprocess.run([exe, str(iters), str(seed)], capture_output=True, text=True)
    counts = dict(zip(r.stdout.split()[-8::2], map(int, r.stdout.split()[-7::2])))
    return r.returncode, counts, r.stdout


def test_certain_decisions_match_the_reference(tmp_path):
    exe = build(tmp_path, "csc")
    for seed in (1, 2, 3):
        rc, counts, out = run(exe, 3_000_000, seed)
        assert rc == 0 and counts["bad"] == 0, out
        assert counts["certain"] > 500_000 and counts["no_line"] > 0, out


@pytest.mark.parametrize("define", ["-DTOL_SCALE=0x1p-53", "-DN_MIN=0x1p-1074"])
def test_weaker_rules_fail(tmp_path, define):
    exe = build(tmp_path, "csc_weak", define)
    rc, counts, out = run(exe, 3_000_000, 1)
    assert rc == 1 and counts["bad"] > 0, out
