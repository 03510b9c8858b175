"""The explicit kernel's shortened division by a launch constant (div_k in kernels.hip:
q = a * RN(1/b), corrected once with the exact remainder a - b q).  Markstein's theorem
makes it the correctly rounded quotient when no intermediate leaves the normal range; the
engine uses it for numerators 0 or in [2^-80, 2^80] and divisors 72, 12, dx^2 and dy^2 in
[2^-20, 2^20].  Here every significand of the numerator at the range's ends and middle is
checked against IEEE division for those divisors and for random ones in the range; a plain
a * RN(1/b) must fail the same check (so the check can fail)."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def bits(x):
    return "%08x" % struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    d = tmp_path_factory.mktemp("fastdiv")
    src = os.path.join(HERE, "fast_div_check.c")
    out = {}
    for name in ("exact", "plain"):
        text = open(src).read()
        if name == "plain":
            text = text.replace("fmaf(fmaf(-b, q, aa), rb, q)", "q")
        path = d / f"{name}.c"
        path.write_text(text)
        exe = str(d / name)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, str(path), "-lm"], check=True)
        out[name] = exe
    return out


def divisors():
    rng = np.random.default_rng(7)
    fixed = [72.0, 12.0, 1.0, 0.25, 0.0625, 2.0 ** -20, 2.0 ** 20, 0.3 ** 2, 1.7 ** 2]
    rand = np.exp2(rng.uniform(-20, 20, 6))
    return [np.float32(v) for v in list(fixed) + list(rand)]


def test_shortened_division_is_correctly_rounded(checker):
    r = subprocess.run([checker["exact"]] + [bits(b) for b in divisors()], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout


def test_uncorrected_product_is_not(checker):
    r = subprocess.run([checker["plain"], bits(72.0), bits(12.0)], capture_output=True, text=True)
    assert r.returncode == 1 and not r.stdout.startswith("0 of")
