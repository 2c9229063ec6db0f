"""The drop-in claim from C: tests/c/so3_roundtrip.c, a C99 program that
includes <so3/so3.h> and links libso3_amd.so the way a caller of the
reference library would, compiles and links here (no GPU needed) and, on
the GPU, round-trips complex and real transforms at L = N = 16 (the
reference harness's configuration, src/c/so3_test_csv.c:132-135) and checks
the direct pair, an adjoint, s2toso3 and the index helpers."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "so3_amd")


def _build(tmp_path):
    exe = str(tmp_path / "so3_roundtrip")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c", "so3_roundtrip.c"), "-L", LIBDIR, "-lso3_amd",
           "-Wl,-rpath," + LIBDIR, "-Wl,--allow-shlib-undefined", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def test_c99_caller_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_c99_caller_round_trip(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "16", "16"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("OK"), r.stdout + r.stderr
