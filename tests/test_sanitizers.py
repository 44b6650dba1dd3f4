"""AddressSanitizer and UndefinedBehaviorSanitizer builds, on the CPU, of the C host part of
libmcgb (csrc/mcgb_host.c), the C driver (host/mcgb_main.c) and the FP64 oracle
(oracle/mcgb_oracle.c): each is compiled with gcc and -fsanitize=address,undefined (no recovery)
together with a checker under tests/sanitize/ and run; any invalid access, leak or undefined
operation aborts the program with a report.  The driver is linked with tests/sanitize/no_gpu.c,
whose GPU entry points fail as the library does without a device, and runs up to that point:
parameter file, start configuration, the first trajectory frames, the error exit."""
import os
import re
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "parellel_mc_mpi_for_liquidcrystalsystems_amd")
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(REPO, "include")
ORC = os.path.join(REPO, "oracle")
SAN = os.path.join(REPO, "tests", "sanitize")
GOLD = os.path.join(REPO, "tests", "golden")
FLAGS = ["-std=gnu11", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-static-libasan", "-Wall"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1")

pytestmark = pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")


def build(out, sources, extra=()):
    cmd = ["gcc", *FLAGS, "-I", INC, "-I", CSRC, "-I", ORC, *sources, "-o", str(out), "-lm", *extra]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return str(out)


def clean(r):
    """No sanitizer report in the output."""
    text = r.stdout + r.stderr
    return not re.search(r"AddressSanitizer|LeakSanitizer|runtime error:", text), text


def test_host_part_under_sanitizers(tmp_path):
    exe = build(tmp_path / "host_check", [os.path.join(SAN, "host_check.c"),
                                          os.path.join(CSRC, "mcgb_host.c")])
    r = subprocess.run([exe, GOLD, str(tmp_path)], capture_output=True, text=True, env=ENV,
                       timeout=300)
    ok, text = clean(r)
    assert r.returncode == 0 and ok and "host_check: ok" in r.stdout, text


def test_oracle_under_sanitizers(tmp_path):
    exe = build(tmp_path / "oracle_check", [os.path.join(SAN, "oracle_check.c"),
                                            os.path.join(ORC, "mcgb_oracle.c")], ["-fopenmp"])
    r = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=600)
    ok, text = clean(r)
    assert r.returncode == 0 and ok and "oracle_check: ok" in r.stdout, text


def test_driver_under_sanitizers(tmp_path):
    exe = build(tmp_path / "mcgb", [os.path.join(PKG, "host", "mcgb_main.c"),
                                    os.path.join(CSRC, "mcgb_host.c"),
                                    os.path.join(SAN, "no_gpu.c")])
    r = subprocess.run([exe, str(tmp_path / "nope.txt")], capture_output=True, text=True, env=ENV,
                       cwd=tmp_path, timeout=120)
    ok, text = clean(r)
    assert r.returncode == 1 and ok and "Error opening the parameters file" in r.stderr, text
    txt = open(os.path.join(GOLD, "params_n8000.txt")).read()
    for k, v in (("num_particles", 1000), ("system_density", 0.02),
                 ("initial_configuration_type", 2), ("total_mc_cycles", 0)):
        txt = re.sub(r"(?m)^%s\s+\S+" % k, "%s %s" % (k, v), txt)
    (tmp_path / "parametersFile.txt").write_text(txt)
    r = subprocess.run([exe, "parametersFile.txt"], capture_output=True, text=True, env=ENV,
                       cwd=tmp_path, timeout=120)
    ok, text = clean(r)
    assert r.returncode == 1 and ok and "cannot create the GPU context" in r.stderr, text
    assert (tmp_path / "Configuraciones.dat").stat().st_size > 0
