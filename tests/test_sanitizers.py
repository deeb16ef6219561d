"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the CPU code: the
oracle (oracle/ch_oracle.c, exercised by tests/sanitize/oracle_driver.c on
every integrator, model, transform and adjoint) and the drop-in's TOML parser
(cahn-hilliard_amd/host/toml_lite.cpp, on the committed inputs and malformed
snippets).  Host code only, built with gcc / g++; a report fails the test."""
import glob
import os
import shutil
import subprocess

import pytest

from conftest import REPO

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g", "-O1"]


def run_clean(cmd, env_extra=None):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0 and "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, \
        r.stderr[-4000:]
    return r


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc absent")
def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_driver"
    subprocess.run(["gcc", "-std=c11"] + SAN + ["-I", os.path.join(REPO, "oracle"),
                    os.path.join(REPO, "tests", "sanitize", "oracle_driver.c"), os.path.join(REPO, "oracle", "ch_oracle.c"),
                    "-lm", "-o", str(exe)], check=True)
    r = run_clean([str(exe)])
    assert r.stdout.startswith("checksum")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ absent")
def test_toml_parser_under_sanitizers(tmp_path):
    exe = tmp_path / "toml_driver"
    host = os.path.join(REPO, "cahn-hilliard_amd", "host")
    subprocess.run(["g++", "-std=c++17"] + SAN + ["-I", host, os.path.join(REPO, "tests", "sanitize", "toml_driver.cpp"),
                    os.path.join(host, "toml_lite.cpp"), "-o", str(exe)], check=True)
    inputs = sorted(glob.glob(os.path.join(REPO, "tests", "golden", "*.toml")))
    assert inputs
    r = run_clean([str(exe)] + inputs)
    assert r.stdout.startswith("parsed")
