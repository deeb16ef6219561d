"""The drop-in drivers' number formatting (cahn-hilliard_amd/host/fmt_general.h)
against the reference's own bundled fmt 7.1.3: "{:.5}" and "{:.8}" of a set of
doubles (tests/golden/fmt_vectors.json, written by make_fmt_vectors.py from the
reference's vendored fmt headers) -- the energy.dat columns and the native /
VTK headers.  CPU only: compiles a small program with g++."""
import json
import os
import subprocess

from conftest import REPO


def test_fmt_general_matches_reference_fmt(tmp_path):
    vec = json.load(open(os.path.join(REPO, "tests", "golden", "fmt_vectors.json")))
    src = tmp_path / "f.cpp"
    src.write_text('#include "fmt_general.h"\n#include <cstdlib>\n'
                   'int main(int argc, char **argv) {\n'
                   '    for(int i = 1; i < argc; i++) {\n'
                   '        double v = std::strtod(argv[i], nullptr);\n'
                   '        std::printf("%s|%s\\n", ch::fmt_general(v, 5).c_str(), ch::fmt_general(v, 8).c_str());\n'
                   '    }\n}\n')
    exe = tmp_path / "f"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(REPO, "cahn-hilliard_amd", "host"), str(src), "-o",
                    str(exe)], check=True)
    vals = [d["value"] for d in vec["doubles"]]
    out = subprocess.run([str(exe)] + [repr(v) for v in vals], capture_output=True, text=True,
                         check=True).stdout.split("\n")
    for d, line in zip(vec["doubles"], out):
        assert line == f"{d['p5']}|{d['p8']}", (d, line)
