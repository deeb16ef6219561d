#!/usr/bin/env python3
"""Write tests/golden/fmt_vectors.json: how the reference's bundled fmt 7.1.3
(extern/spdlog/include/spdlog/fmt/bundled, header-only) prints doubles with
"{:.5}" and "{:.8}" and a step with "{:L}" -- the energy.dat line of
src/main.cpp:121 and the native/VTK headers of src/utils/Printer.cpp.

Runs in the development container only: it compiles a five-line program
against the reference's vendored fmt headers (no reference source is copied)
and records its output for the values below.

Usage: python tests/golden/make_fmt_vectors.py
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
FMT = "/root/reference/extern/spdlog/include"
VALUES = [0.0, -0.0, 0.1, 1.0, 5.0, 10.0, 100.0, 12345.0, 99999.0, 99999.5, 123456.0, 1e6, 1.5e7, -4.615e-3,
          -0.004614974325, 4.519942281538741e-06, -0.07657740583026618, 1e-4, 9.99999e-5, 1e-5, 1.234567891e-7,
          0.30000000000000004, 2.0 / 3.0, -2.5, 0.01, 99.999995, 1e15, 1e16, 1e100, -1.79e308, 5e-324,
          0.123456789012, 3.14159265358979]
STEPS = [0, 7, 100, 9900, 123456789]


def main():
    src = r"""
#define FMT_HEADER_ONLY
#include <spdlog/fmt/bundled/format.h>
#include <cstdio>
#include <cstdlib>
int main(int argc, char **argv) {
    for(int i = 1; i < argc; i++) {
        if(argv[i][0] == 's') { std::printf("%s\n", fmt::format("{:L}", std::atoll(argv[i] + 1)).c_str()); continue; }
        double v = std::strtod(argv[i], nullptr);
        std::printf("%s|%s\n", fmt::format("{:.5}", v).c_str(), fmt::format("{:.8}", v).c_str());
    }
}
"""
    with tempfile.TemporaryDirectory() as tmp:
        cpp, exe = os.path.join(tmp, "f.cpp"), os.path.join(tmp, "f")
        with open(cpp, "w") as f:
            f.write(src)
        subprocess.run(["g++", "-std=c++17", "-O1", "-I" + FMT, cpp, "-o", exe], check=True)
        args = [repr(v) for v in VALUES] + ["s%d" % s for s in STEPS]
        out = subprocess.run([exe] + args, check=True, capture_output=True, text=True).stdout.split("\n")
    doubles = [dict(value=v, p5=out[i].split("|")[0], p8=out[i].split("|")[1]) for i, v in enumerate(VALUES)]
    steps = [dict(step=s, L=out[len(VALUES) + i]) for i, s in enumerate(STEPS)]
    with open(os.path.join(HERE, "fmt_vectors.json"), "w") as f:
        json.dump({"fmt_version": "7.1.3", "doubles": doubles, "steps": steps}, f, indent=1)
    for d in doubles:
        print(d)


if __name__ == "__main__":
    main()
