"""INTEGRATION.md's reference-side binding (section 2: the EulerHIP plugin over
the C ABI) compiles against the reference's own headers (src/integrators/
Integrator.h and what it includes) and include/ch_hip.h, instantiated for 2D
so every member is checked.  CPU only; skipped where the reference sources
are absent (they exist in the development container only)."""
import os
import re
import subprocess

import pytest

from conftest import REPO

REF = "/root/reference"


@pytest.mark.skipif(not os.path.isdir(os.path.join(REF, "src", "integrators")), reason="reference sources absent")
def test_integration_binding_compiles(tmp_path):
    text = open(os.path.join(REPO, "INTEGRATION.md")).read()
    sec = text[text.index("## 2."):text.index("## 3.")]
    code = re.findall(r"```cpp\n(.*?)```", sec, re.S)[0]
    src = tmp_path / "binding.cpp"
    src.write_text(code + "\ntemplate struct ch::EulerHIP<2>;\ntemplate class ch::HIPIntegrator<3>;\n")
    # the snippet's "../integrators/Integrator.h" resolves from any sibling of
    # src/integrators on the include path (its own home would be src/HIP)
    # NDEBUG as the reference's release build (its MultiField asserts do not compile without it)
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-DNDEBUG", "-DNOCUDA", "-DTOML_EXCEPTIONS=0", "-DTOML_ENABLE_FORMATTERS=0",
           "-I", os.path.join(REF, "src", "models"), "-I", os.path.join(REF, "src"), "-I", os.path.join(REF, "extern"),
           "-I", os.path.join(REF, "extern", "spdlog", "include"), "-I", os.path.join(REPO, "include"), str(src)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
