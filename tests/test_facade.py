"""The C++ facade (MPMParticleContainer method surface) compiles against the C
ABI header and links against libexagoop_amd.so; without a device every call
reports MPM_E_NODEVICE through the facade's exception instead of aborting."""
import os
import shutil
import subprocess
import tempfile

import pytest

from exagoop_amd import capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROG = r"""
#include <cstdio>
#include "mpm_container.hpp"
int main() {
    mpm_specs s;
    mpm_specs_default(&s);
    s.ncells[0] = 8; s.ncells[1] = 8; s.ncells[2] = 8;
    for (int d = 0; d < 3; ++d) { s.plo[d] = 0.0; s.phi[d] = 1.0; }
    try {
        exagoop_amd::MPMParticleContainer pc(s);
        double pos[3] = {0.5, 0.5, 0.5};
        double r[MPM_NREAL] = {0};
        int id[MPM_NINT] = {0};
        r[MPM_R_MASS] = 1.0; r[MPM_R_VOLUME] = r[MPM_R_VOLINIT] = 1e-3;
        r[MPM_R_E] = 1e3; r[MPM_R_NU] = 0.3; r[MPM_R_DENSITY] = 1e3; r[MPM_R_JACOBIAN] = 1.0;
        r[MPM_R_DEFGRAD] = r[MPM_R_DEFGRAD + 4] = r[MPM_R_DEFGRAD + 8] = 1.0;
        pc.InitParticles(1, pos, r, id);
        pc.Initialise_Internal_Forces();
        pc.Step(2);
        std::printf("ran %lld\n", (long long)pc.TotalNumberOfParticles());
    } catch (const std::runtime_error &e) {
        std::printf("error %s\n", e.what());
    }
    return 0;
}
"""


def test_facade_compiles_links_and_reports_errors():
    capi.load()
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no g++")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "f.cpp")
        exe = os.path.join(d, "f")
        open(src, "w").write(PROG)
        libdir = os.path.join(ROOT, "exagoop_amd")
        r = subprocess.run([gxx, "-std=c++17", "-I" + os.path.join(ROOT, "include"),
                            "-I" + os.path.join(ROOT, "exagoop_amd", "csrc"), src, "-o", exe,
                            "-L" + libdir, "-lexagoop_amd", "-Wl,-rpath," + libdir],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stderr
        line = out.stdout.strip()
        assert line == "ran 1" or (line.startswith("error mpm_create failed") and
                                   "(-9)" in line), line
