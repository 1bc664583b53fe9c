"""The reference's heat-conduction validation cases on the GPU
(Tests/1D_Heat_Conduction, 1D_Heat_Conduction_Convective,
1D_Heat_Conduction_HeatFlux, 2D_Heat_Conduction): the particle temperatures
at the validators' output time are within their RMS tolerance of the
analytic solutions (set-up and solutions in tests/heat_cases.py).  The 1-D
case is also run on the oracle: the two agree to 1e-10 after its 5000
steps.
"""
import sys

import numpy as np
import pytest

import heat_cases
import oracle
from exagoop_amd import capi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", list(heat_cases.CASES))
def test_reference_heat_validator(gpu, name):
    say(f"[heat] {name}: {heat_cases.CASES[name]['steps']} steps")
    rms, g = heat_cases.run(capi.Context, name)
    say(f"[heat] {name}: RMS {rms:.3e} (tolerance {heat_cases.CASES[name]['tol']:.0e})")
    assert g.time()[1] == heat_cases.CASES[name]["steps"]
    assert rms < heat_cases.CASES[name]["tol"], rms


def test_1d_heat_case_matches_the_oracle(gpu):
    say("[heat] 1D_Heat_Conduction GPU vs oracle")
    _, g = heat_cases.run(capi.Context, "1D_Heat_Conduction")
    _, o = heat_cases.run(lambda s: oracle.Oracle(s, threads=8), "1D_Heat_Conduction")
    tg, to = g.download_thermal(), o.download_thermal()
    # The flux across the bar (y, z) is zero up to rounding: the flux
    # components are measured against the flux vector's norm.
    flux = slice(capi.TH_HEAT_FLUX, capi.TH_HEAT_FLUX + 3)
    qn = np.linalg.norm(to[flux])
    for f in range(capi.NTHERMAL):
        nb = qn if capi.TH_HEAT_FLUX <= f < capi.TH_HEAT_FLUX + 3 else np.linalg.norm(to[f])
        err = np.linalg.norm(tg[f] - to[f]) / (nb if nb > 0 else 1.0)
        assert err <= 1e-10, (f, err)
