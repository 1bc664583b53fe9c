"""The reference's known answers for the mechanical step on the GPU
(set-ups in tests/mech_cases.py), with the oracle run beside it."""
import sys

import numpy as np
import pytest

import mech_cases
import oracle
from exagoop_amd import capi

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(a, b):
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def test_axial_bar_validator_and_parity(gpu):
    say("[mech] axial bar, five periods")
    s, pos, r, idata = mech_cases.axial_bar()
    runs = []
    for ctx in (capi.Context(s), oracle.Oracle(s, threads=1)):
        ctx.upload(pos, r, idata)
        ctx.init_internal_forces()
        half = mech_cases.run_bar(ctx, 47.5)
        runs.append((max(half, mech_cases.run_bar(ctx)), ctx))
    (rms_g, g), (rms_o, o) = runs
    assert rms_g < 0.1 and rms_g < 2e-3
    assert g.time()[1] == o.time()[1]
    pg, rg, _, _ = g.download()
    po, ro, _, _ = o.download()
    assert rel(pg, po) <= 1e-10
    assert rel(rg[capi.R_VEL], ro[capi.R_VEL]) <= 1e-8


def test_dam_break_front_follows_martin_and_moyce(gpu):
    say("[mech] Martin & Moyce dam break")
    s, pos, r, idata = mech_cases.martin_moyce_dam()
    g = capi.Context(s)
    g.upload(pos, r, idata)
    g.init_internal_forces()
    front = mech_cases.run_front(g)
    mech_cases.check_front(front)
    o = oracle.Oracle(s, threads=8)
    o.upload(pos, r, idata)
    o.init_internal_forces()
    g2 = capi.Context(s)
    g2.upload(pos, r, idata)
    g2.init_internal_forces()
    for c in (g2, o):
        c.step(200)
    pg, rg, _, _ = g2.download()
    po, ro, _, _ = o.download()
    assert rel(pg, po) <= 1e-10
    assert rel(rg[capi.R_STRESS:capi.R_STRESS + 6], ro[capi.R_STRESS:capi.R_STRESS + 6]) <= 1e-10
