"""The reference's known answers for the mechanical step on the oracle
(set-ups in tests/mech_cases.py): the axial bar's first mode after five
periods, and the dam-break front against Martin & Moyce's data."""
import pytest

import mech_cases

oracle = pytest.importorskip("oracle")


def test_axial_bar_validator():
    s, pos, r, idata = mech_cases.axial_bar()
    o = oracle.Oracle(s, threads=1)
    o.upload(pos, r, idata)
    o.init_internal_forces()
    # half a period apart: a lost fixed end shows at one of the two
    half = mech_cases.run_bar(o, 47.5)
    rms = mech_cases.run_bar(o)
    assert rms < 0.1  # the reference's tolerance
    assert max(half, rms) < 2e-3  # 5e-4 measured


def test_dam_break_front_follows_martin_and_moyce():
    s, pos, r, idata = mech_cases.martin_moyce_dam()
    o = oracle.Oracle(s, threads=8)
    o.upload(pos, r, idata)
    o.init_internal_forces()
    mech_cases.check_front(mech_cases.run_front(o))
