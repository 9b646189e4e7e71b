"""A reference-format XML deck run end to end on the device (the case reader,
insertion, the reference's driver loop, the RawData writer) against the same
deck run through the FP64 oracle: every RawData line agrees to the written
precision (6 digits) but for the last digit of values that straddle a
rounding boundary, and the final states to 1e-9 of the domain (positions)
and 1e-6 relative (velocities): spheres, boxes and a superquadric, GJK gap
1e-12 on both sides (the superquadric support's pow differs by ulps)."""
import os

import numpy as np
import pytest

from grains3dgpu_amd import case as C
from tests.test_case_cpu import DECK, POS

pytestmark = pytest.mark.gpu


def test_case_on_device_matches_oracle(tmp_path):
    from grains3dgpu_amd import engine
    from oracle import pyoracle
    outs = {}
    for name in ("device", "oracle"):
        d = tmp_path / name
        d.mkdir()
        (d / "deck.xml").write_text(DECK)
        (d / "pos.dat").write_text("\n".join(" ".join(repr(v) for v in p) for p in POS) + "\n")
        c = C.read_case(str(d / "deck.xml"))
        b = engine.Engine(dt=c.dt, precision=c.precision, gjk_rel_tol=1e-12) if name == "device" else \
            pyoracle.Oracle(dt=c.dt, gjk_rel_tol=1e-12)
        assert C.run_case(c, b) == 21
        outs[name] = (b.get_state(), d / "out")
        if name == "device":
            b.close()
    (pe, _, ve, _), de = outs["device"]
    (po, _, vo, _), do = outs["oracle"]
    assert np.abs(pe - po).max() < 1e-9 * 0.024
    assert np.abs(ve - vo).max() <= 1e-6 * np.abs(vo).max()
    for f in sorted(os.listdir(do)):
        a = open(os.path.join(de, f)).read().splitlines()
        b = open(os.path.join(do, f)).read().splitlines()
        assert len(a) == len(b), f
        for la, lb in zip(a, b):
            xa, xb = np.array(la.split(), float), np.array(lb.split(), float)
            np.testing.assert_allclose(xa, xb, rtol=2e-6, atol=1e-12, err_msg=f)
