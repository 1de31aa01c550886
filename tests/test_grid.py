"""setGrid's constants as the package forms them (grid_constants, used by
Grid) against the oracle's setGrid restatement (src/Grid.cpp:5-62), bit
for bit, on grids whose spacings are and are not exact binary fractions:
an inexact spacing tells a division done in double and then narrowed (the
reference's) from one done in FPfield."""
from __future__ import annotations

import numpy as np
import pytest

import oracle as O
from sputnipic_amd.core import grid_constants

GRIDS = [(32, 16, 16), (8, 6, 4), (12, 8, 8), (16, 4, 10), (80, 40, 40), (96, 48, 48),
         (128, 64, 64), (7, 3, 5)]


@pytest.mark.parametrize("prec", ["f64", "f32"])
@pytest.mark.parametrize("cells", GRIDS)
def test_grid_constants_match_oracle(cells, prec):
    cfg = O.GemConfig(nxc=cells[0], nyc=cells[1], nzc=cells[2])
    st = O.State(cfg, prec, build_ic=False)
    k = grid_constants(*cells, cfg.Lx, cfg.Ly, cfg.Lz, prec)
    g = st.grid
    for name in ("invdx", "invdy", "invdz", "invVOL"):
        assert k[name] == getattr(g, name), (name, k[name], getattr(g, name))
    for name in ("dx", "dy", "dz"):
        assert k[name] == getattr(g, name), name
    shape = (g.nxn, g.nyn, g.nzn)
    assert np.array_equal(k["xn"], st.XN.reshape(shape)[:, 0, 0])
    assert np.array_equal(k["yn"], st.YN.reshape(shape)[0, :, 0])
    assert np.array_equal(k["zn"], st.ZN.reshape(shape)[0, 0, :])


def test_fp32_inverse_volume_divides_in_double():
    """8x6x4 cells on 20x10x10: dx * dy * dz = 10.41666.. is not a float;
    the FP32 invVOL is float(1 / 10.41666..), not 1.0f / float(10.41666..)."""
    k = grid_constants(8, 6, 4, 20.0, 10.0, 10.0, "f32")
    vol = 20.0 / 8 * (10.0 / 6) * (10.0 / 4)
    assert k["invVOL"] == float(np.float32(1.0 / vol))
    assert k["invVOL"] != float(np.float32(1.0) / np.float32(vol))
