"""VTU files of the output step (src/cipher.c:1086-1091): leaves written as
pixels/voxels with shared corner points and one Float64 cell array per output
name.  Pinned on the CPU by reading the file back: the cell arrays, the
point count of a uniform forest ((n+1)^dim corners), and every cell's
corner box, whose edge lengths must equal the leaf size h and whose
lower corner the leaf anchor."""
import numpy as np
import pytest

import oracle
from cipher_amd import abi
from cipher_amd import mesh as M
from cipher_amd import vtu
from tests import helpers as H


@pytest.mark.parametrize("dim", [2, 3])
def test_uniform_forest_round_trip(tmp_path, dim):
    f = M.uniform_forest((8,) * dim, [1] * dim, [2.0] * dim, [True] * dim, level=2)
    vals = np.arange(f.ncells, dtype=np.float64) * 0.5
    p = tmp_path / "u.vtu"
    vtu.write(str(p), f, {"phaseid": vals})
    pts, conn, types, data = vtu.read(str(p))
    assert pts.shape[0] == 5 ** dim and conn.shape == (f.ncells, 2 ** dim)
    assert np.all(types == (vtu.VTK_PIXEL if dim == 2 else vtu.VTK_VOXEL))
    assert np.array_equal(data["phaseid"], vals)


def test_amr_forest_geometry_and_fields(tmp_path):
    wl = H.gg2d_amr()
    f = wl.forest
    X = H.initial_state(wl)
    out = oracle.output(wl.cfg, wl.mesh, X, [(abi.OUT_PHASEID, 0), (abi.OUT_TEMPERATURE, 0)])
    p = tmp_path / "amr.vtu"
    vtu.write(str(p), f, {"phaseid": out[:, 0], "temperature": out[:, 1]})
    pts, conn, types, data = vtu.read(str(p))
    assert np.array_equal(data["phaseid"], out[:, 0]) and np.array_equal(data["temperature"], out[:, 1])
    corners = pts[conn][:, :, :2]
    lo, hi = corners.min(axis=1), corners.max(axis=1)
    h = f.cellgeom()
    np.testing.assert_allclose(hi - lo, np.stack([h, h], axis=1), rtol=1e-12)
    np.testing.assert_allclose(lo, f.anchor * (f.size / f.res)[None, :], rtol=1e-12, atol=1e-12)
    # pixel corner order: x fastest, then y
    np.testing.assert_allclose(corners[:, 1] - corners[:, 0], np.stack([h, 0 * h], axis=1), atol=1e-12)
    np.testing.assert_allclose(corners[:, 2] - corners[:, 0], np.stack([0 * h, h], axis=1), atol=1e-12)
    with pytest.raises(ValueError):
        vtu.write(str(p), f, {"bad": out[:-1, 0]})
