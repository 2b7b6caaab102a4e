"""Particle file formats (aboria_amd/io.py): the reference's text stream
(operator<< / operator>>, src/Particles.h:553-571 with src/Traits.h:433-455
and src/Vector.h:746-765) and its VTK unstructured grid
(copy_to_vtk_grid / copy_from_vtk_grid, src/Particles.h:586-684).

Known answers: the header and lines a default std::ostream prints (6
significant digits, vectors as "(x,y,z)", values separated by ", ").
Round trips: exact at precision 17 and for Float64 VTK arrays; Float32 VTK
arrays equal the values rounded to FP32 (the reference's vtkFloatArray);
one VTK_VERTEX cell per particle, points padded to 3 components."""
import io as pyio
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from aboria_amd import io as pio


def test_stream_known_answer():
    buf = pyio.StringIO()
    pos = np.array([[0.1, 0.2, 0.1 + 0.2], [1.0, -2.5, 1e-7]])
    pio.write_stream(buf, pos, id=[7, 3], alive=[1, 1],
                     columns={"scalar1": np.array([1.0, 2.0 / 3.0])})
    assert buf.getvalue().splitlines() == [
        "position, id, alive, scalar1",
        "(0.1,0.2,0.3), 7, 1, 1",
        "(1,-2.5,1e-07), 3, 1, 0.666667",
    ]


def test_stream_round_trip_full_precision():
    rng = np.random.default_rng(3)
    pos = rng.random((50, 2))
    cols = {"a": rng.standard_normal(50), "v": rng.standard_normal((50, 3))}
    buf = pyio.StringIO()
    pio.write_stream(buf, pos, id=np.arange(50) * 3, columns=cols, precision=17)
    buf.seek(0)
    names, out = pio.read_stream(buf)
    assert names == ["position", "id", "alive", "a", "v"]
    assert np.array_equal(out["position"], pos)
    assert np.array_equal(out["id"], np.arange(50) * 3)
    assert np.array_equal(out["a"], cols["a"])
    assert np.array_equal(out["v"], cols["v"])


@pytest.mark.parametrize("binary", [True, False])
@pytest.mark.parametrize("float64", [True, False])
def test_vtu_round_trip(tmp_path, binary, float64):
    rng = np.random.default_rng(4)
    n = 200
    pos = rng.random((n, 2))
    cols = {"density": rng.random(n), "velocity": rng.standard_normal((n, 3))}
    f = str(tmp_path / "p.vtu")
    pio.write_vtu(f, pos, id=np.arange(n)[::-1], columns=cols, float64=float64, binary=binary)
    root = ET.parse(f).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    assert piece.get("NumberOfPoints") == str(n) and piece.get("NumberOfCells") == str(n)
    out = pio.read_vtu(f, dimension=2)
    assert np.array_equal(out["position"], pos)
    assert np.array_equal(out["id"], np.arange(n)[::-1])
    assert np.array_equal(out["alive"], np.ones(n, np.uint8))
    for k, v in cols.items():
        want = v if float64 else v.astype(np.float32)
        assert np.array_equal(out[k], want)
    pts = pio._read_array(piece.find("Points/DataArray"))
    assert pts.shape == (n, 3) and np.all(pts[:, 2] == 0.0)
