"""EVP-FFT microstructure text files (material.py:1035-1091, importers.py:153-200):
the reference's fixtures and expected lines (tests/test_importers.py:17-34,
248-280; tests/test_material.py:611-632; tests/small_model_fields.txt)."""

import numpy as np
import pandas as pd
from numpy.testing import assert_allclose, assert_array_equal

from materialite_amd import Material, Orientation, Sphere, import_evpfft

# rows of the reference's EVP-FFT import fixture (tests/test_importers.py:17-34)
IMPORT_ROWS = [
    [99.80835, 66.40345, 28.98868, 1, 1, 1, 3, 1],
    [99.80835, 66.40345, 28.98868, 2, 1, 1, 3, 1],
    [99.80835, 66.40345, 28.98868, 1, 2, 1, 3, 1],
    [99.80835, 66.40345, 28.98868, 2, 2, 1, 3, 1],
    [-157.91174, 115.84247, -38.84106, 1, 1, 2, 5, 1],
    [-157.91174, 115.84247, -38.84106, 2, 1, 2, 5, 1],
    [149.57114, 93.98035, 66.88121, 1, 2, 2, 2, 1],
    [149.57114, 93.98035, 66.88121, 2, 2, 2, 2, 1],
    [124.95919, 77.12344, 3.69264, 1, 1, 3, 1, 1],
    [14.56741, 69.05068, 44.11161, 2, 1, 3, 9, 1],
    [124.95919, 77.12344, 3.69264, 1, 2, 3, 1, 1],
    [36.92948, 123.74353, 9.99989, 2, 2, 3, 7, 1],
    [124.95919, 77.12344, 3.69264, 1, 1, 4, 1, 1],
    [14.56741, 69.05068, 44.11161, 2, 1, 4, 9, 1],
    [149.27193, 113.16453, -52.42448, 1, 2, 4, 4, 1],
    [149.27193, 113.16453, -52.42448, 2, 2, 4, 4, 1],
]

# tests/small_model_fields.txt of the reference (a 2^3 EVP-FFT input)
SMALL_MODEL = """30.0 45.0 60.0 1 1 1 2 1
30.0 45.0 60.0 2 1 1 2 1
30.0 45.0 60.0 1 2 1 2 1
0.0 0.0 0.0 2 2 1 1 1
30.0 45.0 60.0 1 1 2 2 1
0.0 0.0 0.0 2 1 2 1 1
0.0 0.0 0.0 1 2 2 1 1
0.0 0.0 0.0 2 2 2 1 1
"""


def test_import_evpfft(tmp_path):
    f = tmp_path / "fields.txt"
    f.write_text("\n".join(" ".join(map(str, r)) for r in IMPORT_ROWS))
    material = import_evpfft(f)
    expected = pd.DataFrame(np.array(IMPORT_ROWS), columns=[
        "e1", "e2", "e3", "x_id", "y_id", "z_id", "grain", "phase"]).sort_values(
        by=["x_id", "y_id", "z_id"])
    assert list(material.dimensions) == [2, 2, 4]
    assert_allclose(material.extract("orientation").euler_angles_in_degrees,
                    expected[["e1", "e2", "e3"]].to_numpy())
    assert_array_equal(material.extract(["x_id", "y_id", "z_id"]) + 1,
                       expected[["x_id", "y_id", "z_id"]].to_numpy().astype(int))
    assert_array_equal(material.extract(["grain", "phase"]),
                       expected[["grain", "phase"]].to_numpy().astype(int))


def test_export_small_model_first_line(tmp_path):
    # tests/test_material.py:611-632
    material = (Material(dimensions=[2, 2, 2])
                .create_uniform_field("orientation", Orientation.from_euler_angles([0, 0, 0]))
                .create_uniform_field("grain", 1)
                .create_uniform_field("phase", 1)
                .insert_feature(Sphere(radius=1, centroid=[0, 0, 0]), fields={
                    "grain": 2,
                    "orientation": Orientation.from_euler_angles([30, 45, 60], in_degrees=True)}))
    out = tmp_path / "small_model_fields.txt"
    material.export_to_evpfft(output=out, euler_angles_to_degrees=True)
    lines = out.read_text().splitlines()
    assert lines[0] == "30.00000 45.00000 60.00000 1 1 1 2 1"
    # the whole file is the reference's small-model input, in the exporter's format
    ref = [list(map(float, l.split())) for l in SMALL_MODEL.strip().splitlines()]
    got = [list(map(float, l.split())) for l in lines]
    assert_allclose(got, ref, atol=1e-12)


def test_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    material = (Material(dimensions=[4, 2, 8]).create_voronoi(num_regions=6, label="grain", rng=rng)
                .assign_random_orientations(region_label="grain", rng=rng)
                .create_uniform_field("phase", 1))
    out = tmp_path / "f.txt"
    material.export_to_evpfft(output=out)
    back = import_evpfft(out)
    assert list(back.dimensions) == [4, 2, 8]
    assert_array_equal(back.extract("grain"), material.extract("grain"))
    # 5 decimals of degrees: rotation matrices agree to ~1e-7
    assert_allclose(back.extract("orientation").rotation_matrix,
                    material.extract("orientation").rotation_matrix, atol=2e-7)


def test_degenerate_euler_angles():
    # R22 = +1 and -1 (tensor.py:1937-1956)
    o = Orientation.from_euler_angles([[0.3, 0.0, 0.0], [0.4, np.pi, 0.0]])
    e = o.euler_angles
    assert_allclose(e[0], [0.3, 0.0, 0.0], atol=1e-12)
    assert_allclose(e[1], [0.4, np.pi, 0.0], atol=1e-12)
