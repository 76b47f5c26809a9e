"""ParaxialSimulation::save (src/paraxialSimulation.cpp:137-193,286-317,454-511) of the
drop-in classes, read back with HDF5's own h5dump.  Runs on the CPU: the driver
saves user-defined modules with known results from a 3D (FFTSolver3D) and a 2D
(CrankNicholson) simulation that are set up but not propagated.

Checked against the reference's format: group /data with double attributes xmin,
xmax, dx, zmin, zmax, dz, ymin, ymax, dy, wavenumber and the fixed-length string
attributes solver and description; one dataset /data/<module name> per module with
the dataspace of include/hdf5DataspaceCreator.hpp:30-66 (vector {n}, matrix
{n_cols, n_rows}, cube {n_rows, n_cols, n_slices}) over the column-major buffer,
H5T_IEEE_F64LE with the module's attributes (int -> H5T_STD_I32LE), or H5T_STD_U8LE
without attributes for the 8-bit cubes."""
import os
import re
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "dropin", "dropin_driver")
H5LIB = os.path.join(os.path.dirname(HERE), "paxpro_amd", "libpaxpro_h5.so")
H5DUMP = "/opt/conda/bin/h5dump"

pytestmark = pytest.mark.skipif(not (os.path.exists(DRIVER) and os.path.exists(H5LIB) and
                                     os.path.exists(H5DUMP)),
                                reason="drop-in driver, HDF5 writer or h5dump not available")


def dump(path, *args):
    r = subprocess.run([H5DUMP] + list(args) + [path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return r.stdout


def header(path):
    """{object path: text of its block} for the groups' attributes and datasets."""
    text = dump(path, "-H")
    out = {}
    for m in re.finditer(r'(ATTRIBUTE|DATASET) "([^"]+)" \{(.*?)\n\s{0,9}\}', text, re.S):
        out[m.group(2)] = m.group(3)
    return text, out


def data(path, dset, dtype, tmp):
    b = str(tmp / (dset.replace("/", "_") + ".bin"))
    dump(path, "-d", dset, "-b", "LE", "-o", b)
    return np.fromfile(b, dtype=dtype)


def dims(block):
    m = re.search(r"DATASPACE\s+SIMPLE \{ \( ([0-9, ]+) \)", block)
    return tuple(int(v) for v in m.group(1).split(","))


def attr_value(path, name):
    text = dump(path, "-a", name)
    return re.search(r"\(0\): (.*)", text).group(1).strip()


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    tmp = tmp_path_factory.mktemp("h5")
    out = str(tmp / "known.h5")
    r = subprocess.run([DRIVER, "h5known", out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ("knownVector", "knownField2D", "knownMatrix", "knownCube", "knownBytes"):
        assert f"Dataset {name} added to HDF5 file" in r.stderr
    return tmp, out, out + ".2d.h5"


def test_group_attributes(files):
    _, f3, f2 = files
    text, blocks = header(f3)
    assert 'GROUP "data"' in text
    for name, val in (("xmin", -100.0), ("xmax", 100.0), ("dx", 12.5), ("ymin", -100.0), ("ymax", 100.0),
                      ("dy", 12.5), ("zmin", 0.0), ("zmax", 50.0), ("dz", 5.0),
                      ("wavenumber", 2 * np.pi / 0.2)):
        assert "H5T_IEEE_F64LE" in blocks[name] and "SCALAR" in blocks[name], name
        assert float(attr_value(f3, "/data/" + name)) == pytest.approx(val, rel=1e-5), name
    for name, s in (("solver", "FFTSolver3D"), ("description", "known modules")):
        b = blocks[name]
        assert "H5T_C_S1" in b and f"STRSIZE {len(s)};" in b and "H5T_STR_NULLTERM" in b
        assert attr_value(f3, "/data/" + name) == f'"{s}"'
    text2, blocks2 = header(f2)
    assert attr_value(f2, "/data/solver") == '"CrankNicholson"'
    assert "description" not in blocks2   # empty description: no attribute


def test_matrix_dataset_is_written_transposed_with_attributes(files, tmp_path):
    _, f3, _ = files
    _, blocks = header(f3)
    b = blocks["knownMatrix"]
    assert "H5T_IEEE_F64LE" in b and dims(b) == (4, 3)      # {n_cols, n_rows}
    got = data(f3, "/data/knownMatrix", np.float64, tmp_path).reshape(4, 3)
    r, c = np.arange(3)[None, :], np.arange(4)[:, None]
    np.testing.assert_array_equal(got, 10.0 * r + c)        # row = column index of the result
    assert attr_value(f3, "/data/knownMatrix/count") == "7"
    assert float(attr_value(f3, "/data/knownMatrix/scale")) == 2.5
    assert "H5T_STD_I32LE" in dump(f3, "-a", "/data/knownMatrix/count")


def test_cube_datasets_keep_the_column_major_buffer(files, tmp_path):
    _, f3, _ = files
    _, blocks = header(f3)
    b = blocks["knownCube"]
    assert "H5T_IEEE_F64LE" in b and dims(b) == (2, 3, 4)   # {n_rows, n_cols, n_slices}
    got = data(f3, "/data/knownCube", np.float64, tmp_path)
    r, c, s = np.meshgrid(np.arange(2), np.arange(3), np.arange(4), indexing="ij")
    np.testing.assert_array_equal(got, (100.0 * r + 10.0 * c + s).ravel(order="F"))
    assert attr_value(f3, "/data/knownCube/slices") == "4"
    b = blocks["knownBytes"]
    assert "H5T_STD_U8LE" in b and dims(b) == (2, 2, 3) and "ATTRIBUTE" not in b
    np.testing.assert_array_equal(data(f3, "/data/knownBytes", np.uint8, tmp_path), 20 * np.arange(12) + 1)


def test_2d_vector_matrix_and_empty_results(files, tmp_path):
    _, _, f2 = files
    _, blocks = header(f2)
    assert dims(blocks["knownVector"]) == (5,)
    np.testing.assert_array_equal(data(f2, "/data/knownVector", np.float64, tmp_path), 0.5 * np.arange(5) - 1.0)
    assert dims(blocks["knownField2D"]) == (4, 3)
    # a projection quantity of a 2D simulation is a vector; this module has none: empty
    assert dims(blocks["knownMatrix"]) == (0,)
