"""Output file formats (CPU): the sparse Tecplot writer restated from the
reference (src/funlbm/file/tecplot/dump.py:13-22) on the reference's own
example matrix, and the Tecplot POINT and legacy VTK files read back by an
independent parser."""

import numpy as np

from funlbm_amd.lbm import output


def test_sparse_tecplot_of_the_reference_example(tmp_path):
    # the example of src/funlbm/file/tecplot/dump.py:25-28
    dense = np.array([[[0, 1, 0], [2, 0, 3]], [[0, 4, 0], [5, 0, 6]]])
    path = output.tecplot_sparse(dense, str(tmp_path / "sparse.dat"))
    want = ('TITLE = "Sparse Matrix Data"\n'
            'VARIABLES = "Z", "Y", "X", "Value"\n'
            "ZONE I=3, J=2, K=2, F=POINT\n"
            "0 0 1 1\n0 1 0 2\n0 1 2 3\n1 0 1 4\n1 1 0 5\n1 1 2 6\n")
    assert open(path).read() == want
    # float data: indices and values stacked as floats, as numpy prints them
    path = output.tecplot_sparse(np.array([[[0.0, 0.5]]]), str(tmp_path / "f.dat"))
    assert open(path).read().splitlines()[3] == "0.0 0.0 1.0 0.5"


def _fields(shape=(4, 3, 5)):
    rng = np.random.default_rng(0)
    u = rng.standard_normal(shape + (3,)).astype(np.float32)
    rho = (1 + 0.01 * rng.standard_normal(shape)).astype(np.float32)
    return u, rho


def test_tecplot_points_round_trip(tmp_path):
    u, rho = _fields()
    path = output.tecplot_points([u, rho], ["ux", "uy", "uz", "rho"], str(tmp_path / "t.dat"), "flow_u")
    lines = open(path).read().splitlines()
    assert lines[0] == 'TITLE = "flow_u"'
    assert lines[1] == 'VARIABLES = "x", "y", "z", "ux", "uy", "uz", "rho"'
    assert lines[2] == "ZONE I=4, J=3, K=5, F=POINT"
    data = np.loadtxt(path, skiprows=3)
    assert data.shape == (4 * 3 * 5, 7)
    # I (x) fastest, then J, then K
    assert data[1, :3].tolist() == [1, 0, 0] and data[4, :3].tolist() == [0, 1, 0]
    x, y, z = data[:, 0].astype(int), data[:, 1].astype(int), data[:, 2].astype(int)
    assert np.array_equal(data[:, 3:6].astype(np.float32), u[x, y, z])
    assert np.array_equal(data[:, 6].astype(np.float32), rho[x, y, z])


def _read_legacy_vtk(path):
    """Minimal independent reader: header lines, then blocks by keyword."""
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        end = raw.index(b"\n", pos)
        s = raw[pos:end].decode()
        pos = end + 1
        return s

    assert line().startswith("# vtk DataFile Version")
    line()
    mode = line()
    meta, blocks = {}, {}

    def floats(n):
        nonlocal pos
        if mode == "BINARY":
            a = np.frombuffer(raw, dtype=">f4", count=n, offset=pos).astype(np.float32)
            pos += 4 * n + 1
            return a
        vals = []
        while len(vals) < n:
            vals += [float(v) for v in line().split()]
        return np.array(vals, np.float32)

    while pos < len(raw):
        words = line().split()
        if not words:
            continue
        key = words[0]
        if key in ("DATASET", "DIMENSIONS", "ORIGIN", "SPACING", "POINT_DATA"):
            meta[key] = words[1:]
        elif key == "VECTORS":
            n = int(meta["POINT_DATA"][0])
            blocks[words[1]] = floats(3 * n).reshape(n, 3)
        elif key == "SCALARS":
            line()  # LOOKUP_TABLE
            blocks[words[1]] = floats(int(meta["POINT_DATA"][0]))
        elif key == "POINTS":
            n = int(words[1])
            blocks["POINTS"] = floats(3 * n).reshape(n, 3)
        elif key == "VERTICES":
            n = int(words[1])
            if mode == "BINARY":
                blocks["VERTICES"] = np.frombuffer(raw, dtype=">i4", count=2 * n, offset=pos)
                pos += 8 * n + 1
            else:
                blocks["VERTICES"] = np.array([int(v) for _ in range(n) for v in line().split()])
    return mode, meta, blocks


def test_vtk_grid_binary_and_ascii(tmp_path):
    u, rho = _fields()
    for binary in (True, False):
        path = output.vtk_grid(str(tmp_path / f"g{binary}.vtk"), {"u": u}, {"rho": rho}, binary=binary)
        mode, meta, blocks = _read_legacy_vtk(path)
        assert mode == ("BINARY" if binary else "ASCII")
        assert meta["DATASET"] == ["STRUCTURED_POINTS"] and meta["DIMENSIONS"] == ["4", "3", "5"]
        # point (i, j, k) is entry i + 4 (j + 3 k)
        i, j, k = 3, 1, 4
        assert np.array_equal(blocks["u"][i + 4 * (j + 3 * k)], u[i, j, k])
        assert blocks["rho"][i + 4 * (j + 3 * k)] == rho[i, j, k]
        assert np.array_equal(blocks["u"].reshape(5, 3, 4, 3).transpose(2, 1, 0, 3), u)


def test_vtk_points(tmp_path):
    pts = np.random.default_rng(1).uniform(0, 10, size=(7, 3)).astype(np.float32)
    val = np.arange(7, dtype=np.float32) * 0.5
    for binary in (True, False):
        path = output.vtk_points(str(tmp_path / f"p{binary}.vtk"), pts, {"u": val}, binary=binary)
        mode, meta, blocks = _read_legacy_vtk(path)
        assert np.array_equal(blocks["POINTS"], pts)
        assert np.array_equal(blocks["u"], val)
        assert blocks["VERTICES"].reshape(7, 2).tolist() == [[1, i] for i in range(7)]
