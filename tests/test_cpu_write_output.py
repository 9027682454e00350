"""CSV and VTU output (vfvm_write_csv / vfvm_write_vtu, restating src/base/write_csv.rs:5-100
and src/base/write_vtu.rs) on the CPU: the
reference's fixed 9-decimal format checked against Python's correctly rounded formatting of
the same doubles, the full-precision mode reading back bit for bit, vertex and boundary-node
order, vector fields and the WriteCsv error."""
import numpy as np
import pytest

import vfvm


def ref_fmt(v):
    # Rust's "{:.9}": fixed point, 9 decimals, exact value rounded half to even
    if np.isnan(v):
        return "NaN"
    if np.isinf(v):
        return "inf" if v > 0 else "-inf"
    return f"{v:.9f}"


def test_scalar_domain_reference_format_and_order(tmp_path):
    m = vfvm.Mesh.lattice(5, 0, 0.3)
    pos = m.get("pos")
    rng = np.random.default_rng(1)
    vals = rng.standard_normal(m.num_tot) * 10.0 ** rng.integers(-12, 6, m.num_tot)
    vals[:3] = [np.nan, np.inf, -0.0]
    path = m.write_csv(vals, str(tmp_path / "temp.csv"), ts=7)
    assert path.endswith("temp_0007.csv")
    lines = open(path).read().splitlines()
    assert lines[0] == "x,y,z,value" and len(lines) == 1 + m.num_tot
    for v in range(m.num_tot):  # points, then nodes
        assert lines[1 + v] == ",".join(ref_fmt(a) for a in (*pos[v], vals[v]))


def test_full_precision_round_trip(tmp_path):
    m = vfvm.Mesh.lattice(4, 3, 0.3)
    vals = np.random.default_rng(2).standard_normal((3, m.num_tot)) * 1e-7
    path = m.write_csv(vals, str(tmp_path / "vel.csv"), ts=0, digits=0)
    data = np.loadtxt(path, delimiter=",", skiprows=1)
    assert open(path).readline().strip() == "x,y,z,value_x,value_y,value_z"
    assert np.array_equal(data[:, :3], m.get("pos")) and np.array_equal(data[:, 3:].T, vals)


def test_boundary_nodes_and_errors(tmp_path):
    m = vfvm.Mesh.lattice(4, 0, 0.3)
    b = 2
    nodes = m.num_pts + m.bnd_nodes(b)
    vals = np.arange(len(nodes), dtype=float) * 0.5
    path = m.write_csv(vals, str(tmp_path / "flux.txt"), ts=12, bnd=b)
    data = np.loadtxt(path, delimiter=",", skiprows=1)
    assert np.abs(data[:, :3] - m.get("pos")[nodes]).max() <= 5e-10 and np.array_equal(data[:, 3], vals)
    with pytest.raises(vfvm.VFVChemError) as e:
        m.write_csv(vals, str(tmp_path / "missing_dir" / "x.csv"), bnd=b)
    assert e.value.code == 50
    with pytest.raises(ValueError):
        m.write_csv(vals, str(tmp_path / "z.csv"))  # boundary-length values for the whole domain
    with pytest.raises(vfvm.VFVChemError) as e:
        m.write_csv(vals, str(tmp_path / "y.csv"), bnd=9)
    assert e.value.code == 1


def _vtu_arrays(path):
    import xml.etree.ElementTree as ET
    root = ET.parse(path).getroot()
    piece = root.find("UnstructuredGrid/Piece")
    arr = {}
    for da in piece.iter("DataArray"):
        arr[da.get("Name") or "points"] = np.array(da.text.split(), float)
    return piece, arr


def test_vtu_domain_and_boundary(tmp_path):
    """VTU layout of write_vtu.rs: points, the tetrahedra (type 10) or one vertex cell per
    boundary node (type 1), point data "value" in the reference's 9-decimal format."""
    m = vfvm.Mesh.lattice(4, 0, 0.3)
    vals = np.random.default_rng(4).standard_normal((3, m.num_tot))
    path = m.write_csv(vals, str(tmp_path / "vel.vtu"), ts=3)
    assert path.endswith("vel_0003.vtu")
    piece, a = _vtu_arrays(path)
    tets = m.get("tets")
    assert int(piece.get("NumberOfPoints")) == m.num_tot and int(piece.get("NumberOfCells")) == len(tets)
    assert np.abs(a["points"].reshape(-1, 3) - m.get("pos")).max() <= 5e-10
    assert np.array_equal(a["connectivity"].reshape(-1, 4), tets)
    assert np.array_equal(a["offsets"], 4 * np.arange(1, len(tets) + 1)) and np.all(a["types"] == 10)
    assert np.abs(a["value"].reshape(-1, 3) - vals.T).max() <= 5e-10
    nodes = m.num_pts + m.bnd_nodes(4)
    bv = np.linspace(0.0, 1.0, len(nodes))
    piece, a = _vtu_arrays(m.write_csv(bv, str(tmp_path / "qb.vtu"), bnd=4, digits=0))
    assert int(piece.get("NumberOfCells")) == len(nodes) and np.all(a["types"] == 1)
    assert np.array_equal(a["connectivity"], np.arange(len(nodes))) and np.array_equal(a["value"], bv)
    assert np.array_equal(a["points"].reshape(-1, 3), m.get("pos")[nodes])


def test_huge_values_are_written_in_full(tmp_path):
    """Fixed-point expansions longer than any small buffer (1e100, -1e300 with 9 decimals) are
    written whole, as Rust's "{:.9}" writes them; the full-precision mode reads them back."""
    m = vfvm.Mesh.lattice(3, 0, 0.3)
    vals = np.zeros(m.num_tot)
    vals[0], vals[1], vals[2] = 1e100, -1e300, np.finfo(np.float64).max
    path = m.write_csv(vals, str(tmp_path / "big.csv"), ts=0)
    lines = open(path).read().splitlines()
    for v in range(3):
        assert lines[1 + v].split(",")[3] == ref_fmt(vals[v])
    assert len(lines[2].split(",")[3]) == len(f"{-1e300:.9f}") == 312
    path = m.write_csv(vals, str(tmp_path / "bigfull.csv"), ts=0, digits=0)
    data = np.loadtxt(path, delimiter=",", skiprows=1)
    assert np.array_equal(data[:, 3], vals)
