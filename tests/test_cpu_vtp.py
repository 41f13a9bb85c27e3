"""VTK XML point-cloud reading and writing of the command-line tool (no GPU).

The reader must extract what the reference's vtp_point_cloud_reader extracts
(src/rti/io/vtp_point_cloud_reader.hpp:23-95) from the data-array encodings a
vtkXMLPolyDataWriter produces: ascii; inline binary (base64, header and data
in one or in separate segments); appended base64 or raw; UInt32 or UInt64
headers; uncompressed or zlib blocks. The files here are written by a small
generator that follows the VTK XML file-format layout.
"""
import os
import struct
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from tests.vtk_files import write_vtk_cloud

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBE = os.path.join(ROOT, "tests", "cpp", "vtp_probe")


@pytest.fixture(scope="module", autouse=True)
def probe():
    if not os.path.exists(PROBE):
        subprocess.run(["bash", os.path.join(ROOT, "tests", "cpp", "build.sh")], check=True)


def run_probe(path, out=None):
    args = [PROBE, path] + ([out] if out else [])
    r = subprocess.run(args, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    n = int(lines[0])
    data = np.array([[float(x) for x in ln.split()] for ln in lines[1:]], dtype=np.float64).reshape(n, 7)
    return data


def cloud(n=57, seed=0, unit=True):
    rng = np.random.default_rng(seed)
    pts = rng.normal(size=(n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    if unit:
        nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    rad = (0.01 + rng.random(n) * 0.05).astype(np.float32)
    return pts, nrm, rad


CASES = [
    dict(fmt="ascii"),
    dict(fmt="binary", header="UInt32", joined=True),
    dict(fmt="binary", header="UInt32"),
    dict(fmt="binary", header="UInt64"),
    dict(fmt="binary", header="UInt64", compress=True),
    dict(fmt="binary", header="UInt32", compress=True),
    dict(fmt="appended", header="UInt32", encoding="base64"),
    dict(fmt="appended", header="UInt64", encoding="base64", compress=True),
    dict(fmt="appended", header="UInt64", encoding="raw"),
    dict(fmt="appended", header="UInt32", encoding="raw", compress=True),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(f"{k}={v}" for k, v in c.items()))
def test_reader_encodings(tmp_path, case):
    pts, nrm, rad = cloud()
    path = str(tmp_path / "in.vtp")
    write_vtk_cloud(path, pts, nrm, rad, **case)
    got = run_probe(path).astype(np.float32)  # printed with 9 digits: exact float32
    np.testing.assert_array_equal(got[:, 0:3], pts)
    np.testing.assert_array_equal(got[:, 3], rad)
    np.testing.assert_allclose(got[:, 4:7], nrm, atol=1e-7)


def test_sqrt_of_area_and_normalisation(tmp_path):
    pts, nrm, sqa = cloud(seed=3, unit=False)
    nrm[0] = [0.0, 0.0, 1.0]  # already unit length: kept as is
    path = str(tmp_path / "in.vtp")
    write_vtk_cloud(path, pts, nrm, sqa, radius_name="sqrt-of-area", normals_name="nrm", fmt="binary")
    got = run_probe(path)
    # radius = sqrt-of-area * sqrt(3)/2 * (1 + 1/32), in double, then float
    want = (sqa.astype(np.float64) * (np.sqrt(3.0) / 2 * (1 + 1.0 / 32))).astype(np.float32)
    np.testing.assert_array_equal(got[:, 3].astype(np.float32), want)
    # normals off unit length by more than 1e-6 are divided by their length (float)
    n32 = nrm.astype(np.float32)
    ln = np.sqrt((n32 * n32).sum(1, dtype=np.float32)).astype(np.float32)
    np.testing.assert_allclose(got[:, 4:7], n32 / ln[:, None], rtol=2e-7, atol=1e-7)
    np.testing.assert_array_equal(got[0, 4:7], [0.0, 0.0, 1.0])


def test_normals_found_without_declaration(tmp_path):
    pts, nrm, rad = cloud(seed=4)
    path = str(tmp_path / "in.vtp")
    write_vtk_cloud(path, pts, nrm, rad, fmt="ascii", normals_name="Normals", declare_normals=False)
    got = run_probe(path)
    np.testing.assert_allclose(got[:, 4:7], nrm, atol=1e-7)


def test_missing_radius_is_an_error(tmp_path):
    pts, nrm, rad = cloud(seed=5)
    path = str(tmp_path / "in.vtp")
    write_vtk_cloud(path, pts, nrm, rad, radius_name="thickness")
    r = subprocess.run([PROBE, path], capture_output=True, text=True)
    assert r.returncode != 0 and "radius" in r.stderr


CLI = os.path.join(ROOT, "rti_amd", "cli", "rti_cli")


def _cli(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_cli_arguments_checked_before_the_trace(tmp_path):
    """dev.cpp:51-180: -i is required; .vtk input and --triangles end the
    program with the reference's messages, before any trace."""
    if not os.path.exists(CLI):
        pytest.skip("rti_cli is not built")
    r = _cli()
    assert r.returncode != 0 and "usage" in r.stdout
    r = _cli("-i", str(tmp_path / "x.vtk"))
    assert r.returncode != 0 and "convert the input file to vtp" in r.stdout
    pts, nrm, rad = cloud(seed=7)
    path = str(tmp_path / "in.vtp")
    write_vtk_cloud(path, pts, nrm, rad)
    r = _cli("-i", path, "--triangles")
    assert r.returncode != 0 and "not supported" in r.stderr
    r = _cli("-i", path, "-s", "1.5", "--triangles")
    assert "reset to sane value" in r.stdout and "sticking coefficient == 1" in r.stdout


def test_writer_layout_and_round_trip(tmp_path):
    pts, nrm, rad = cloud(n=31, seed=6)
    src = str(tmp_path / "in.vtp")
    out = str(tmp_path / "out.vtp")
    write_vtk_cloud(src, pts, nrm, rad, fmt="binary", compress=True, header="UInt64")
    a = run_probe(src, out)
    b = run_probe(out)  # our reader on our writer's output
    np.testing.assert_array_equal(a, b)
    root = ET.parse(out).getroot()  # well-formed XML, reference layout
    assert root.attrib["type"] == "PolyData"
    piece = root.find("PolyData/Piece")
    assert piece.attrib["NumberOfPoints"] == "31" and piece.attrib["NumberOfVerts"] == "31"
    cd = piece.find("CellData")
    assert cd.attrib["Normals"] == "Normals"
    arr = {d.attrib["Name"]: d for d in cd.findall("DataArray")}
    assert set(arr) == {"Normals", "radius", "value", "hitcnt"}
    assert arr["hitcnt"].attrib["type"] == "UInt32" and arr["value"].attrib["type"] == "Float64"
    np.testing.assert_array_equal(np.array(arr["value"].text.split(), float), 0.5 * np.arange(31))
    np.testing.assert_array_equal(np.array(arr["hitcnt"].text.split(), int), np.arange(31))
    conn = piece.find("Verts/DataArray[@Name='connectivity']").text.split()
    offs = piece.find("Verts/DataArray[@Name='offsets']").text.split()
    assert conn == [str(i) for i in range(31)] and offs == [str(i + 1) for i in range(31)]
    fd = root.find("PolyData/FieldData/Array")
    codes = [int(x) for x in fd.text.split()]
    assert fd.attrib["Name"] == "cmd" and codes[-1] == 0
    assert bytes(codes[:-1]).decode() == 'probe <&> "quoted"'


@pytest.mark.parametrize("compress", [False, True])
def test_corrupt_binary_headers_are_errors(tmp_path, compress):
    """A byte count or block count near 2^64 in an appended raw array (it
    would wrap the bounds arithmetic) is reported as truncated data, not
    read past the buffer."""
    pts, nrm, rad = cloud(seed=6)
    path = tmp_path / "in.vtp"
    write_vtk_cloud(str(path), pts, nrm, rad, fmt="appended", header="UInt64", encoding="raw", compress=compress)
    data = bytearray(path.read_bytes())
    start = data.index(b"_", data.index(b"<AppendedData")) + 1
    # the first array's header: byte count (plain) or block count (zlib)
    data[start:start + 8] = struct.pack("<Q", 0xFFFFFFFFFFFFFFF8)
    path.write_bytes(bytes(data))
    r = subprocess.run([PROBE, str(path)], capture_output=True, text=True)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert "ends early" in r.stderr
