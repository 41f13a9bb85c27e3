"""The command-line tracer end to end (the reference's main/dev.cpp): a .vtp
point cloud in, the trace on the GPU, a .vtp out whose cell data equal the
library's results for the same inputs (periodic x and y, cosine source on the
top of the centre box padded once by the largest radius, diffuse reflection),
and the boundary-box .vtp."""
import os
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from rti_amd import run_scene
from tests.helpers import small_trench
from tests.vtk_files import write_vtk_cloud

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "rti_amd", "cli", "rti_cli")


def _arrays(path):
    piece = ET.parse(path).getroot().find("PolyData/Piece")
    out = {}
    for sec in ("CellData", "Points"):
        for d in piece.find(sec).findall("DataArray"):
            out[d.attrib["Name"]] = np.array(d.text.split(), dtype=np.float64)
    return piece, out


def test_cli_output_equals_library(tmp_path):
    sc = small_trench()
    xyzr, nrm = sc["xyzr"], sc["normals"]
    src = str(tmp_path / "trench.vtp")
    write_vtk_cloud(src, xyzr[:, :3], nrm, xyzr[:, 3], fmt="binary", compress=True, header="UInt64")
    out = str(tmp_path / "result")  # the tool appends .vtp
    print("[gpu] rti_cli on a trench point cloud", flush=True)
    r = subprocess.run([CLI, "-i", src, "-o", out, "-r", "150000", "-s", "0.2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Appending .vtp" in r.stdout
    piece, got = _arrays(out + ".vtp")
    n = xyzr.shape[0]
    assert int(piece.attrib["NumberOfPoints"]) == n
    ztop = np.float32(xyzr[:, 2].max() + xyzr[:, 3].max())
    lib = run_scene(sc, n_rays=150_000, sticking=0.2, bc_x=1, bc_y=1, src_z=float(ztop), seed=1, moments=1)
    np.testing.assert_array_equal(got["value"], lib["raw"])
    np.testing.assert_array_equal(got["hitcnt"], lib["counts"].astype(np.float64))
    np.testing.assert_array_equal(got["exposed-area"], lib["areas"].astype(np.float64))
    np.testing.assert_array_equal(got["relative-error"], lib["rel_err"])
    np.testing.assert_allclose(got["variance-of-variance"], lib["vov"], rtol=1e-9)
    np.testing.assert_array_equal(got["radius"], xyzr[:, 3].astype(np.float64))
    np.testing.assert_array_equal(got["Points"].reshape(n, 3).astype(np.float32), xyzr[:, :3])
    # boundary box: x/y extent of the centres, z from the lowest centre to the source plane
    bpiece, b = _arrays(out + ".bounding-box.vtp")
    assert bpiece.attrib["NumberOfPolys"] == "8"
    v = b["Points"].reshape(8, 3).astype(np.float32)
    lo, hi = xyzr[:, :3].min(0), xyzr[:, :3].max(0)
    assert set(v[:, 0]) == {lo[0], hi[0]} and set(v[:, 1]) == {lo[1], hi[1]}
    assert set(v[:, 2]) == {lo[2], ztop}
    nn = b["Normals"].reshape(8, 3)
    np.testing.assert_allclose(np.linalg.norm(nn, axis=1), 1.0, rtol=1e-12)
