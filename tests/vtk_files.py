"""Writes VTK XML PolyData point clouds in the encodings vtkXMLPolyDataWriter
produces (ascii; inline binary; appended base64 or raw; UInt32 or UInt64
headers; optional zlib blocks), following the VTK XML file-format layout.
Test input generator for the command-line tool's reader."""
import base64
import struct
import zlib

import numpy as np


def _block(arr, header, compress, block_size=64):
    """(header bytes, data bytes) of one array in VTK's binary layout."""
    raw = arr.tobytes()
    hfmt = "<I" if header == "UInt32" else "<Q"
    if not compress:
        return struct.pack(hfmt, len(raw)), raw
    blocks = [raw[i:i + block_size] for i in range(0, len(raw), block_size)] or [b""]
    comp = [zlib.compress(b) for b in blocks]
    last = len(blocks[-1]) if len(blocks[-1]) != block_size else 0
    h = struct.pack("<" + hfmt[1] * (3 + len(comp)), len(comp), block_size, last, *[len(c) for c in comp])
    return h, b"".join(comp)


def write_vtk_cloud(path, pts, nrm, rad, *, fmt="ascii", header="UInt32", compress=False, joined=False,
                    encoding="base64", radius_name="radius", normals_name="Normals", declare_normals=True):
    n = pts.shape[0]
    arrays = [("Points", pts.astype(np.float32), 3, "Points"),
              (normals_name, nrm.astype(np.float64), 3, "CellData"),
              (radius_name, rad.astype(np.float32), 1, "CellData"),
              ("noise", np.arange(n, dtype=np.int32), 1, "CellData")]
    appended = b""

    def encode(arr):
        nonlocal appended
        h, d = _block(arr, header, compress)
        if fmt == "ascii":
            return " ".join(repr(float(x)) if arr.dtype.kind == "f" else str(int(x)) for x in arr.ravel()), {}
        if fmt == "binary":
            if joined and not compress:
                return base64.b64encode(h + d).decode(), {}
            return base64.b64encode(h).decode() + base64.b64encode(d).decode(), {}
        # appended
        off = len(appended)
        chunk = (h + d) if encoding == "raw" else (base64.b64encode(h) + base64.b64encode(d))
        appended += chunk
        return None, {"offset": str(off)}

    vtype = {np.dtype(np.float32): "Float32", np.dtype(np.float64): "Float64", np.dtype(np.int32): "Int32"}
    body = {"Points": [], "CellData": []}
    for name, arr, nc, sec in arrays:
        text, extra = encode(arr)
        attrs = f'type="{vtype[arr.dtype]}" Name="{name}" NumberOfComponents="{nc}" format="{fmt}"'
        for k, v in extra.items():
            attrs += f' {k}="{v}"'
        if text is None:
            body[sec].append(f"<DataArray {attrs}/>")
        else:
            body[sec].append(f"<DataArray {attrs}>\n{text}\n</DataArray>")
    comp = ' compressor="vtkZLibDataCompressor"' if compress else ""
    cd_attr = f' Normals="{normals_name}"' if declare_normals else ""
    conn = " ".join(str(i) for i in range(n))
    offs = " ".join(str(i + 1) for i in range(n))
    head = (f'<?xml version="1.0"?>\n<VTKFile type="PolyData" version="1.0" byte_order="LittleEndian" '
            f'header_type="{header}"{comp}>\n<PolyData>\n<Piece NumberOfPoints="{n}" NumberOfVerts="{n}" '
            f'NumberOfLines="0" NumberOfStrips="0" NumberOfPolys="0">\n'
            f'<CellData{cd_attr}>\n' + "\n".join(body["CellData"]) + "\n</CellData>\n<Points>\n" +
            "\n".join(body["Points"]) + "\n</Points>\n<Verts>\n"
            f'<DataArray type="Int64" Name="connectivity" format="ascii">{conn}</DataArray>\n'
            f'<DataArray type="Int64" Name="offsets" format="ascii">{offs}</DataArray>\n'
            "</Verts>\n</Piece>\n</PolyData>\n")
    with open(path, "wb") as fh:
        fh.write(head.encode())
        if fmt == "appended":
            fh.write(f'<AppendedData encoding="{encoding}">\n_'.encode() + appended + b"\n</AppendedData>\n")
        fh.write(b"</VTKFile>\n")
