"""VTK UnstructuredGrid output of DG fields (src/vtu.mojo, src/vtu_2d.mojo,
src/frame_writer.mojo): every element writes its own NP nodes, so the
discontinuities between elements stay visible.

Format (as the reference's writers): XML with one appended raw-binary blob
after a literal '_', each array preceded by its UInt32 byte count; points
as Float32 x, y, z (z = 0 on triangles); connectivity 0..K*NP-1; offsets
(e+1)*NP; one cell type per element -- tetrahedra: VTK_QUADRATIC_TETRA (24)
at NP = 10, else VTK_LAGRANGE_TETRAHEDRON (71); triangles: VTK_TRIANGLE (5),
VTK_QUADRATIC_TRIANGLE (22) at NP = 6, else VTK_LAGRANGE_TRIANGLE (69).
The nodes are written in the reference element's order, as the
reference's writers do: vertices, then the edge nodes in VTK's edge order
(0-1, 1-2, 2-0, 0-3, 1-3, 2-3, each from its first vertex), then face and
interior nodes; at NP = 10 and 6 that is exactly VTK's quadratic cell.
Point fields are Float32 scalars named by the caller; the first is the
default.

A ParaView collection (.pvd) ties frames to their times (write_pvd), and
FrameWriter writes one frame per call from a Solver or a batch instance."""
from __future__ import annotations

import os
from typing import Dict, List, Sequence

import numpy as np

VTK_TRIANGLE, VTK_QUADRATIC_TRIANGLE, VTK_LAGRANGE_TRIANGLE = 5, 22, 69
VTK_QUADRATIC_TETRA, VTK_LAGRANGE_TETRAHEDRON = 24, 71


def cell_type(ndim: int, nodes_per_elem: int) -> int:
    if ndim == 3:
        return VTK_QUADRATIC_TETRA if nodes_per_elem == 10 else VTK_LAGRANGE_TETRAHEDRON
    if nodes_per_elem == 3:
        return VTK_TRIANGLE
    return VTK_QUADRATIC_TRIANGLE if nodes_per_elem == 6 else VTK_LAGRANGE_TRIANGLE


def frame_name(prefix: str, i: int, width: int = 5) -> str:
    """prefix + zero-padded i + '.vtu' (src/vtu_2d.mojo:314-327)."""
    return f"{prefix}{i:0{width}d}.vtu"


def write_vtu(path: str, elem_node_xyz: np.ndarray, fields: Dict[str, np.ndarray]) -> None:
    """One frame: elem_node_xyz [K, NP, 2 or 3]; fields name -> [K, NP]."""
    xyz = np.asarray(elem_node_xyz, dtype=np.float64)
    K, NP, nd = xyz.shape
    if not fields:
        raise ValueError("write_vtu needs at least one field")
    pts = np.zeros((K * NP, 3), dtype="<f4")
    pts[:, :nd] = xyz.reshape(-1, nd)
    arrays = []  # (xml element, raw little-endian bytes)
    names = list(fields)
    for name in names:
        f = np.asarray(fields[name], dtype=np.float64)
        if f.shape != (K, NP):
            raise ValueError(f"field {name!r} has shape {f.shape}, expected {(K, NP)}")
        arrays.append(("point", f'<DataArray type="Float32" Name="{name}" format="appended" offset="{{}}"/>',
                       f.astype("<f4").tobytes()))
    arrays.append(("points", '<DataArray type="Float32" NumberOfComponents="3" format="appended" offset="{}"/>',
                   pts.tobytes()))
    arrays.append(("cells", '<DataArray type="Int32" Name="connectivity" format="appended" offset="{}"/>',
                   np.arange(K * NP, dtype="<i4").tobytes()))
    arrays.append(("cells", '<DataArray type="Int32" Name="offsets" format="appended" offset="{}"/>',
                   (np.arange(1, K + 1, dtype="<i4") * NP).tobytes()))
    arrays.append(("cells", '<DataArray type="UInt8" Name="types" format="appended" offset="{}"/>',
                   np.full(K, cell_type(nd, NP), dtype=np.uint8).tobytes()))
    off, blob, xml = 0, [], {"point": [], "points": [], "cells": []}
    for where, tag, raw in arrays:
        xml[where].append(tag.format(off))
        blob.append(np.uint32(len(raw)).astype("<u4").tobytes() + raw)
        off += 4 + len(raw)
    head = ['<?xml version="1.0"?>',
            '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" header_type="UInt32">',
            "<UnstructuredGrid>", f'<Piece NumberOfPoints="{K * NP}" NumberOfCells="{K}">',
            f'<PointData Scalars="{names[0]}">', *xml["point"], "</PointData>",
            "<Points>", *xml["points"], "</Points>", "<Cells>", *xml["cells"], "</Cells>",
            "</Piece>", "</UnstructuredGrid>", '<AppendedData encoding="raw">']
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as fh:
        fh.write(("\n".join(head) + "\n_").encode())
        for b in blob:
            fh.write(b)
        fh.write(b"\n</AppendedData>\n</VTKFile>\n")


def read_vtu(path: str) -> dict:
    """Parse a file written by write_vtu: points [N, 3], connectivity,
    offsets, types, and the point fields (a round-trip check)."""
    import re
    data = open(path, "rb").read()
    cut = data.index(b"_", data.index(b"<AppendedData"))
    head = data[:cut].decode()
    blob = data[cut + 1:]
    out = {"fields": {}}
    dtypes = {"Float32": "<f4", "Int32": "<i4", "UInt8": "u1"}
    for m in re.finditer(r'<DataArray type="(\w+)"(?: Name="([^"]*)")?(?: NumberOfComponents="(\d+)")? '
                         r'format="appended" offset="(\d+)"/>', head):
        typ, name, ncomp, off = m.group(1), m.group(2), m.group(3), int(m.group(4))
        n = int(np.frombuffer(blob[off:off + 4], dtype="<u4")[0])
        arr = np.frombuffer(blob[off + 4:off + 4 + n], dtype=dtypes[typ])
        if ncomp:
            out["points"] = arr.reshape(-1, int(ncomp))
        elif name in ("connectivity", "offsets", "types"):
            out[name] = arr
        else:
            out["fields"][name] = arr
    return out


def write_pvd(path: str, vtu_paths: Sequence[str], times: Sequence[float]) -> None:
    """ParaView collection of frames and their times (src/vtu.mojo:251-283);
    the frame paths are relative to the .pvd's directory."""
    if len(vtu_paths) != len(times):
        raise ValueError("write_pvd: vtu_paths/times length mismatch")
    lines = ['<?xml version="1.0"?>', '<VTKFile type="Collection" version="0.1" byte_order="LittleEndian">',
             "<Collection>"]
    lines += [f'<DataSet timestep="{t!r}" group="" part="0" file="{p}"/>' for p, t in zip(vtu_paths, times)]
    lines += ["</Collection>", "</VTKFile>"]
    with open(path, "w") as fh:
        fh.write("\n".join(lines) + "\n")


class FrameWriter:
    """Frames of chosen state components during a run (src/frame_writer.mojo:
    86-206): write(t) downloads the state of `source` -- a Solver, or a
    (Batch, instance id) pair -- and writes output_dir/prefixNNNNN.vtu;
    finalize() writes the .pvd of every frame written."""

    def __init__(self, source, elem_node_xyz: np.ndarray, output_dir: str = "output", prefix: str = "frame_",
                 components: Dict[str, int] = None):
        self.source = source
        self.xyz = np.asarray(elem_node_xyz)
        self.output_dir = output_dir
        self.prefix = prefix
        self.components = components or {"q0": 0}
        self.paths: List[str] = []
        self.times: List[float] = []
        os.makedirs(output_dir, exist_ok=True)

    def _state(self) -> np.ndarray:
        if isinstance(self.source, tuple):
            batch, ident = self.source
            return batch.get_state(ident)
        return self.source.get_state()

    def write(self, t: float) -> str:
        q = self._state()
        name = frame_name(self.prefix, len(self.paths))
        write_vtu(os.path.join(self.output_dir, name), self.xyz,
                  {k: q[..., c] for k, c in self.components.items()})
        self.paths.append(name)
        self.times.append(float(t))
        return os.path.join(self.output_dir, name)

    def finalize(self, pvd_name: str = "frames.pvd") -> str:
        path = os.path.join(self.output_dir, pvd_name)
        write_pvd(path, self.paths, self.times)
        return path
