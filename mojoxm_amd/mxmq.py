"""MXMQ state dumps: the reference's per-rank binary format
(test/mpi_advection_test.mojo:14-21, read by test/diff_mpi_dumps.py:26-63),
for exchanging states with the reference's tools.

Layout, little-endian:
    u32 magic = 0x514D584D ('MXMQ'), u32 version, u32 num_owned_elements,
    u32 num_components, u32 nodes_per_elem,
    u32[num_owned] global element ids (order of the values below),
    values [num_owned][nodes_per_elem][num_components]
Version 1 (the reference's) stores f32 values.  Version 2 is this
package's extension with f64 values, for FP64 states; the reference's
reader rejects it, so write version 1 for the reference's tools."""
from __future__ import annotations

import glob
import os
import struct

import numpy as np

MAGIC = 0x514D584D
HEADER = struct.Struct("<IIIII")
DTYPES = {1: np.float32, 2: np.float64}


def write_dump(path: str, q: np.ndarray, elem_ids=None, version: int = 1) -> None:
    """q: [num_owned, nodes_per_elem, num_components]; elem_ids default to
    0..num_owned-1 (one rank owning the whole mesh)."""
    if version not in DTYPES:
        raise ValueError("MXMQ version must be 1 (f32) or 2 (f64)")
    q = np.asarray(q)
    if q.ndim != 3:
        raise ValueError("q must be [elements, nodes, components]")
    n, npn, nc = q.shape
    ids = np.arange(n, dtype=np.uint32) if elem_ids is None else np.asarray(elem_ids, dtype=np.uint32)
    if ids.shape != (n,):
        raise ValueError("one global id per element")
    d = os.path.dirname(path)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(path, "wb") as f:
        f.write(HEADER.pack(MAGIC, version, n, nc, npn))
        f.write(ids.astype("<u4").tobytes())
        f.write(np.ascontiguousarray(q, dtype=np.dtype(DTYPES[version]).newbyteorder("<")).tobytes())


def read_dump(path: str):
    """Returns (global ids [n], q [n, nodes_per_elem, num_components])."""
    with open(path, "rb") as f:
        raw = f.read()
    if len(raw) < HEADER.size:
        raise ValueError(f"{path}: truncated header")
    magic, version, n, nc, npn = HEADER.unpack_from(raw)
    if magic != MAGIC or version not in DTYPES:
        raise ValueError(f"{path}: bad header (magic={magic:#x}, version={version})")
    dt = np.dtype(DTYPES[version]).newbyteorder("<")
    off = HEADER.size
    need = off + 4 * n + dt.itemsize * n * npn * nc
    if len(raw) != need:
        raise ValueError(f"{path}: {len(raw)} bytes, header says {need}")
    ids = np.frombuffer(raw, dtype="<u4", count=n, offset=off).astype(np.uint32)
    q = np.frombuffer(raw, dtype=dt, count=n * npn * nc, offset=off + 4 * n).reshape(n, npn, nc)
    return ids, q.astype(DTYPES[version])


def load_directory(d: str) -> dict:
    """Union over the final_q_rank_*.bin dumps of a directory: global id ->
    values [nodes_per_elem * num_components]; a duplicate owner is an error
    (test/diff_mpi_dumps.py:26-63)."""
    paths = sorted(glob.glob(os.path.join(d, "final_q_rank_*.bin")))
    if not paths:
        raise ValueError(f"no final_q_rank_*.bin under {d!r}")
    out = {}
    for p in paths:
        ids, q = read_dump(p)
        for k, gid in enumerate(ids):
            gid = int(gid)
            if gid in out:
                raise ValueError(f"duplicate owner for global_elem_id={gid} in {d}")
            out[gid] = q[k].reshape(-1)
    return out


def compare_directories(a: str, b: str, tol_abs: float = 1e-5, tol_rel: float = 1e-4):
    """The reference's dump comparison (test/diff_mpi_dumps.py:66-108):
    the same global element set, then max |a - b| and max |a - b| /
    max(|a|, 1) over matched elements; passes when both are within their
    tolerances.  Returns (passed, max_abs, max_rel)."""
    da, db = load_directory(a), load_directory(b)
    if set(da) != set(db):
        raise ValueError("the two dump sets cover different global elements")
    max_abs = max_rel = 0.0
    for gid, va in da.items():
        va64, vb64 = va.astype(np.float64), db[gid].astype(np.float64)
        diff = np.abs(va64 - vb64)
        max_abs = max(max_abs, float(diff.max(initial=0.0)))
        denom = np.maximum(np.abs(va64), 1.0)
        max_rel = max(max_rel, float((diff / denom).max(initial=0.0)))
    return (max_abs <= tol_abs and max_rel <= tol_rel), max_abs, max_rel
