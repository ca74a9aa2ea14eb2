"""Generate integer fixtures of the 3D Kuhn mesh from the reference's literal
tables, read as text.

The reference defines the Kuhn split by four small literal tables written as
chains of `if ... return` in src/local_mesh.mojo:
    kuhn_vertex(t, k)        :64-114   cube corner of tet t's vertex k
    corner_dx/dy/dz(c)       :117-132  corner c's offset along x, y, z
    tet_face_info(t, f)      :168-217  (face type, owner-cell offset, side)
    face_type_corner(ft, i)  :220-292  owner-cube corner at canonical position i
This script parses those function bodies as text (a small reader of
`if <name> == <int> [or ...]:` / `return <value>` chains; nothing of the
reference is imported or run) and then builds the connectivity of a periodic
2x2x2 mesh for P = 1..3 by geometry alone:
    element e = cell * 6 + t, cell = i + Nx (j + Ny k)
    face id   = owner cell * 12 + face type, owner = cell + offset (wrapped)
    canonical face node m: barycentric weights over the face type's three
    corners (vertices, edges 01/12/20, interior; src/local_mesh.mojo:401-455)
    -> an integer lattice point; face_elem_node is the element-local node at
    that point (modulo the periodic wrap), elem_canon_to_ref the reference
    face slot holding that node (face map from refelem_golden.npz).
The normals follow the reference's rule (:700-739): the corners' cross
product, turned away from the side-0 tet's opposite vertex.

Run from the repository root (reads /root/reference, or the path given):

    python tests/golden/make_mesh_golden.py [/path/to/reference]

It writes tests/golden/kuhn_mesh_golden.npz.
"""
from __future__ import annotations

import itertools
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "kuhn_mesh_golden.npz")
sys.path.insert(0, HERE)
from make_golden import tet_nodes  # noqa: E402  (the reference's node order, restated there)

_COND = re.compile(r"(\w+)\s*(==|>=|<=|>|<)\s*(-?\d+)")


def function_body(text: str, name: str) -> list[str]:
    """Lines of `def name(...)` up to the next top-level statement."""
    lines = text.splitlines()
    for k, ln in enumerate(lines):
        if ln.startswith(f"def {name}("):
            body = []
            for ln2 in lines[k + 1:]:
                if ln2 and not ln2[0].isspace():
                    break
                body.append(ln2)
            return body
    raise KeyError(name)


def _holds(cond: str, env: dict) -> bool:
    terms = [t.strip() for t in cond.split(" or ")]
    out = False
    for t in terms:
        m = _COND.fullmatch(t)
        if not m:
            raise ValueError(f"unsupported condition {t!r}")
        a, op, b = env[m.group(1)], m.group(2), int(m.group(3))
        out |= {"==": a == b, ">=": a >= b, "<=": a <= b, ">": a > b, "<": a < b}[op]
    return out


def _value(expr: str):
    expr = expr.strip()
    m = re.fullmatch(r"\w+\((.*)\)", expr)
    if m:
        return tuple(int(x) for x in m.group(1).split(","))
    return int(expr)


def evaluate(body: list[str], env: dict):
    """Walk an if/return chain for the argument values in env."""
    skip_deeper = None
    for ln in body:
        s = ln.split("#", 1)[0].rstrip()
        if not s.strip():
            continue
        ind = len(s) - len(s.lstrip())
        if skip_deeper is not None:
            if ind > skip_deeper:
                continue
            skip_deeper = None
        st = s.strip()
        if st.startswith("if ") and st.endswith(":"):
            if not _holds(st[3:-1], env):
                skip_deeper = ind
            continue
        if st.startswith("return "):
            return _value(st[len("return "):])
        if st.startswith("var ") or st.startswith("comptime "):
            raise ValueError(f"unsupported statement {st!r}")
    raise ValueError("no return reached")


def literal_tables(ref_root: str) -> dict:
    text = open(os.path.join(ref_root, "src", "local_mesh.mojo")).read()
    kv = function_body(text, "kuhn_vertex")
    tfi = function_body(text, "tet_face_info")
    ftc = function_body(text, "face_type_corner")
    cd = [function_body(text, f"corner_d{a}") for a in "xyz"]
    return {
        "kuhn_vertex": np.array([[evaluate(kv, {"t": t, "k": k}) for k in range(4)] for t in range(6)], np.int32),
        "tet_face_info": np.array([[evaluate(tfi, {"t": t, "f": f}) for f in range(4)] for t in range(6)], np.int32),
        "face_type_corner": np.array([[evaluate(ftc, {"ft": ft, "idx": i}) for i in range(3)] for ft in range(12)],
                                     np.int32),
        "corner_offset": np.array([[evaluate(cd[a], {"c": c}) for a in range(3)] for c in range(8)], np.int32),
    }


def canon_face_bary(P: int) -> list[tuple[int, int, int]]:
    """Canonical face-node order: vertices 0, 1, 2; interior of edges 01, 12,
    20 (each walked from its first corner); face interior, (b0, b1)
    lexicographically descending."""
    out = [(P, 0, 0), (0, P, 0), (0, 0, P)]
    for k in range(1, P):
        out.append((P - k, k, 0))
    for k in range(1, P):
        out.append((0, P - k, k))
    for k in range(1, P):
        out.append((k, 0, P - k))
    for b0 in range(P - 2, 0, -1):
        for b1 in range(P - b0 - 1, 0, -1):
            out.append((b0, b1, P - b0 - b1))
    return out


def build_mesh(tab: dict, P: int, n=(2, 2, 2), L=(1.0, 1.0, 1.0)) -> dict:
    kv, tfi, ftc, co = tab["kuhn_vertex"], tab["tet_face_info"], tab["face_type_corner"], tab["corner_offset"]
    Nx, Ny, Nz = n
    ncell = Nx * Ny * Nz
    K, Nf = 6 * ncell, 12 * ncell
    exps = tet_nodes(P)
    NP, NFP = len(exps), (P + 1) * (P + 2) // 2
    fmap = np.load(os.path.join(HERE, "refelem_golden.npz"))[f"fmap_3d_p{P}"].reshape(4, NFP)
    cell_ijk = [(c % Nx, (c // Nx) % Ny, c // (Nx * Ny)) for c in range(ncell)]
    period = np.array([Nx * P, Ny * P, Nz * P])

    # element node lattice points (units of cell / P)
    node_pt = np.zeros((K, NP, 3), np.int64)
    for c, (i, j, k) in enumerate(cell_ijk):
        for t in range(6):
            V = co[kv[t]]  # tet vertices' corner offsets [4][3]
            for nn, a in enumerate(exps):
                node_pt[c * 6 + t, nn] = P * np.array([i, j, k]) + np.asarray(a) @ V
    elem_faces = np.zeros((K, 4), np.int32)
    elem_face_side = np.zeros((K, 4), np.int32)
    face_elem = -np.ones((Nf, 2), np.int32)
    for c, (i, j, k) in enumerate(cell_ijk):
        for t in range(6):
            e = c * 6 + t
            for f in range(4):
                ft, di, dj, dk, side = tfi[t, f]
                oc = (i + di) % Nx + Nx * (((j + dj) % Ny) + Ny * ((k + dk) % Nz))
                elem_faces[e, f] = oc * 12 + ft
                elem_face_side[e, f] = side
                assert face_elem[oc * 12 + ft, side] == -1
                face_elem[oc * 12 + ft, side] = e
    assert (face_elem >= 0).all()
    bary = canon_face_bary(P)
    face_elem_node = np.zeros((Nf, 2, NFP), np.int32)
    for fid in range(Nf):
        c, ft = divmod(fid, 12)
        corners = co[ftc[ft]]
        for m, b in enumerate(bary):
            x = P * np.array(cell_ijk[c]) + np.asarray(b) @ corners
            for s in range(2):
                e = face_elem[fid, s]
                hit = np.nonzero(((node_pt[e] - x) % period == 0).all(axis=1))[0]
                assert hit.size == 1
                face_elem_node[fid, s, m] = hit[0]
    elem_canon_to_ref = np.zeros((K, 4, NFP), np.int32)
    for e in range(K):
        for f in range(4):
            fid, s = elem_faces[e, f], elem_face_side[e, f]
            for m in range(NFP):
                r = np.nonzero(fmap[f] == face_elem_node[fid, s, m])[0]
                assert r.size == 1
                elem_canon_to_ref[e, f, m] = r[0]
    # normals and areas per face type, physical cell size h
    h = np.array([L[0] / Nx, L[1] / Ny, L[2] / Nz])
    ft_n, ft_a = np.zeros((12, 3)), np.zeros(12)
    for ft in range(12):
        p = co[ftc[ft]] * h
        cr = np.cross(p[1] - p[0], p[2] - p[0])
        nrm = cr / np.linalg.norm(cr)
        t0, f0 = next((t, f) for t in range(6) for f in range(4) if tfi[t, f, 0] == ft and tfi[t, f, 4] == 0)
        v4 = co[kv[t0, f0]] * h
        if (v4 - p[0]) @ nrm > 0:
            nrm = -nrm
        ft_n[ft], ft_a[ft] = nrm, 0.5 * np.linalg.norm(cr)
    fts = np.arange(Nf) % 12
    return {"elem_faces": elem_faces, "elem_face_side": elem_face_side, "face_elem": face_elem,
            "face_elem_node": face_elem_node, "elem_canon_to_ref": elem_canon_to_ref,
            "face_normal": ft_n[fts], "face_area": ft_a[fts]}


def main(ref_root: str = "/root/reference") -> None:
    tab = literal_tables(ref_root)
    out = dict(tab)
    for P in (1, 2, 3):
        for name, a in build_mesh(tab, P).items():
            out[f"{name}_p{P}"] = a
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}: " + ", ".join(f"{k}{list(v.shape)}" for k, v in out.items() if not k[-1].isdigit()))


if __name__ == "__main__":
    main(*sys.argv[1:])
