"""MXMQ state dumps (mojoxm_amd/mxmq.py) against the reference's format
description (test/mpi_advection_test.mojo:14-21) and comparison rules
(test/diff_mpi_dumps.py:26-108), and the diagnostics CSV writer's layout
(src/diagnostics.mojo:95-223) on a stand-in source (no GPU)."""
import struct

import numpy as np
import pytest

from mojoxm_amd import mxmq
from mojoxm_amd.diagnostics import DiagComponents, Diagnostics, DiagnosticsWriter


def test_header_bytes_and_layout(tmp_path):
    q = np.arange(3 * 10 * 1, dtype=np.float64).reshape(3, 10, 1) / 7.0
    p = tmp_path / "final_q_rank_0.bin"
    mxmq.write_dump(str(p), q, elem_ids=[5, 1, 9])
    raw = p.read_bytes()
    assert raw[:4] == b"MXMQ"
    assert struct.unpack("<IIIII", raw[:20]) == (0x514D584D, 1, 3, 1, 10)
    assert np.array_equal(np.frombuffer(raw[20:32], "<u4"), [5, 1, 9])
    assert len(raw) == 20 + 12 + 30 * 4
    assert np.array_equal(np.frombuffer(raw[32:], "<f4"), q.ravel().astype(np.float32))


@pytest.mark.parametrize("version,dtype", [(1, np.float32), (2, np.float64)])
def test_round_trip(tmp_path, version, dtype):
    rng = np.random.default_rng(0)
    q = rng.standard_normal((7, 10, 5))
    p = str(tmp_path / "d.bin")
    mxmq.write_dump(p, q, version=version)
    ids, r = mxmq.read_dump(p)
    assert np.array_equal(ids, np.arange(7)) and r.dtype == dtype
    assert np.array_equal(r, q.astype(dtype))


def test_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.bin"
    p.write_bytes(struct.pack("<IIIII", 0x12345678, 1, 0, 1, 1))
    with pytest.raises(ValueError):
        mxmq.read_dump(str(p))
    p.write_bytes(struct.pack("<IIIII", 0x514D584D, 1, 2, 1, 10) + b"\0" * 8)
    with pytest.raises(ValueError):
        mxmq.read_dump(str(p))
    with pytest.raises(ValueError):
        mxmq.write_dump(str(p), np.zeros((2, 3, 1)), version=3)


def test_directory_union_and_compare(tmp_path):
    rng = np.random.default_rng(1)
    q = rng.uniform(0.5, 1.5, (8, 10, 1))
    a, b, c = tmp_path / "a", tmp_path / "b", tmp_path / "c"
    mxmq.write_dump(str(a / "final_q_rank_0.bin"), q)                        # np = 1
    mxmq.write_dump(str(b / "final_q_rank_0.bin"), q[:3], elem_ids=[0, 1, 2])  # np = 2
    mxmq.write_dump(str(b / "final_q_rank_1.bin"), q[3:] * (1 + 2e-6), elem_ids=range(3, 8))
    ok, mabs, mrel = mxmq.compare_directories(str(a), str(b))
    assert ok and 0 < mabs < 1e-5 and 0 < mrel < 1e-5
    ok, _, _ = mxmq.compare_directories(str(a), str(b), tol_abs=1e-7)
    assert not ok  # both tolerances must hold
    mxmq.write_dump(str(c / "final_q_rank_0.bin"), q[:4], elem_ids=range(4))
    mxmq.write_dump(str(c / "final_q_rank_1.bin"), q[3:], elem_ids=range(3, 8))
    with pytest.raises(ValueError, match="duplicate owner"):
        mxmq.load_directory(str(c))
    mxmq.write_dump(str(c / "final_q_rank_1.bin"), q[4:7], elem_ids=range(4, 7))
    with pytest.raises(ValueError, match="different global elements"):
        mxmq.compare_directories(str(a), str(c))


class _Source:
    def __init__(self):
        self.calls = 0

    def diagnostics(self):
        self.calls += 1
        nc = 5
        v = np.arange(3 * nc + 2, dtype=np.float64) + self.calls
        return Diagnostics.from_flat(v, nc)


def test_diagnostics_writer_csv(tmp_path):
    comps = (DiagComponents().linear("mass", 0).linear("energy", 4).squared("mx2", 1).maxabs("rho_max", 0)
             .kinetic_energy().enstrophy())
    path = tmp_path / "out" / "diagnostics.csv"
    w = DiagnosticsWriter(_Source(), str(path), comps)
    w.record(0.0)
    row = w.record(0.5)
    lines = path.read_text().splitlines()
    assert lines[0] == "time,mass,energy,mx2,rho_max,kinetic_energy,enstrophy"
    # second call: value k + 2 at flat index k (linear 0..4, squared 5..9, max 10..14, KE 15, ens 16)
    assert row == [2.0, 6.0, 8.0, 12.0, 17.0, 18.0]
    assert [float(x) for x in lines[2].split(",")] == [0.5] + row
