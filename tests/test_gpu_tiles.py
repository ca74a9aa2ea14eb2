"""Tile sizes of the stage kernel (mxm_dg_tile_config, and the tiling choice
of mxm_dg_autotune).

The FP32-MFMA form on triangles at P = 3 has two tilings: 24 elements on
256 threads, and the wide tile of 32 elements on 320 threads (5 waves).
An element's arithmetic does not depend on its slot in a tile: a face
inside a tile is evaluated once and negated for the partner, a face
between tiles by each side with the same left and right states and the
same normal, and each MFMA output column sums the same products in the
same order.  So both tilings must give the same bits, for every 2D
physics, with the limiter, and on meshes whose edge tiles are partly
filled.  The reductions over tiles agree exactly (largest signal speed,
max |q|) or to the compensated sums' rounding (integrals)."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph

from test_gpu_parity import CASES_2D, random_state

pytestmark = pytest.mark.gpu

NARROW, WIDE = (False, 24, 256), (True, 32, 320)


def advance(n, phys, bc, q0, wide, nsteps=4, limiter=False, P=3):
    bcs = M.BoundaryConditions2D(*bc) if bc else None
    with M.Solver(n, (1.0, 0.9), phys, P=P, bcs=bcs, precision=M.PREC_FP32_MFMA) as s:
        s.tile = wide
        assert s.tile == (WIDE if wide else NARROW)
        s.set_state(q0)
        if limiter:
            s.enable_cell_limiter(True, 0.1)
        s.step_ssprk3(0.2 / max(n) / (2 * P + 1) / 2.0, nsteps)
        return s.get_state(), s.max_wave_speed(), s.diagnostics().flat()


def seeded(phys, n, seed, P=3):
    rng = np.random.default_rng(seed)
    q = random_state(phys, 2 * n[0] * n[1], (P + 1) * (P + 2) // 2, rng, 2)
    return q.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("case", range(len(CASES_2D)))
def test_wide_tile_gives_the_same_bits(case):
    # 9 x 7 squares: the edge tiles of both tilings are partly filled
    phys, bc = CASES_2D[case]
    n = (9, 7)
    q0 = seeded(phys, n, 100 + case)
    a, sa, da = advance(n, phys, bc, q0, False)
    b, sb, db = advance(n, phys, bc, q0, True)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
    assert sa == sb
    nc = phys.NUM_COMPONENTS
    assert np.array_equal(da[2 * nc:3 * nc], db[2 * nc:3 * nc])  # max |q_c|
    np.testing.assert_allclose(db[:2 * nc], da[:2 * nc], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(db[3 * nc:], da[3 * nc:], rtol=1e-12, equal_nan=True)


@pytest.mark.parametrize("phys", [ph.Euler(flux_type=ph.FLUX_HLLE), ph.Euler2D(hllc=True)])
def test_wide_tile_gives_the_same_bits_limited(phys):
    n = (12, 10)
    q0 = seeded(phys, n, 7)
    a, _, _ = advance(n, phys, [1, 1, 2, 2], q0, False, nsteps=6, limiter=True)
    b, _, _ = advance(n, phys, [1, 1, 2, 2], q0, True, nsteps=6, limiter=True)
    assert np.array_equal(a, b)


def test_tile_changes_keep_the_state():
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    n = (10, 6)
    q0 = seeded(phys, n, 3)
    with M.Solver(n, (1.0, 1.0), phys, P=3, precision=M.PREC_FP32_MFMA) as s:
        assert s.tile == NARROW
        s.set_state(q0)
        for wide in (True, True, False, True):
            s.tile = wide
            assert s.tile == (WIDE if wide else NARROW)
            assert np.array_equal(s.get_state(), q0)


def test_autotune_keeps_the_state_and_the_bits():
    # the C2 vortex mesh at 32^2: whichever tiling and configuration the
    # autotuner keeps, the state it timed on is unchanged and the steps
    # that follow give the bits of the untuned solver
    pb = M.problems.euler_vortex_2d(N=32)
    mesh = M.LocalMesh2D(*pb.n, *pb.L, P=pb.P)
    q0 = pb.ic(mesh.elem_node_xyz).astype(np.float32).astype(np.float64)
    out = []
    for tune in (False, True):
        with M.Solver(pb.n, pb.L, pb.physics, P=pb.P, precision=M.PREC_FP32_MFMA) as s:
            s.set_state(q0)
            if tune:
                cfg, us = s.autotune(pb.dt, reps=5)
                assert s.launch_config == cfg and len(us) == 4 and min(us) > 0
                assert s.tile in (NARROW, WIDE)
                assert np.array_equal(s.get_state(), q0)
            s.step_ssprk3(pb.dt, 5)
            out.append(s.get_state())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("dim,P,prec", [(2, 3, M.PREC_FP64), (2, 2, M.PREC_FP32_MFMA), (3, 3, M.PREC_FP32_MFMA)])
def test_no_wide_tile_elsewhere(dim, P, prec):
    n = (4, 4) if dim == 2 else (2, 2, 2)
    phys = ph.Euler(flux_type=ph.FLUX_RUSANOV)
    with M.Solver(n, (1.0,) * dim, phys, P=P, precision=prec) as s:
        wide, e, t = s.tile
        assert not wide and t == 256
        s.tile = False
        with pytest.raises(M.MxmError):
            s.tile = True
        assert s.tile == (wide, e, t)
