"""The batch engine (mxm_batch_*): every instance of a batch must give the
same state, bit for bit, as the same problem advanced alone by a Solver
with the same arithmetic form -- whatever the other instances, their step
counts, and instances joining or leaving mid-run.  The Solver path is tied
to the FP64 oracle by test_gpu_parity.py, so this pins the batch to the
oracle as well; one instance is also checked against the oracle directly.
Each batched step replays a HIP graph, so the multi-step comparisons
check every replay against the eager single-solver launches."""
import numpy as np
import pytest

import mojoxm_amd as M
from mojoxm_amd import physics as ph
from mojoxm_amd.batch import Batch, footprint

from test_gpu_parity import random_state, rel_l2

pytestmark = pytest.mark.gpu


def alone(n, L, phys, P, q0, dt, nsteps, precision, bcs=None, limiter=False, eps=0.1):
    with M.Solver(n, L, phys, P=P, bcs=bcs, precision=precision) as s:
        s.set_state(q0)
        if limiter:
            s.enable_cell_limiter(True, eps)
        s.step_ssprk3(dt, nsteps)
        return s.get_state()


def make_cases(rng):
    """(n, L, physics, P, dt, nsteps, bcs, limiter) of a mixed batch: three
    kernel groups, different sizes, parameters and step counts."""
    cases = []
    for k, nn in enumerate([(6, 4), (9, 7), (4, 4)]):
        cases.append((nn, (1.0, 0.9), ph.Euler(flux_type=ph.FLUX_HLLEC, gamma=1.4 - 0.05 * k), 3, 2e-3 * (k + 1),
                      3 + 2 * k, None, False))
    for k, nn in enumerate([(8, 3), (5, 5)]):
        cases.append((nn, (1.0, 1.0), ph.MHDGLM2D(c_h=1.2 + 0.2 * k, alpha_d=(1.2 + 0.2 * k) / 0.18), 2, 1.5e-3,
                       4 + k, None, False))
    cases.append(((3, 2, 2), (1.0, 1.0, 1.0), ph.Advection(1.0, -0.5, 0.3), 2, 3e-3, 5, None, False))
    cases.append(((10, 3), (1.0, 0.3), ph.Euler(flux_type=ph.FLUX_HLLE), 2, 1e-3, 4,
                  M.BoundaryConditions2D(2, 2, 1, 1), True))
    return cases


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_batch_matches_alone(precision):
    rng = np.random.default_rng(3)
    cases = make_cases(rng)
    states = []
    with Batch(precision) as b:
        ids = []
        for (n, L, phys, P, dt, ns, bcs, lim) in cases:
            K = int(np.prod(n)) * (6 if len(n) == 3 else 2)
            NP = (P + 1) * (P + 2) * (P + 3) // 6 if len(n) == 3 else (P + 1) * (P + 2) // 2
            q0 = random_state(phys, K, NP, rng, len(n))
            if precision == M.PREC_FP32_MFMA:
                q0 = q0.astype(np.float32).astype(np.float64)
            states.append(q0)
            ids.append(b.add(n, L, phys, P, q0, dt, ns, bcs=bcs, limiter=lim))
        done = []
        for _ in range(10):
            b.step(1)
            done += b.poll()
        assert sorted(done) == sorted(ids)
        for k, (n, L, phys, P, dt, ns, bcs, lim) in enumerate(cases):
            assert b.status(ids[k]) == ns
            got = b.get_state(ids[k])
            want = alone(n, L, phys, P, states[k], dt, ns, precision, bcs=bcs, limiter=lim)
            assert np.array_equal(got, want), (k, rel_l2(got, want))


def test_batch_joins_and_leaves():
    """An instance added mid-run, and one removed mid-run, leave the others
    (and themselves) unchanged."""
    rng = np.random.default_rng(5)
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    n, L, P, dt = (6, 6), (1.0, 1.0), 3, 2e-3
    q = [random_state(phys, 72, 10, rng, 2) for _ in range(3)]
    with Batch(M.PREC_FP64) as b:
        a = b.add(n, L, phys, P, q[0], dt, 6)
        c = b.add(n, L, phys, P, q[1], dt, 6)
        b.step(2)
        d = b.add(n, L, phys, P, q[2], dt, 3)  # joins after 2 steps of a, c
        b.remove(c)
        b.step(4)
        assert sorted(b.poll()) == sorted([a, d])
        assert np.array_equal(b.get_state(a), alone(n, L, phys, P, q[0], dt, 6, M.PREC_FP64))
        assert np.array_equal(b.get_state(d), alone(n, L, phys, P, q[2], dt, 3, M.PREC_FP64))


@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_batch_running_prefix_shrinks(precision):
    """Many instances of one group with step counts from 1 to 30: the arena
    is ordered longest-running first and the launch grid shrinks to the
    running prefix between launches (several times here, in chunks of 1, 3
    and 8 steps, with one instance joining mid-run).  Every instance must
    still equal its solve alone, and finished ones keep their state."""
    rng = np.random.default_rng(21)
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    sizes = [(3, 3), (5, 4), (4, 6), (7, 3), (2, 2), (6, 6), (3, 5), (8, 2), (4, 4), (5, 5), (2, 7), (6, 3)]
    steps = [30, 1, 7, 13, 2, 24, 5, 17, 9, 3, 28, 11]
    dt, P = 1e-3, 2
    cases = []
    with Batch(precision) as b:
        ids = []
        for nn, ns in zip(sizes, steps):
            q0 = random_state(phys, 2 * nn[0] * nn[1], 6, rng, 2)
            if precision == M.PREC_FP32_MFMA:
                q0 = q0.astype(np.float32).astype(np.float64)
            cases.append((nn, q0, ns))
            ids.append(b.add(nn, (1.0, 1.0), phys, P, q0, dt, ns))
        done = set()
        for chunk in (1, 3, 8, 1, 8, 3):
            b.step(chunk)
            done |= set(b.poll())
        q0 = random_state(phys, 2 * 5 * 3, 6, rng, 2)
        if precision == M.PREC_FP32_MFMA:
            q0 = q0.astype(np.float32).astype(np.float64)
        cases.append(((5, 3), q0, 6))
        ids.append(b.add((5, 3), (1.0, 1.0), phys, P, q0, dt, 6))
        for chunk in (8, 1, 3):
            b.step(chunk)
            done |= set(b.poll())
        assert done == set(ids)
        for i, (nn, q, ns) in zip(ids, cases):
            assert np.array_equal(b.get_state(i), alone(nn, (1.0, 1.0), phys, P, q, dt, ns, precision)), (nn, ns)


def test_batch_against_oracle(oracle):
    rng = np.random.default_rng(9)
    phys = ph.MHDGLM2D(c_h=1.3, alpha_d=1.3 / 0.18)
    n, L, P, dt = (5, 4), (1.0, 1.0), 3, 1e-3
    q0 = random_state(phys, 40, 10, rng, 2)
    with Batch(M.PREC_FP64) as b:
        i = b.add(n, L, phys, P, q0, dt, 3)
        b.add((7, 7), L, phys, P, random_state(phys, 98, 10, rng, 2), dt, 5)
        b.step(5)
        got = b.get_state(i)
    want = oracle.Mesh(2, P, n, L).step(phys.phys_id, phys.params(), q0, dt, 3)
    assert rel_l2(got, want) < 1e-12


def test_batch_memory_matches_footprints():
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    with Batch(M.PREC_FP32_MFMA) as b:
        total = 0
        for nn in [(8, 8), (12, 6)]:
            K = nn[0] * nn[1] * 2
            b.add(nn, (1.0, 1.0), phys, 3, np.ones((K, 10, 5)), 1e-3, 1)
            total += footprint(nn, (1.0, 1.0), phys, 3, M.PREC_FP32_MFMA)
        b.step(1)
        used = b.memory()
    # the group adds its operators and a shared-stride padding of the job tables
    assert total <= used <= total + 64 * 1024


def test_batch_failed_instance_is_switched_off():
    """An instance whose state holds a NaN is flagged by the stage kernel,
    switched off at the end of its first step (its counter stays at 0) and
    reported by poll as failed; the other instances of its group run on,
    each equal to its solve alone."""
    rng = np.random.default_rng(5)
    phys = ph.Maxwell2D(c=1.0)
    bcs = M.BoundaryConditions2D(1, 1, 1, 1)
    dt, P, ns = 2e-3, 3, 6
    qs = [random_state(phys, 2 * 6 * 5, 10, rng, 2) for _ in range(3)]
    bad = qs[1].copy()
    bad[7, 3, 2] = np.nan
    with Batch(M.PREC_FP64) as b:
        ids = [b.add((6, 5), (1.0, 1.0), phys, P, q, dt, ns, bcs=bcs) for q in (qs[0], bad, qs[2])]
        b.step(1)
        first = b.poll()
        assert first == [ids[1]] and b.failed(ids[1])
        assert b.status(ids[1]) == 0
        b.step(ns)
        assert sorted(b.poll()) == sorted([ids[0], ids[2]])
        assert not b.failed(ids[0]) and not b.failed(ids[2])
        for i, q in ((ids[0], qs[0]), (ids[2], qs[2])):
            assert np.array_equal(b.get_state(i), alone((6, 5), (1.0, 1.0), phys, P, q, dt, ns, M.PREC_FP64, bcs=bcs))


@pytest.mark.parametrize("limiter", [False, True])
@pytest.mark.parametrize("precision", [M.PREC_FP64, M.PREC_FP32_MFMA])
def test_batch_reuses_free_ranges_and_compacts(precision, limiter):
    """A removed instance's tiles are taken by a later instance that fits
    them, with no new device memory (no compaction); one that fits nowhere
    compacts the group to exactly its instances' footprints.  Every
    instance equals its solve alone, whichever way it was placed.  With the
    limiter, the factors still pending on the instances' states move with
    them through the compaction, and the new instance in the freed range
    starts with none of the removed one's."""
    rng = np.random.default_rng(9)
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    dt, P = 1e-3, 3
    word = np.float32 if precision == M.PREC_FP32_MFMA else np.float64

    def state(nn):
        return random_state(phys, 2 * nn[0] * nn[1], 10, rng, 2).astype(word).astype(np.float64)

    sizes = {"a": (12, 10), "b": (6, 6), "c": (8, 5), "d": (6, 5), "e": (16, 16)}
    q = {k: state(v) for k, v in sizes.items()}
    steps = {"a": 9, "b": 4, "c": 7, "d": 5, "e": 3}
    fp = {k: footprint(v, (1.0, 1.0), phys, P, precision, limiter) for k, v in sizes.items()}
    with Batch(precision) as b:
        ids = {k: b.add(sizes[k], (1.0, 1.0), phys, P, q[k], dt, steps[k], limiter=limiter) for k in "abc"}
        b.step(2)
        m1 = b.memory()
        assert fp["a"] + fp["b"] + fp["c"] <= m1 <= fp["a"] + fp["b"] + fp["c"] + 64 * 1024
        b.remove(ids.pop("b"))
        ids["d"] = b.add(sizes["d"], (1.0, 1.0), phys, P, q["d"], dt, steps["d"], limiter=limiter)  # b's tiles
        b.step(3)
        assert b.memory() == m1
        ids["e"] = b.add(sizes["e"], (1.0, 1.0), phys, P, q["e"], dt, steps["e"], limiter=limiter)  # compaction
        b.step(2)
        m2 = b.memory()
        assert fp["a"] + fp["c"] + fp["d"] + fp["e"] <= m2 <= fp["a"] + fp["c"] + fp["d"] + fp["e"] + 64 * 1024
        b.step(10)
        assert set(b.poll()) == set(ids.values())
        for k, i in ids.items():
            assert np.array_equal(b.get_state(i), alone(sizes[k], (1.0, 1.0), phys, P, q[k], dt, steps[k],
                                                        precision, limiter=limiter)), k
    if limiter:  # the limiter acted on these states
        args = (sizes["a"], (1.0, 1.0), phys, P, q["a"], dt, steps["a"], precision)
        assert not np.array_equal(alone(*args, limiter=True), alone(*args))


def test_batch_lagged_poll_reports_everything():
    """poll(lagged=True) waits only for the work before the most recent
    step() call; polling that way after every chunk still reports every
    instance once, and the results equal the solves alone."""
    rng = np.random.default_rng(13)
    phys = ph.ShallowWater2D(hll=True)
    dt, P = 1e-3, 2
    steps = [3, 11, 6, 17, 1]
    qs = [random_state(phys, 2 * 5 * 4, 6, rng, 2) for _ in steps]
    with Batch(M.PREC_FP64) as b:
        ids = [b.add((5, 4), (1.0, 1.0), phys, P, q, dt, ns) for q, ns in zip(qs, steps)]
        seen = []
        for _ in range(8):
            b.step(3)
            seen += b.poll(lagged=True)
        seen += b.poll()
        assert sorted(seen) == sorted(ids)
        for i, q, ns in zip(ids, qs, steps):
            assert np.array_equal(b.get_state(i), alone((5, 4), (1.0, 1.0), phys, P, q, dt, ns, M.PREC_FP64))


def test_batch_without_graphs_matches_alone():
    """Direct stream launches (graphs off, the mode for kernel tracers) run
    the same kernels: every instance still equals its solve alone."""
    rng = np.random.default_rng(17)
    cases = make_cases(rng)[:4]
    with Batch(M.PREC_FP32_MFMA) as b:
        b.graphs(False)
        ids, qs = [], []
        for n, L, phys, P, dt, ns, bcs, lim in cases:
            q0 = random_state(phys, int(np.prod(n)) * (6 if len(n) == 3 else 2),
                              (P + 1) * (P + 2) * (P + 3) // 6 if len(n) == 3 else (P + 1) * (P + 2) // 2, rng,
                              len(n)).astype(np.float32).astype(np.float64)
            qs.append(q0)
            ids.append(b.add(n, L, phys, P, q0, dt, ns, bcs=bcs, limiter=lim))
        b.step(max(c[5] for c in cases))
        assert sorted(b.poll()) == sorted(ids)
        for i, q0, (n, L, phys, P, dt, ns, bcs, lim) in zip(ids, qs, cases):
            assert np.array_equal(b.get_state(i), alone(n, L, phys, P, q0, dt, ns, M.PREC_FP32_MFMA, bcs=bcs,
                                                        limiter=lim))


def test_batch_headroom_places_later_instances_without_a_rebuild():
    """With headroom, an arena rebuilt because an instance did not fit gets
    spare tiles and slots: the next small instance is placed in them with no
    new device memory.  Every instance equals its solve alone."""
    rng = np.random.default_rng(23)
    phys = ph.Euler(flux_type=ph.FLUX_HLLEC)
    dt, P = 1e-3, 2
    sizes = {"a": (10, 8), "c": (6, 6), "d": (4, 4)}
    q = {k: random_state(phys, 2 * v[0] * v[1], 6, rng, 2) for k, v in sizes.items()}
    steps = {"a": 7, "c": 5, "d": 3}
    fp = {k: footprint(v, (1.0, 1.0), phys, P, M.PREC_FP64) for k, v in sizes.items()}
    with Batch(M.PREC_FP64) as b:
        b.headroom(0.5)
        ids = {"a": b.add(sizes["a"], (1.0, 1.0), phys, P, q["a"], dt, steps["a"])}
        b.step(1)  # first arena: exact
        assert b.memory() <= fp["a"] + 64 * 1024
        ids["c"] = b.add(sizes["c"], (1.0, 1.0), phys, P, q["c"], dt, steps["c"])
        b.step(1)  # does not fit: rebuilt with headroom
        m2 = b.memory()
        assert m2 > fp["a"] + fp["c"] + 64 * 1024
        ids["d"] = b.add(sizes["d"], (1.0, 1.0), phys, P, q["d"], dt, steps["d"])
        b.step(1)  # placed in the headroom
        assert b.memory() == m2
        b.step(10)
        assert set(b.poll()) == set(ids.values())
        for k, i in ids.items():
            assert np.array_equal(b.get_state(i), alone(sizes[k], (1.0, 1.0), phys, P, q[k], dt, steps[k],
                                                        M.PREC_FP64)), k
