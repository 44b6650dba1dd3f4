"""Bit-reproducible resume: a run of k sweeps, saved through the resume file (fixed-point state
plus the record: seed, sweep index, step sizes, adaptation counters, running energy) and
continued for 20 - k sweeps in a new context, ends bit for bit where 20 sweeps in one context end:
positions, orientations, step sizes and running energy, in FP64 and FP32, NVT with a colloid
and NPT, one replica and two replicas resumed in reverse order."""
import numpy as np
import pytest

from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def _params(precision, npt, seed, N=1300, rho=0.25):
    kw = dict(npt=True, pressure=2.0, npt_max_dlnv=0.01) if npt else dict(colloid_radius=3.0)
    return m.gb_params(N, rho, 1.0, precision=precision, seed=seed, energy_check_interval=0,
                       max_displacement=0.8, max_rotation=0.1, adapt_moves=2000, **kw)


def _start(npt, seed, N=1300, rho=0.25):
    return layered_system(N, rho, colloid_radius=0.0 if npt else 3.0, seed=seed)


def _same(a, b, what):
    (qa, ua, La), (qb, ub, Lb) = a, b
    for d in range(3):
        assert np.array_equal(qa[d], qb[d]), f"{what}: positions, axis {d}"
        assert ua[d].tobytes() == ub[d].tobytes(), f"{what}: orientations, axis {d}"
    assert La.tobytes() == Lb.tobytes(), f"{what}: box"


@pytest.mark.parametrize("precision,npt", [(64, False), (32, False), (64, True), (32, True)])
def test_resumed_sweeps_equal_twenty_straight(tmp_path, precision, npt):
    print(f"resume precision={precision} npt={npt}", flush=True)
    p = _params(precision, npt, seed=4321)
    q, u, L = _start(npt, seed=12)
    with m.Context(p) as a:
        a.set_state_fixed(q, u, L)
        a.sweep(20)
        want, want_rec = a.get_state_fixed(), a.get_resume().as_dict()
    assert want_rec["sweep_index"] == 20 and want_rec["adaptations"] >= 5
    path = str(tmp_path / "fortSphere.1.mcgb")
    with m.Context(p) as b:
        b.set_state_fixed(q, u, L)
        k = 0
        while k < 7 or b.get_resume().adapt_attempted == 0:   # save inside an adaptation window
            b.sweep(1)
            k += 1
        assert k < 20
        qb, ub, Lb = b.get_state_fixed()
        rec = b.get_resume()
        assert rec.sweep_index == k and rec.seed == 4321
        m.write_resume(path, qb, ub, Lb, rec)
    q2, u2, L2, rec2 = m.read_resume(path, 1300)
    with m.Context(p) as c:
        c.resume_state(q2, u2, L2, rec2)
        assert c.get_resume().as_dict() == rec2.as_dict()
        c.sweep(20 - k)
        got, got_rec = c.get_state_fixed(), c.get_resume().as_dict()
    _same(want, got, "resumed run")
    assert got_rec == want_rec
    if npt:
        assert not np.array_equal(want[2], L)             # the box moved


def test_two_replicas_resumed_in_reverse_order(tmp_path):
    ps = [_params(64, True, seed=s) for s in (71, 72)]
    starts = [_start(True, seed=s) for s in (31, 32)]

    def fresh():
        ctx = m.Context(ps)
        for r, (q, u, L) in enumerate(starts):
            ctx.set_state_fixed(q, u, L, replica=r)
        return ctx

    with fresh() as a:
        a.sweep(20)
        want = [(a.get_state_fixed(r), a.get_resume(r).as_dict()) for r in range(2)]
    with fresh() as b:
        b.sweep(9)
        for r in range(2):
            m.write_resume(str(tmp_path / f"r{r}"), *b.get_state_fixed(r), b.get_resume(r))
    with m.Context(ps) as c:
        for r in (1, 0):                     # resuming replica 0 leaves replica 1's chain alone
            q, u, L, rec = m.read_resume(str(tmp_path / f"r{r}"), 1300)
            c.resume_state(q, u, L, rec, replica=r)
        c.sweep(11)
        for r in range(2):
            _same(want[r][0], c.get_state_fixed(r), f"replica {r}")
            assert c.get_resume(r).as_dict() == want[r][1], f"replica {r}"


def test_resume_rejects_a_foreign_record():
    p = _params(64, False, seed=5)
    q, u, L = _start(False, seed=5)
    with m.Context(p) as c:
        c.set_state_fixed(q, u, L)
        c.sweep(2)
        rec = c.get_resume()
        st = c.get_state_fixed()
        rec.seed = 6
        with pytest.raises(m.McgbError, match="seed"):
            c.resume_state(*st, rec)
        rec.seed = 5
        rec.sweep_index = 2**32
        with pytest.raises(m.McgbError, match="out of range"):
            c.resume_state(*st, rec)
        rec.sweep_index = 2
        rec.max_displacement = 0.0
        with pytest.raises(m.McgbError, match="out of range"):
            c.resume_state(*st, rec)
        assert c.get_resume().sweep_index == 2              # nothing was uploaded


@pytest.mark.parametrize("precision", [32, 64])
def test_rejected_set_state_leaves_the_replica_unchanged(precision):
    """A set_state that fails its checks (a zero orientation; a current step size of half the
    new box) leaves state, energy, counters and the chain exactly as they were: the context goes
    on as a twin that never saw the calls."""
    N, rho = 1300, 0.25
    q, u, L = layered_system(N, rho, seed=4)
    p = m.gb_params(N, rho, 1.0, precision=precision, seed=77, energy_check_interval=0)
    with m.Context(p) as a, m.Context(p) as b:
        for ctx in (a, b):
            ctx.set_state_fixed(q, u, L)
            ctx.sweep(3)
            ctx.set_step_sizes(0.46 * L[0], 0.1)
        q0, u0, L0 = a.get_state_fixed()
        st0 = a.stats()
        bad = [x.copy() for x in u0]
        for d in range(3):
            bad[d][7] = 0.0
        with pytest.raises(m.McgbError) as e:
            a.set_state_fixed(q0, bad, L0)
        assert e.value.code == m.ERR_STATE
        with pytest.raises(m.McgbError) as e:       # 0.46 L >= half of a box of 0.9 L
            a.set_state_fixed([x[::-1].copy() for x in q0], u0, 0.9 * L0)
        assert e.value.code == m.ERR_GEOMETRY
        q1, u1, L1 = a.get_state_fixed()
        for d in range(3):
            assert np.array_equal(q1[d], q0[d]) and np.array_equal(u1[d], u0[d])
        assert np.array_equal(L1, L0) and a.stats() == st0
        for ctx in (a, b):
            ctx.set_step_sizes(0.1, 0.1)
            ctx.sweep(2)
        qa, ua, _ = a.get_state_fixed()
        qb, ub, _ = b.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
        assert a.stats() == b.stats()
