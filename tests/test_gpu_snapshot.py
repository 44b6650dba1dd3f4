"""Asynchronous sweeps and device-side snapshots: a snapshot equals mcgb_get_state at the point it
was enqueued (bit for bit, FP32 and FP64 storage, replicas, NPT boxes), sweeps enqueued after it
do not change it, and asynchronous sweeps give the same trajectory as blocking ones."""
import numpy as np
import pytest

from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def _ctx(precision, npt=False, seed=11, N=3000, nrep=1):
    ps, states = [], []
    for i in range(nrep):
        q, u, L = layered_system(N + 500 * i, 0.25, seed=seed + i)
        ps.append(m.gb_params(N + 500 * i, 0.25, 1.0, precision=precision, seed=700 + seed,
                              npt=npt, pressure=1.5))
        states.append((q, u, L))
    ctx = m.Context(ps)
    for i, (q, u, L) in enumerate(states):
        ctx.set_state_fixed(q, u, L, replica=i)
    return ctx


def _same(a, b):
    (pa, ua, ba), (pb, ub, bb) = a, b
    return (all(np.array_equal(x, y) for x, y in zip(pa, pb)) and
            all(np.array_equal(x, y) for x, y in zip(ua, ub)) and np.array_equal(ba, bb))


@pytest.mark.parametrize("precision", [32, 64])
def test_snapshot_equals_get_state(precision):
    print(f"snapshot precision {precision}", flush=True)
    with _ctx(precision, nrep=2) as ctx:
        ctx.sweep(3)
        for r in range(2):
            ctx.snapshot_begin(replica=r)
            snap = ctx.snapshot_wait(replica=r)
            assert _same(snap, ctx.get_state(replica=r)), f"replica {r}"
            # the device conversion is the host's: x = q * (L / 2^32) in IEEE double
            q, u, L = ctx.get_state_fixed(replica=r)
            host = ([q[d].astype(np.float64) * (L[d] / 2.0 ** 32) for d in range(3)], u, L)
            assert _same(snap, host), f"replica {r}"


def test_get_state_keeps_a_snapshot_in_flight():
    with _ctx(32) as ctx:
        ctx.sweep(2)
        ctx.snapshot_begin()
        ctx.sweep(3)
        now = ctx.get_state()                  # its own pinned record
        snap = ctx.snapshot_wait()
        assert not np.array_equal(snap[0][0], now[0][0])
        assert _same(now, ctx.get_state())


def test_snapshot_is_taken_before_later_sweeps():
    with _ctx(32) as ctx:
        ctx.sweep(2)
        before = ctx.get_state()
        ctx.snapshot_begin()
        ctx.sweep_async(6)
        snap = ctx.snapshot_wait()              # the copy, not the six sweeps after it
        assert _same(snap, before)
        ctx.sync()
        after = ctx.get_state()
        assert not np.array_equal(after[0][0], before[0][0])
        ctx.snapshot_begin()
        assert _same(ctx.snapshot_wait(), after)
        with pytest.raises(m.McgbError):        # nothing in flight any more
            ctx.snapshot_wait()


def test_async_sweeps_match_blocking_sweeps():
    with _ctx(32) as a, _ctx(32) as b:
        a.sweep(9)
        for _ in range(3):
            b.sweep_async(3)
        b.sync()
        assert _same(a.get_state(), b.get_state())
        sa, sb = a.stats(), b.stats()
        assert sa["accepted"] == sb["accepted"] and sa["energy"] == sb["energy"]


def test_snapshot_box_follows_npt():
    with _ctx(64, npt=True, N=4000) as ctx:
        ctx.sweep(20)
        ctx.snapshot_begin()
        pos, u, box = ctx.snapshot_wait()
        ref = ctx.get_state()
        assert np.array_equal(box, ref[2]) and _same((pos, u, box), ref)
        assert ctx.stats()["volume_accepted"] > 0
