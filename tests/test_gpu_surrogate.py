"""The learned first stage on the GPU: matrix-core inference and training gradients against the
CPU oracle of the network (oracle/surrogate_ref.py), emitted pair energies against the FP64
pair oracle and the FP64 total energy, and delayed acceptance against the plain sweep."""
import numpy as np
import pytest

from oracle import oracle_ctypes as oc
from oracle import surrogate_ref as sr
from parellel_mc_mpi_for_liquidcrystalsystems_amd import mcgb as m
from tests.systems import layered_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def need_gpu():
    if m.device_count() < 1:
        pytest.fail("no GPU visible to HIP: the gpu tests need an MI355X")


def small_ctx(N=2000, rho=0.3, T=1.0, R=0.0, precision=32, seed=0, pseed=11):
    p = m.gb_params(N, rho, T, colloid_radius=R, precision=precision, seed=pseed,
                    energy_check_interval=0)
    q, u, L = layered_system(N, rho, colloid_radius=R, seed=seed)
    ctx = m.Context(p)
    ctx.set_state_fixed(q, u, L)
    return ctx


def features(n, seed):
    rng = np.random.default_rng(seed)
    return np.stack([rng.uniform(0.5, 4.0, n), rng.uniform(-1, 1, n), rng.uniform(-1, 1, n),
                     rng.uniform(-1, 1, n)], axis=1).astype(np.float32)


def random_params(seed, scale=0.3):
    return np.random.default_rng(seed).normal(scale=scale, size=sr.NPARAM).astype(np.float32)


def test_matrix_core_inference_matches_oracle():
    with small_ctx() as ctx:
        for p in (random_params(4), None):
            if p is None:
                ctx.surrogate_init(7)
                p = ctx.surrogate_get()
            else:
                ctx.surrogate_set(p)
            f = features(5000, 9)
            ref = sr.forward(p, f)
            refb = sr.forward_bf16(p, f)
            scale = np.abs(ref).max() + 1e-3
            y32 = ctx.surrogate_eval(f, mfma=False)
            np.testing.assert_allclose(y32, ref, rtol=1e-4, atol=1e-5 * scale)
            ymf = ctx.surrogate_eval(f, mfma=True)
            # against the BF16-rounded network (any operand-layout slip gives O(scale) errors)
            err = np.abs(ymf - refb) / scale
            assert np.median(err) < 1e-4 and err.max() < 1e-2, (np.median(err), err.max())
            # and within BF16 accuracy of the exact network
            assert np.abs(ymf - ref).max() < 0.05 * scale


def test_training_gradient_matches_oracle():
    with small_ctx() as ctx:
        p = random_params(5)
        f = features(1000, 2)                        # 16 workgroups, the last one partial
        t = np.random.default_rng(3).normal(scale=2.0, size=1000).astype(np.float32)
        t[:30] = 1.0e30                              # capped pairs: clipped targets
        g, loss = ctx.surrogate_grad(f, t, clip=10.0, params=p)
        lref, gref = sr.loss_and_grad(p.astype(np.float64), f.astype(np.float64),
                                      t.astype(np.float64), 10.0)
        assert loss == pytest.approx(lref, rel=1e-4)
        np.testing.assert_allclose(g, gref, rtol=2e-3, atol=2e-5 * np.abs(gref).max())


def test_emitted_pairs_are_the_exact_pair_energies():
    N = 1500
    with small_ctx(N=N, precision=64) as ctx:
        ctx.sweep(3)
        ctx.pairs_reserve(400000)
        n = ctx.emit_pairs(stride=1)
        rec = ctx.pairs_get()
        assert rec.shape == (n, 5)
        # every ordered in-cutoff pair once: the energies sum to twice the FP64 total
        e = ctx.total_energy()
        capped = rec[:, 4] >= 1e29
        assert capped.sum() == 2 * e[1]
        assert rec[~capped, 4].astype(np.float64).sum() == pytest.approx(
            2 * e[3], rel=1e-5, abs=1e-3)
        # count by brute force over the minimum images
        q, u, L = ctx.get_state_fixed()
        s = L[0] / 2.0 ** 32
        cnt = 0
        for i in range(N):
            d = [(q[k] - q[k][i]).astype(np.int32).astype(np.float64) * s for k in range(3)]
            r2 = d[0] ** 2 + d[1] ** 2 + d[2] ** 2
            cnt += int((r2 < 16.0).sum()) - 1
        assert n == cnt
        assert (rec[:, 0] < 4.0).all() and (np.abs(rec[:, 1:4]) <= 1.0 + 1e-6).all()
        # energies from the features through the oracle's pair function (sample)
        c = oc.consts()
        rng = np.random.default_rng(0)
        sel = rng.choice(np.flatnonzero(~capped), 1500, replace=False)
        r, a, b, cc, U = (rec[sel, k].astype(np.float64) for k in range(5))
        ok = 1 - a * a > 1e-3
        y = (cc - a * b) / np.sqrt(np.maximum(1 - a * a, 1e-300))
        z2 = 1 - b * b - y * y
        ok &= z2 > 1e-4
        ei = np.stack([a, np.sqrt(np.maximum(1 - a * a, 0)), 0 * a], axis=1)[ok]
        ej = np.stack([b, y, np.sqrt(np.maximum(z2, 0))], axis=1)[ok]
        dr = np.stack([r, 0 * r, 0 * r], axis=1)[ok]
        uo, co = oc.gb_pair(c, ei, ej, dr)
        assert ok.sum() > 1000 and not co.any()
        fin = np.isfinite(uo)
        assert fin.mean() > 0.99
        # features are FP32: the pair's condition number bounds the difference
        rho = 1.0 / np.maximum(r[ok] - 3.0 + 1.0, 0.3)
        err = np.abs(U[ok] - uo) / (1e-4 * (1 + 4 * rho ** 12))
        assert err[fin].max() <= 1.0, (err[fin].max(), U[ok][fin][np.argmax(err[fin])])


def _equilibrated(N=2000, T=1.0, sweeps=200, precision=32, seed=0):
    ctx = small_ctx(N=N, T=T, precision=precision, seed=seed)
    ctx.sweep(sweeps)
    return ctx


def test_zero_network_delayed_acceptance_is_the_plain_sweep():
    # stage 1 with a zero surrogate always passes (dE~ = 0 <= -T log U'); stage 2 then is the
    # plain Metropolis test with the same draw, so the chains are identical
    for prec in (64, 32):
        a = small_ctx(N=1500, precision=prec)
        b = small_ctx(N=1500, precision=prec)
        with a, b:
            b.surrogate_set(np.zeros(sr.NPARAM, dtype=np.float32))
            b.da_enable(True)
            a.sweep(4)
            b.sweep(4)
            qa, ua, _ = a.get_state_fixed()
            qb, ub, _ = b.get_state_fixed()
            for d in range(3):
                assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
            sa, sb = a.stats(), b.stats()
            assert sa["accepted"] == sb["accepted"] and sb["da_stage1_rejected"] == 0
            assert sb["da_stage2_rejected"] == (sb["attempted"] - sb["accepted"] -
                                                sb["rejected_left_cell"] - sb["rejected_hard_core"])


def test_judge_mode_does_not_change_decisions():
    a = small_ctx(N=1500)
    b = small_ctx(N=1500)
    with a, b:
        for c in (a, b):
            c.surrogate_init(3)
        a.da_enable(True, judge=False)
        b.da_enable(True, judge=True)
        a.sweep(3)
        b.sweep(3)
        qa, ua, _ = a.get_state_fixed()
        qb, ub, _ = b.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
        ha, hb = a.da_histogram(), b.da_histogram()
        sb = b.stats()
        evaluated = sb["attempted"] - sb["rejected_left_cell"] - sb["rejected_hard_core"]
        assert hb.sum() <= evaluated and hb.sum() > ha.sum()


def _means(ctx, sweeps):
    ctx.series_clear()
    ctx.sweep(sweeps)
    e, p2, _ = ctx.series()
    out = []
    for x in (e, p2):
        tau, _ = m.tau_int(x)
        out.append((x.mean(), x.std() * np.sqrt(max(tau, 1.0) / len(x))))
    return out


def test_delayed_acceptance_with_a_poor_network_samples_the_same_distribution():
    # an untrained network changes only the efficiency: <E>/N and P2 agree with the plain
    # sweep within 2 combined standard errors (autocorrelation-corrected)
    a = _equilibrated()
    b = _equilibrated()
    with a, b:
        b.surrogate_init(21)
        b.da_enable(True)
        (ea, sea), (pa, spa) = _means(a, 3000)
        (eb, seb), (pb, spb) = _means(b, 3000)
        assert abs(ea - eb) < 2 * np.hypot(sea, seb), (ea, eb, sea, seb)
        assert abs(pa - pb) < 2 * np.hypot(spa, spb), (pa, pb, spa, spb)
        s = b.stats()
        assert s["da_stage1_rejected"] > 0 and s["da_stage2_rejected"] > 0


def test_trained_network_screens_out_exact_evaluations():
    # what the trained network gains over the plain sweep: a share of the exact evaluations is
    # skipped (stage-1 rejections) while few stage-1 acceptances are overturned by stage 2
    with _equilibrated(sweeps=100) as ctx:
        ctx.pairs_reserve(2_000_000)
        for k in range(4):
            ctx.sweep(25)
            ctx.emit_pairs(stride=2, offset=k % 2)
        ctx.surrogate_init(1)
        losses = ctx.surrogate_train(20000, batch=4096, lr=3e-3, clip=10.0, seed=5, wfloor=1.0)
        assert losses[-200:].mean() < 0.02 * losses[:10].mean()
        ctx.reset_counters()
        ctx.sweep(20)
        plain = ctx.stats()
        ctx.da_enable(True, judge=True)
        ctx.reset_counters()
        ctx.da_histogram(clear=True)
        ctx.sweep(20)
        s = ctx.stats()
        h = ctx.da_histogram()
    evaluated = s["attempted"] - s["rejected_left_cell"] - s["rejected_hard_core"]
    s1, s2 = s["da_stage1_rejected"], s["da_stage2_rejected"]
    plain_rej = 1 - plain["accepted"] / (plain["attempted"] - plain["rejected_left_cell"] -
                                         plain["rejected_hard_core"])
    assert s1 / evaluated > 0.6 * plain_rej          # most rejections need no exact energy
    assert s2 / max(1, evaluated - s1) < 0.4         # stage 2 overturns a minority of stage 1
    centre = h[24:40].sum()                          # |dE~ - dE| < 1 kT for most moves
    assert centre > 0.55 * h.sum()


@pytest.mark.parametrize("poison", ["overflow", "huge"])
def test_poisoned_network_falls_back_to_the_plain_kernel(poison):
    """With the guard on, a poisoned network is switched off after one window and the rest of
    the sweeps run the plain kernel: the chain moves again, and its running energy equals the
    FP64 recompute.  Poisons: weights x1e30, finite but overflowing in the hidden layers (what
    a diverged training run leaves): non-finite outputs, stage 1 passes nothing; weights x1000:
    the exact stage overturns nearly every stage-1 acceptance.  (mcgb_surrogate_set itself
    refuses non-finite weights.)"""
    with small_ctx(N=1500) as ctx:
        ctx.surrogate_init(3)
        w = ctx.surrogate_get()
        w *= 1e30 if poison == "overflow" else 1000.0
        ctx.surrogate_set(w)
        if poison == "overflow":
            # the poison forces non-finite surrogate energies (inf - inf: NaN) into stage 1,
            # where the ordered compare rejects them
            y = ctx.surrogate_eval(features(4096, 5), mfma=True)
            bad = np.mean(~np.isfinite(y))
            print(f"non-finite surrogate outputs: {bad:.3f} (NaN {np.mean(np.isnan(y)):.3f})",
                  flush=True)
            assert bad > 0.9
        ctx.da_enable(True)
        ctx.da_guard(5, 0.5, 0.02)
        ctx.sweep(5)                               # one window: delayed acceptance
        s5 = ctx.stats()
        st = ctx.da_status()
        print(f"poison {poison}: {st}", flush=True)
        assert not st["on"] and st["fallbacks"] == 1, st
        if poison == "overflow":
            assert st["stage1_acceptance"] < 0.02
        else:
            assert st["stage2_rejection"] > 0.9
        ctx.sweep(10)                              # the plain kernel
        s = ctx.stats()
        assert s["da_stage1_rejected"] == s5["da_stage1_rejected"]
        assert s["da_stage2_rejected"] == s5["da_stage2_rejected"]
        assert s["accepted"] - s5["accepted"] > 0.2 * (s["attempted"] - s5["attempted"])
        tot = ctx.total_energy()
        assert s["cap_pairs"] == tot[1]
        assert s["energy_finite"] == pytest.approx(tot[3], rel=1e-6, abs=1e-3)
        q, u, _ = ctx.get_state()
        assert all(np.all(np.isfinite(x)) for x in (*q, *u))


def test_guard_keeps_a_working_network_and_the_plain_chain_for_a_zero_network():
    """The guard judges from counts only: a zero network (stage 1 passes everything, stage 2 is
    the plain test) rejects at stage 2 what plain Metropolis rejects, so a loose bound keeps
    it on and a tight one makes it fall back -- to the identical chain."""
    a = small_ctx(N=1500)
    b = small_ctx(N=1500)
    with a, b:
        b.surrogate_set(np.zeros(sr.NPARAM, dtype=np.float32))
        b.da_enable(True)
        b.da_guard(2, 0.99, 0.0)
        a.sweep(6)
        b.sweep(6)
        assert b.da_status()["on"] and b.da_status()["fallbacks"] == 0
        b.da_guard(2, 0.1, 0.0)
        a.sweep(4)
        b.sweep(4)
        st = b.da_status()
        assert not st["on"] and st["fallbacks"] == 1 and st["stage2_rejection"] > 0.1
        qa, ua, _ = a.get_state_fixed()
        qb, ub, _ = b.get_state_fixed()
        for d in range(3):
            assert np.array_equal(qa[d], qb[d]) and np.array_equal(ua[d], ub[d])
