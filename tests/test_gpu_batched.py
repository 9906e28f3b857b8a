"""Batched heterogeneous simulations (GPU tests).

Every instance of a batch must reproduce its stand-alone preset run (same
Philox key and epoch, so the same paths): the batched FP64 terminal mean and
variance agree with pmean/pvar of the stand-alone process and with the
oracle's FP64 twin to FP32-state tolerance.  The captured graph is replayed
several times, with eager work in between, and every replay must equal the
eager launch bit for bit (the reduction order is fixed).
"""
import numpy as np
import pytest

import sdepy_amd as sa
from sdepy_amd import batched as B
from oracle import euler as ora

pytestmark = pytest.mark.gpu


def _specs(seed):
    g = np.random.default_rng(seed)
    specs, twins = [], []
    for i in range(24):
        paths = int(g.integers(50, 3000))
        steps = int(g.integers(12, 80))
        rs = 1000 + i
        if i % 3 == 0:
            kw = dict(x0=1., mu=g.uniform(0, .1), sigma=g.uniform(.1, .5))
            mk = lambda kw=kw, paths=paths, steps=steps, rs=rs: sa.lognorm_process(  # noqa: E731
                paths=paths, steps=steps, dtype=np.float32,
                rng=np.random.default_rng(rs), **kw)
            kind = 'lognorm'
        elif i % 3 == 1:
            kw = dict(factors=2, x0=np.array([[.01], [-.02]]), theta=0.,
                      k=g.uniform(.01, .3, (2, 1)), sigma=g.uniform(.005, .02, (2, 1)),
                      rho=g.uniform(-.9, .9))
            mk = lambda kw=kw, paths=paths, steps=steps, rs=rs: sa.hull_white_process(  # noqa: E731
                paths=paths, steps=steps, dtype=np.float32,
                rng=np.random.default_rng(rs), **kw)
            kind = 'hull_white'
        else:
            kw = dict(x0=.5, theta=1., k=2., sigma=.3)
            mk = lambda kw=kw, paths=paths, steps=steps, rs=rs: sa.ornstein_uhlenbeck_process(  # noqa: E731
                paths=paths, steps=steps, dtype=np.float32,
                rng=np.random.default_rng(rs), **kw)
            kind = 'ou'
        specs.append((mk(), (0., 1.)))
        twins.append((mk, kind, kw, paths, steps))
    return specs, twins


def test_batch_matches_standalone_and_oracle(gpu):
    specs, twins = _specs(5)
    bt = B.BatchedTerminalMoments(specs, autotune=False, graph=False)
    mean, var = bt.run()
    for i, (mk, kind, kw, paths, steps) in enumerate(twins):
        p = mk()
        x = p((0., 1.))
        m_ref = float(x.pmean(dtype=np.float64)[-1].reshape(-1)[0])
        v_ref = float(x.pvar(dtype=np.float64)[-1].reshape(-1)[0])
        scale = max(abs(m_ref), np.sqrt(v_ref), 1e-3)
        assert abs(mean[i] - m_ref) < 1e-5 * scale, (i, kind)
        assert abs(var[i] - v_ref) < 1e-4 * v_ref + 1e-12, (i, kind)
        # oracle FP64 twin of the same draws
        key = p.sources['dw'].key
        vs = (2,) if kind == 'hull_white' else ()
        corr = kw.get('rho')
        src = ora.PhiloxWiener(key, 0, paths, vs, rho=corr)
        okw = {k: v for k, v in kw.items() if k not in ('rho',)}
        preset = {'lognorm': 'lognorm', 'hull_white': 'hull_white', 'ou': 'ou'}[kind]
        ref = ora.integrate(preset, (0., 1.), paths=paths, steps=steps,
                            sources={'dw': src}, **okw)
        xt = ref[-1]
        sc = max(abs(xt.mean()), xt.std(), 1e-3)
        assert abs(mean[i] - xt.mean()) < 2e-5 * sc, (i, kind)
        assert abs(var[i] - xt.var()) < 2e-4 * xt.var() + 1e-12, (i, kind)


def test_batch_launch_shape_does_not_change_results(gpu, tmp_path, monkeypatch):
    """Every (grid, paths per claim, claim order) gives bit-identical moments:
    each 64-path tile writes its own slot and the slots are merged in order;
    the tuner returns a point of the space and caches it by the batch shape."""
    monkeypatch.setenv('SDEPY_AMD_TUNE_CACHE', str(tmp_path / 'tune.json'))
    ref = None
    for tw in B.BatchedTerminalMoments.TILE_WAVES:
        for order in B.BatchedTerminalMoments.ORDERS:
            for blocks in (64, 1024):
                # fresh processes: a batch takes the next epoch of each source
                bt = B.BatchedTerminalMoments(_specs(5)[0], blocks=blocks, tile_waves=tw,
                                              order=order, graph=False)
                m, v = bt.run()
                if ref is None:
                    ref = (m, v)
                assert np.array_equal(m, ref[0]) and np.array_equal(v, ref[1]), \
                    (tw, order, blocks)
    tuned = B.BatchedTerminalMoments(_specs(5)[0], autotune=True, graph=False)
    assert tuned.blocks in tuned.candidates
    assert tuned.tile_waves in tuned.TILE_WAVES and tuned.order in tuned.ORDERS
    assert len(tuned.tuning) >= len(tuned.candidates) + 3
    again = B.BatchedTerminalMoments(_specs(5)[0], autotune=True, graph=False)
    assert not again.tuning     # taken from the cache
    assert (again.blocks, again.tile_waves, again.order) == \
        (tuned.blocks, tuned.tile_waves, tuned.order)
    m, v = tuned.run()
    assert np.array_equal(m, ref[0]) and np.array_equal(v, ref[1])


def test_batch_graph_replays_equal_eager(gpu):
    torch = gpu
    specs, _ = _specs(9)
    bt = B.BatchedTerminalMoments(specs, autotune=True, graph=True)
    assert bt.blocks in bt.candidates
    eager = B.BatchedTerminalMoments(_specs(9)[0], blocks=bt.blocks,
                                     tile_waves=bt.tile_waves, order=bt.order, graph=False)
    m0, v0 = eager.run()
    for rep in range(4):
        m, v = bt.run()
        assert np.array_equal(m, m0) and np.array_equal(v, v0), rep
        # unrelated eager work and a sync between replays
        a = torch.ones(1 << 20, device='cuda')
        float(a.sum())
        torch.cuda.synchronize()


def test_batch_rejects_unsupported(gpu):
    with pytest.raises(NotImplementedError):
        B.BatchedTerminalMoments([(sa.full_heston_process(paths=10), (0., 1.))])
    with pytest.raises(NotImplementedError):
        B.BatchedTerminalMoments([(sa.lognorm_process(paths=10, vshape=2), (0., 1.))])
