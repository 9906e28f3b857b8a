"""The measured workloads at their full sizes (GPU tests).

The other GPU tests hold the kernels to the oracle at sizes the oracle runs
in seconds.  These run the benchmark's own configurations, where the oracle
cannot follow every path, and check what does not depend on the size:

* OU-4 correlated, 10^6 paths x 1000 steps, FP32 full path (16 GB): paths
  picked across the range (first and last path, wave and block edges, the
  middle, random ones) equal the oracle's twin of those very paths (Philox
  counters are per global path index); the row moments reduced as the rows
  are stored equal a separate FP64 pass over the same 16 GB; every row's
  mean and variance sit within 5 SE of the exact moments of the Euler
  recursion (a linear recursion: closed form).
* Heston, 4M paths x 365 steps FP32, price + 7 Greeks: reverse and forward
  mode agree far below the SE; the FP64 adjoint's price equals the FP64
  forward + finite-difference run's (same paths); the FP32 price is within a
  small fraction of the SE of the FP64 one; every FP32 Greek is within 3 of
  its SE of the FP64 central finite differences (the accuracy gate the
  benchmark reports).
* Merton, 2M paths x 252 steps FP32, price + 5 Greeks: FP32 and FP64 runs of
  the same paths agree far below the SE; log-Euler with compound-Poisson
  increments is exact in distribution, so the FP64 price and each Greek are
  within 4 SE of Merton's (1976) Poisson-weighted series and its derivatives
  (central differences of the series, whose truncation error is far below
  the SE).
* The 1024 heterogeneous GBM / Hull-White instances in one batch: chosen
  instances equal their stand-alone runs, the smallest the oracle's twin.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from sdepy_amd import greeks as G
from oracle import closed_forms as cf
from oracle import euler as ora

pytestmark = pytest.mark.gpu


def test_ou4_full_path_1M_x_1000(gpu):
    torch = gpu
    g = np.random.default_rng(1234)
    th, kk, sg = (g.uniform(-1, 1, (4, 1)), g.uniform(.5, 3, (4, 1)),
                  g.uniform(.1, .5, (4, 1)))
    A = g.standard_normal((4, 4))
    S = A @ A.T + 4 * np.eye(4)
    d = 1 / np.sqrt(np.diag(S))
    C = S * d[:, None] * d[None, :]
    P, tt = 10 ** 6, np.linspace(0., 1., 1001)
    p = sa.ornstein_uhlenbeck_process(paths=P, vshape=4, dtype=np.float32, x0=0.,
                                      theta=th, k=kk, sigma=sg, corr=C,
                                      rng=np.random.default_rng(1))
    print('ou4 full path 1M x 1000 fp32, fused row moments', flush=True)
    run = p.prepare(tt, moments=True)
    run.launch()
    torch.cuda.synchronize()
    assert tuple(run.xx_dev.shape) == (1001, 4, P)

    # chosen paths against the oracle's twin of the same global path indices
    idx = np.unique(np.concatenate((
        [0, 1, 63, 64, 255, 256, 257, 65535, 65536, 499_999, 500_000,
         524_287, 524_288, P - 257, P - 256, P - 65, P - 64, P - 1],
        np.random.default_rng(11).integers(0, P, 14))))
    sub = run.xx_dev[:, :, torch.as_tensor(idx, device=run.xx_dev.device)].cpu().numpy()
    prob = run.passes[0].probs[0]
    src = ora.PhiloxWiener(p.sources['dw'].key, int(prob.epoch_w), idx.size, (4,),
                           corr=C, path_index=idx)
    ref = ora.integrate('ou', tt, paths=idx.size, vshape=(4,), sources={'dw': src},
                        x0=0., theta=th, k=kk, sigma=sg)
    assert np.abs(sub - ref).max() < 5e-5

    # fused row moments against a separate FP64 pass over the same buffer
    mean = np.asarray(run.pmean(dtype=np.float64), np.float64).reshape(1001, 4)
    var = np.asarray(run.pvar(dtype=np.float64), np.float64).reshape(1001, 4)
    rows = run.xx_dev.reshape(1001 * 4, P)
    sep_m, sep_v = [], []
    for r0 in range(0, rows.shape[0], 128):
        blk = rows[r0:r0 + 128].double()
        sep_m.append(blk.mean(dim=1).cpu().numpy())
        sep_v.append(blk.var(dim=1, unbiased=False).cpu().numpy())
        del blk
    sep_m = np.concatenate(sep_m).reshape(1001, 4)
    sep_v = np.concatenate(sep_v).reshape(1001, 4)
    assert np.abs(mean - sep_m).max() < 1e-12 * max(1., np.abs(sep_m).max())
    assert np.allclose(var, sep_v, rtol=1e-10, atol=1e-18)

    # every row against the exact moments of the Euler recursion
    means, covs = cf.ou_euler_moments(np.diff(tt), 0., th.ravel(), kk.ravel(),
                                      sg.ravel(), C)
    ev = np.diagonal(covs, axis1=1, axis2=2)                # (1001, 4)
    assert np.all(mean[0] == 0.) and np.all(var[0] == 0.)
    zm = np.abs(mean[1:] - means[1:]) / np.sqrt(ev[1:] / P)
    zv = np.abs(var[1:] - ev[1:]) / (ev[1:] * np.sqrt(2. / P))
    assert zm.max() < 5. and zv.max() < 5., (zm.max(), zv.max())
    del run, rows, p
    torch.cuda.empty_cache()


HPRM = dict(x0=100., r=.02, sigma=1., k=2., theta=.04, xi=.5, rho=-.7, y0=.04)


def test_heston_greeks_4M_x_365(gpu):
    torch = gpu
    kw = dict(paths=4_000_000, steps=365, strike=100., maturity=1.,
              key=(0x5eed, 0xc0ffee), epoch=0, **HPRM)
    print('heston 4M x 365: fp32 adjoint, fp32 tangent, fp64 adjoint, fp64 FD', flush=True)
    a = G.heston_greeks(dtype=np.float32, **kw)
    t = G.heston_greeks(dtype=np.float32, method='tangent', **kw)
    a64 = G.heston_greeks(dtype=np.float64, **kw)
    fd = G.HestonFD(dtype=np.float64, h=1e-4, **kw)
    fd.launch()
    price, price_se, fdg = fd.result()
    torch.cuda.synchronize()
    assert abs(t.price - a.price) < 1e-3 * a.price_se
    assert abs(a64.price - price) < 1e-10 * price
    assert abs(a.price - a64.price) < .05 * a64.price_se
    for name in G.HESTON_GREEKS:
        assert abs(t.greeks[name] - a.greeks[name]) < 1e-3 * a.greeks_se[name], name
        z = abs(a.greeks[name] - fdg[name]) / a.greeks_se[name]
        assert z < 3., (name, a.greeks[name], fdg[name], z)
    del fd
    torch.cuda.empty_cache()


def test_merton_greeks_2M_x_252(gpu):
    prm = dict(x0=1., mu=.03, sigma=.2, lam=1., a=-.1, b=.15)
    kw = dict(r=.03, strike=1., maturity=1., steps=252, paths=2_000_000,
              key_w=(0x3e47, 0x11), key_j=(0x3e47, 0x12), epoch=0, **prm)
    print('merton 2M x 252: fp32 and fp64 Greeks', flush=True)
    r32 = G.merton_greeks(dtype=np.float32, **kw)
    r64 = G.merton_greeks(dtype=np.float64, **kw)
    assert abs(r32.price - r64.price) < .05 * r64.price_se
    for name in G.LOG_GREEKS:
        assert abs(r32.greeks[name] - r64.greeks[name]) < .05 * r64.greeks_se[name], name

    def series(**over):
        q = dict(prm, **over)
        return cf.merton_call(1., 1., discount=.03, **q)

    ref = series()
    z = abs(r64.price - ref) / r64.price_se
    assert z < 4., ('price', r64.price, ref, z)
    for name in G.LOG_GREEKS:
        h = 1e-5 * max(1., abs(prm[name]))
        dref = (series(**{name: prm[name] + h}) - series(**{name: prm[name] - h})) / (2 * h)
        z = abs(r64.greeks[name] - dref) / r64.greeks_se[name]
        assert z < 4., (name, r64.greeks[name], dref, z)


def test_batched_1024_instances(gpu):
    """The benchmark's 1024 heterogeneous GBM / 2-factor Hull-White instances
    (7.5e10 path-steps) in one batch: chosen instances (the most paths and
    the most steps of each kind, the fewest paths, random ones) equal their
    stand-alone preset runs on the same Philox paths, and the smallest one
    the oracle's FP64 twin."""
    from sdepy_amd import batched as B
    print('batched 1024 instances', flush=True)
    specs = B.heterogeneous_specs(1024, seed=1234)
    bt = B.BatchedTerminalMoments(specs, autotune=False, graph=False)
    mean, var = bt.run()
    assert mean.shape == var.shape == (1024,)
    assert np.all(np.isfinite(mean)) and np.all(np.isfinite(var)) and np.all(var > 0)
    fresh = B.heterogeneous_specs(1024, seed=1234)
    paths = np.array([p.paths for p, _ in fresh])
    steps = np.array([p.steps for p, _ in fresh])
    half = len(fresh) // 2
    pick = {int(np.argmax(paths[:half])), half + int(np.argmax(paths[half:])),
            int(np.argmax(steps[:half])), half + int(np.argmax(steps[half:])),
            int(np.argmin(paths))}
    pick |= set(int(i) for i in np.random.default_rng(5).integers(0, len(fresh), 4))
    for i in sorted(pick):
        p, tl = fresh[i]
        print('instance', i, p.paths, p.steps, flush=True)
        x = p(tl)
        m_ref = float(x.pmean(dtype=np.float64)[-1].reshape(-1)[0])
        v_ref = float(x.pvar(dtype=np.float64)[-1].reshape(-1)[0])
        scale = max(abs(m_ref), np.sqrt(v_ref), 1e-3)
        assert abs(mean[i] - m_ref) < 1e-5 * scale, (i, mean[i], m_ref)
        assert abs(var[i] - v_ref) < 1e-4 * v_ref + 1e-12, (i, var[i], v_ref)
    # the smallest instance against the oracle's FP64 twin of its draws
    i = int(np.argmin(paths))
    p = fresh[i][0]
    key, prm = p.sources['dw'].key, p._args
    if i < half:
        src = ora.PhiloxWiener(key, 0, p.paths, ())
        ref = ora.integrate('lognorm', (0., 1.), paths=p.paths, steps=p.steps,
                            sources={'dw': src}, x0=1., mu=prm['mu'], sigma=prm['sigma'])
    else:
        src = ora.PhiloxWiener(key, 0, p.paths, (2,), rho=prm['rho'])
        ref = ora.integrate('hull_white', (0., 1.), paths=p.paths, steps=p.steps,
                            sources={'dw': src}, factors=2, x0=0., theta=0.,
                            k=prm['k'], sigma=prm['sigma'])
    xt = ref[-1].reshape(-1)
    sc = max(abs(xt.mean()), xt.std(), 1e-3)
    assert abs(mean[i] - xt.mean()) < 2e-5 * sc
    assert abs(var[i] - xt.var()) < 2e-4 * xt.var() + 1e-12
