"""Jump sizes of every sdepy jump law drawn inside the kernels (GPU tests).

cpoisson_source's ``y`` (infrastructure.py:2017-2040) may be norm_rv,
uniform_rv, exp_rv or double_exp_rv (:1653-1796), constant or callable in t.
The Euler kernel and the stand-alone fill draw the counts and the sizes of all
four from the jump stream; the oracle's Philox twin (PhiloxCPoisson with the
law) restates the same draws on the CPU.  FP64 runs must equal the oracle to
1e-12 relative, FP32 runs to 5e-5, and the jump info exactly, for
kou_jumpdiff_process, jumpdiff_process's default uniform(0, 1) sizes, an
exponential law on the negative side, a time-dependent double exponential, a
backward pass (i0 != 0), sparse stored rows (the jump_defer form) and a
grouped launch of 9 components; and no increments are recorded on the host.
"""
import numpy as np
import pytest

import sdepy_amd as sa
from sdepy_amd import _lib
from oracle import euler as ora

pytestmark = pytest.mark.gpu

RTOL64 = 1e-12


def _rel(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


def _pa(t):
    return .3 + .4 * t


CASES = {
    # name: (preset, its keywords, the oracle's law keywords)
    'kou': (sa.kou_jumpdiff_process, dict(lam=5., a=.2, b=.3, pa=.6),
            dict(law='double_exp', lam=5., a=.2, b=.3, pa=.6)),
    'uniform_default': (sa.jumpdiff_process, dict(lam=3.),
                        dict(law='uniform', lam=3., a=0., b=1.)),
    'uniform': (sa.jumpdiff_process, dict(lam=4., y=sa.uniform_rv(-.2, .1)),
                dict(law='uniform', lam=4., a=-.2, b=.1)),
    'exp_negative': (sa.jumpdiff_process, dict(lam=4., y=sa.exp_rv(-.25)),
                     dict(law='exp', lam=4., a=-.25)),
    'kou_time_dependent': (sa.kou_jumpdiff_process, dict(lam=5., a=.2, b=.3, pa=_pa),
                           dict(law='double_exp', lam=5., a=.2, b=.3, pa=_pa)),
}


class _PassEpochs:
    """Oracle source mixin: each integration pass draws the epoch that the
    preset's source hands out to it (0 for the first pass, 1 for the next:
    a backward and a forward pass when i0 != 0)."""
    passes = 0

    def begin_pass(self, tag):
        super().begin_pass(tag)
        self.epoch = self.passes
        self.passes += 1


class _Wiener(_PassEpochs, ora.PhiloxWiener):
    pass


class _CPoisson(_PassEpochs, ora.PhiloxCPoisson):
    pass


def _run(name, dtype, tt, steps, i0=0, vshape=(), P=600, seed=41):
    cls, kw, law = CASES[name]
    info = {}
    p = cls(paths=P, vshape=vshape, dtype=dtype, steps=steps, i0=i0, info=info,
            rng=np.random.default_rng(seed), x0=1., mu=.03, sigma=.2, **kw)
    kw_, kj = p.sources['dw'].key, p.sources['dj'].key
    run = p.prepare(tt)
    prob = run.passes[0].probs[0]
    # drawn in the kernel: no recorded increments, the law's code
    assert not prob.replay_dj
    assert prob.jump_law == {'normal': _lib.JUMP_NORMAL, 'uniform': _lib.JUMP_UNIFORM,
                             'exp': _lib.JUMP_DOUBLE_EXP,
                             'double_exp': _lib.JUMP_DOUBLE_EXP}[law['law']]
    x = run.run()
    ref_info = {}
    ref = ora.integrate(
        'jumpdiff', tt, paths=P, vshape=vshape, steps=steps, i0=i0, info=ref_info,
        sources={'dw': _Wiener(kw_, 0, P, vshape),
                 'dj': _CPoisson(kj, 0, P, vshape, **law)},
        x0=1., mu=.03, sigma=.2)
    return x, ref, info, ref_info


@pytest.mark.parametrize('name', sorted(CASES))
@pytest.mark.parametrize('dtype', [np.float64, np.float32])
def test_jump_law_matches_oracle(gpu, name, dtype):
    print('jump law', name, np.dtype(dtype).name, flush=True)
    tt = np.linspace(0., 1., 6)
    x, ref, info, ref_info = _run(name, dtype, tt, 41)
    assert ref_info['jump_count'].sum() > 300
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert np.allclose(info['jump_rate'], ref_info['jump_rate'], rtol=1e-12)
    assert _rel(x, ref) < (RTOL64 if dtype == np.float64 else 5e-5)


@pytest.mark.parametrize('name', ['kou', 'uniform', 'exp_negative'])
def test_jump_law_backward_and_deferred(gpu, name):
    """i0 = 2 (a backward pass with the jump stream of its own tag, the sizes
    negated) and a terminal-only timeline with 200 steps (the host picks the
    jump_defer form: the sizes are drawn at the next stored row)."""
    print('jump law backward/deferred', name, flush=True)
    x, ref, info, ref_info = _run(name, np.float64, np.linspace(0., 1., 5), 33, i0=2)
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert _rel(x, ref) < RTOL64
    cls, kw, _ = CASES[name]
    p = cls(paths=600, x0=1., mu=.03, sigma=.2, steps=201, rng=np.random.default_rng(2), **kw)
    assert p.prepare((0., 1.)).passes[0].probs[0].jump_defer == 1
    x, ref, info, ref_info = _run(name, np.float64, (0., 1.), 201, seed=2)
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert _rel(x, ref) < RTOL64


def test_kou_grouped_components(gpu):
    """vshape (9,): two grouped launches (8 + 1 components), each drawing the
    jump slots s*9 + component of the whole state."""
    x, ref, info, ref_info = _run('kou', np.float64, np.linspace(0., 1., 4), 25,
                                  vshape=(9,), P=300)
    assert np.array_equal(info['jump_count'], ref_info['jump_count'])
    assert _rel(x, ref) < RTOL64


@pytest.mark.parametrize('name', sorted(CASES))
def test_cpoisson_source_fill_matches_oracle(gpu, name):
    """A compound Poisson source called on its own (sdegpu_cpoisson_fill)
    draws what the oracle's source draws slot by slot, and its dn_value."""
    _, kw, law = CASES[name]
    y = kw.get('y')
    if name.startswith('kou'):
        y = sa.double_exp_rv(a=kw['a'], b=kw['b'], pa=kw['pa'])
    P, vshape = 500, (2,)
    src = sa.cpoisson_source(paths=P, vshape=vshape, dtype=np.float64, lam=kw['lam'],
                             y=y, rng=np.random.default_rng(9))
    ref = ora.PhiloxCPoisson(src.key, 0, P, vshape, **law)
    for k, (t, dt) in enumerate(((0., .25), (.25, .5), (.75, .25))):
        if k == 0:
            got = src(np.array([t]), np.array([dt]))[0]
            want = ref(t, dt)
        else:
            # one call per slot: epoch k, slot 0 of the oracle's stream
            got = src(t, dt)
            ref = ora.PhiloxCPoisson(src.key, k, P, vshape, **law)
            want = ref(t, dt)
        assert np.array_equal(np.asarray(src.dn_value).reshape(ref.dn_value.shape),
                              ref.dn_value)
        assert np.abs(src.dn_value).sum() > 50
        np.testing.assert_allclose(got, want, rtol=RTOL64, atol=0)
