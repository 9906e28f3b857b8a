"""Benchmark of the path-integration hot path on one MI355X.

Workload (one "step"): Heston full-truncation Euler, 4,000,000 paths x 365
steps over one year, FP32 state, European call K=100: one launch of the
reverse-mode adjoint kernel, which returns the price and the 7 pathwise
Greeks d/d(x0, r, k, theta, xi, rho, v0) with FP64 accumulation.

Printed as one JSON line:
  value       path-steps/s of the FP32 forward+adjoint (4M*365 per step)
  tangent     the same set by forward-mode derivatives carried along the one
              forward pass (no checkpoints, no recomputation), its time and
              its difference from the adjoint in SE units
  fp64_fd     the same set (price + 7 Greeks) from 15 FP64 forward runs with
              central finite differences on common random numbers, its
              throughput and the speedup of the adjoint over it
  accuracy    |FP32 adjoint - FP64 FD| / SE for the price and every Greek
  epochs      the same over 8 independent epochs: each estimator's mean and
              the SE from the spread across epochs (the pathwise v0/k/theta/
              xi/rho estimators have heavy tails here, so one run's per-path
              SE is unreliable), and the FP32-minus-FP64 adjoint differences
  roofline    the adjoint kernel: algorithmic FP32 flops / kernel time vs the
              MI355X FP32 vector peak (the kernel is VALU-bound, not HBM or
              MFMA); HIP-event timing on the launch stream
  ou_fullpath the OU-4 correlated full-path kernel (1M x 1000, FP32), the
              HBM-bound case: stored bytes / kernel time vs 8 TB/s, and the
              largest deviation from the FP64 run of the same paths;
              fused_moments: the same run with the FP64 moments of every row
              reduced by the kernel as it stores (prepare(moments=True)),
              against the kernel plus a separate moments pass
  merton      Merton jump-diffusion 2M x 252 FP32 adjoint (price + d/d x0,
              mu, sigma, a, b), with the FP64 run of the same paths in SE units
              and its VALU issue rate;
              and the preset forward on the same configuration, terminal row
              only and full path of 253 rows
  summaries   moments (pmean/pvar and montecarlo's 1..4), min/max, 100-bin
              histograms of every row of the OU
              output, chf of its last row, tmax and FP64 tvar along the
              timeline and vmean across the 4 values, on the device (HBM
              read passes)
  gbm_c1      GBM 10k x 252 FP64 full path + pmean/pvar (latency-bound):
              kernel time, the end-to-end preset call with the host process's
              pmean/pvar, and the same through prepare(moments=True), whose
              kernel reduces the rows as it stores them
  batched     1024 heterogeneous GBM / HW2F instances in one persistent
              launch; per-instance FP32 moments vs FP64 runs of the same paths;
              VALU issue rate vs the nominal ceiling (SQ_INSTS_VALU per launch
              from rocprof)
  valu_issue  (tangent, merton, batched) VALU wave-instructions per launch
              (rocprof SQ_INSTS_VALU) per second of kernel time, against the
              nominal one wave64 instruction per CU per clock: the measured
              rate of dependent FMA chains; FP32 add/mul and integer
              instructions issue faster, so a mix can exceed it
  cpu_baseline the oracle's NumPy restatement of sdepy's loop (same
              operations as the reference, NumPy draws) on the host, one core,
              price + 7 Greeks by 15 forward runs on a bounded path sample;
              cpu_baseline_parallel the same path-sharded over up to 16 cores
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ('path-steps/s, Heston 4M×365 FP32 fwd+adjoint vs FP64 fwd+FD, '
          'at ≤3 SE')
HPRM = dict(x0=100., r=.02, sigma=1., k=2., theta=.04, xi=.5, rho=-.7, y0=.04)
PATHS, NSTEPS, STRIKE, MATURITY = 4_000_000, 365, 100., 1.

# algorithmic FP32 flops per path-step of the adjoint kernel (FMA = 2):
#   forward pass 26 (Euler 16 + Box-Muller arithmetic 10), segment recompute
#   of the variance 18 (Box-Muller 10 + variance update 8), reverse sweep 30
ADJ_FLOPS_PER_PATH_STEP = 74
FP32_VECTOR_PEAK_TFLOPS = 157.3
HBM_PEAK_GBS = 8000.
# nominal VALU issue rate: one wave64 vector instruction per clock per CU,
# 256 CUs at 2.4 GHz (FMA chains issue at 0.91-0.96 of it; simpler integer
# and FP32 add/mul/mov instructions faster)
VALU_ISSUE_PEAK_PER_S = 256 * 2.4e9

# HBM bytes per launch from rocprofv3 FETCH_SIZE / WRITE_SIZE (KiB per
# dispatch), scaled by what the same counters read for a 1 GiB device copy
# (FETCH reads half the copied bytes: x2.0; WRITE is exact: x1.0), and the
# kernel's VALU wave-instructions per launch (SQ_INSTS_VALU); the summaries
# they come from are in profiles/
ADJ_TRAFFIC_BYTES = 1383912141        # k_heston_adjoint<float, float, 8>, 4M x 365
ADJ_VALU_WAVE_INSTS = 2003390625      # of which transcendental:
ADJ_VALU_TRANS_WAVE_INSTS = 251187500  # (issued beside the other VALU work)
OU_TRAFFIC_BYTES = 16026172416         # k_euler<OU, float, 4>, 1M x 1000
BATCH_VALU_WAVE_INSTS = 31445184494    # k_batch_persistent, the 1024-instance set, 8 tiles per claim
MERTON_VALU_WAVE_INSTS = 249545552     # k_loglin_greeks<float>, 2M x 252
TANGENT_VALU_WAVE_INSTS = 1364890625   # k_heston_tangent<float, float>, 4M x 365


def _events(torch):
    return (torch.cuda.Event(enable_timing=True),
            torch.cuda.Event(enable_timing=True))


def timed(torch, fn, steps, warmup, stream=None):
    """Mean time of fn() in ms from HIP events recorded on the stream the
    kernels are launched on (torch's current stream unless given)."""
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = _events(torch)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return e0.elapsed_time(e1) / steps, wall


def cpu_baseline(paths, steps, seed=1234):
    """The oracle (NumPy restatement of sdepy's Euler loop with NumPy
    multivariate-normal draws, as sdepy makes them) timed on one host core:
    15 forward runs give the price and 7 central-difference Greeks."""
    from oracle import euler as ora
    names = ('x0', 'r', 'k', 'theta', 'xi', 'rho', 'y0')
    base = {k: HPRM[k] for k in names}
    runs = [dict(base)]
    for n in names:
        for sg in (1, -1):
            d = dict(base)
            d[n] = base[n] * (1 + sg * 1e-4)
            runs.append(d)
    t0 = time.perf_counter()
    for d in runs:
        rng = np.random.RandomState(seed)   # common random numbers
        src = ora.NumpyWiener(rng, paths, (2,), rho=d['rho'])
        x, _ = ora.integrate('full_heston', (0., MATURITY), paths=paths,
                             steps=steps + 1, sources={'dw': src},
                             x0=d['x0'], mu=d['r'], sigma=1., y0=d['y0'],
                             theta=d['theta'], k=d['k'], xi=d['xi'])
        np.maximum(x[-1] - STRIKE, 0.).mean()
    dt = time.perf_counter() - t0
    return {'value': paths * steps / dt, 'unit': 'path-steps/s',
            'cores': 1, 'kind': 'port',
            'sample': ('{} paths x {} steps, FP64, price + 7 Greeks by 15 CRN '
                       'forward runs of the oracle, {:.1f} s').format(
                           paths, steps, dt),
            'host_cpu': _cpu_info(),
            'forward_path_steps_per_s': 15 * paths * steps / dt}


def cpu_baseline_subprocess(paths, steps):
    """cpu_baseline in a child process pinned to one thread."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS='1', OPENBLAS_NUM_THREADS='1',
               MKL_NUM_THREADS='1')
    code = ('import json, bench; print(json.dumps(bench.cpu_baseline({}, {})))'
            .format(int(paths), int(steps)))
    res = subprocess.run([sys.executable, '-c', code], cwd=ROOT, env=env,
                         capture_output=True, text=True, check=True)
    return json.loads(res.stdout.strip().splitlines()[-1])


def cpu_baseline_parallel(paths, steps, workers):
    """cpu_baseline in `workers` concurrent single-threaded child processes,
    each on its own sample of `paths` paths (path-sharded, independent
    seeds): the aggregate path-steps/s of the host's cores."""
    import subprocess
    env = dict(os.environ, OMP_NUM_THREADS='1', OPENBLAS_NUM_THREADS='1',
               MKL_NUM_THREADS='1')
    procs = []
    t0 = time.perf_counter()
    for i in range(workers):
        code = ('import json, bench; print(json.dumps(bench.cpu_baseline({}, {}, {})))'
                .format(int(paths), int(steps), 1234 + i))
        procs.append(subprocess.Popen([sys.executable, '-c', code], cwd=ROOT, env=env,
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                      text=True))
    res = []
    for pr in procs:
        o, e = pr.communicate()
        if pr.returncode != 0:
            raise RuntimeError('CPU baseline worker failed: ' + e[-500:])
        res.append(json.loads(o.strip().splitlines()[-1]))
    wall = time.perf_counter() - t0
    return {'value': workers * paths * steps / wall, 'unit': 'path-steps/s',
            'cores': workers, 'kind': 'port',
            'sample': ('{} processes x {} paths x {} steps, FP64, price + 7 Greeks '
                       'by 15 CRN forward runs each, {:.1f} s wall including '
                       'process start').format(workers, paths, steps, wall),
            'per_process_value': [r['value'] for r in res],
            'host_cpu': _cpu_info()}


def _device_summaries(torch, run):
    """Time the cross-path summaries on a prepared run's device output (the
    C2 full path, 4004 rows x 1M paths FP32 = 16 GB): FP64 moments of every
    row (pmean/pvar), a 100-bin histogram of every row (montecarlo counts),
    min/max of every row, and chf at 8 arguments of the last time point.
    Each is a read pass over its input; bytes read / time vs 8 TB/s."""
    import ctypes
    from sdepy_amd import _lib
    lib = _lib.lib()
    rows = run.xx_dev.reshape(-1, run.xx_dev.shape[-1])
    R, P = int(rows.shape[0]), int(rows.shape[1])
    nbytes = R * P * 4
    res = {}
    mean = _lib.empty((R,), np.float64)
    m2 = _lib.empty((R,), np.float64)
    ws_n = lib.sdegpu_moments_workspace(R, P)
    ws = _lib.workspace(ws_n)
    st = _lib.stream()

    def moments():
        _lib.check(lib.sdegpu_path_moments(_lib.ptr(rows), _lib.F32, R, P,
                                           _lib.ptr(mean), _lib.ptr(m2),
                                           _lib.ptr(ws), ws_n, st))
    ms, _ = timed(torch, moments, 5, 2)
    res['moments'] = {'ms': ms, 'gb_per_s': nbytes / (ms * 1e-3) / 1e9}
    mc_n = lib.sdegpu_mc_workspace(R, P)
    mws = _lib.workspace(mc_n)
    mm = _lib.empty((R, 3), np.float64)

    def minmax():
        _lib.check(lib.sdegpu_mc_minmax(_lib.ptr(rows), _lib.F32, R, P, 0,
                                        _lib.ptr(mm), _lib.ptr(mws), mc_n, st))
    ms, _ = timed(torch, minmax, 5, 2)
    res['minmax'] = {'ms': ms, 'gb_per_s': nbytes / (ms * 1e-3) / 1e9}
    # montecarlo moments 1..4 about each row's mean (FP64 sums)
    mom = _lib.empty((R, 5), np.float64)

    def mc_moments():
        _lib.check(lib.sdegpu_mc_moments(_lib.ptr(rows), _lib.F32, R, P, 0,
                                         _lib.ptr(mean), _lib.ptr(mom), _lib.ptr(mws),
                                         mc_n, st))
    ms, _ = timed(torch, mc_moments, 5, 2)
    res['mc_moments_1_4'] = {'ms': ms, 'gb_per_s': nbytes / (ms * 1e-3) / 1e9}
    h = _lib.to_host(mm)
    # np.histogram's edges for 100 bins (a row of equal values widens to +-0.5)
    edges = np.stack([np.histogram_bin_edges(np.array([lo, hi]), bins=100)
                      for lo, hi in h[:, :2]])
    de = _lib.to_device(edges)
    counts = _lib.empty((R, 101), np.int64)

    def hist():
        _lib.check(lib.sdegpu_mc_histogram(_lib.ptr(rows), _lib.F32, R, P, 0,
                                           _lib.ptr(de), 100, 1, _lib.ptr(counts),
                                           st))
    ms, _ = timed(torch, hist, 5, 2)
    c = _lib.to_host(counts)
    res['histogram_100_bins'] = {'ms': ms, 'gb_per_s': nbytes / (ms * 1e-3) / 1e9,
                                 'all_values_counted': bool((c.sum(axis=1) == P).all())}
    last = rows[-4:]
    u = _lib.to_device(np.linspace(-4., 4., 8))
    cn = lib.sdegpu_chf_workspace(4, P, 8)
    cws = _lib.workspace(cn)
    cout = _lib.empty((4, 8, 2), np.float64)

    def chf():
        _lib.check(lib.sdegpu_path_chf(_lib.ptr(last), _lib.F32, 4, P, _lib.ptr(u),
                                       8, _lib.ptr(cout), _lib.ptr(cws), cn, st))
    ms, _ = timed(torch, chf, 5, 2)
    res['chf_8_args_last_row'] = {'ms': ms, 'values_per_s': 4 * P * 8 / (ms * 1e-3)}
    # along the timeline (tmax, tvar in FP64: two read passes) and across the
    # 4 values (vmean), NumPy's reduction order; bytes = reads + writes
    n_t, nout, _ = run.xx_dev.shape
    cols = nout * P
    tout = _lib.empty((1, cols), np.float64)

    def tred(op, acc):
        return lambda: _lib.check(lib.sdegpu_axis_reduce(
            _lib.ptr(run.xx_dev), _lib.F32, 1, n_t, cols, 0, cols, None, op, acc, 0,
            _lib.ptr(tout), st))
    for name, op, acc, passes, osz in (('tmax', _lib.REDUCE_MAX, _lib.F32, 1, 4),
                                       ('tvar_f64', _lib.REDUCE_VAR, _lib.F64, 2, 8)):
        ms, _ = timed(torch, tred(op, acc), 5, 2)
        res[name] = {'ms': ms, 'gb_per_s': (passes * nbytes + cols * osz) / (ms * 1e-3) / 1e9}
    vout = _lib.empty((n_t, P), np.float32)
    comps = (ctypes.c_int32 * nout)(*range(nout))

    def vmean():
        _lib.check(lib.sdegpu_axis_reduce(_lib.ptr(run.xx_dev), _lib.F32, n_t, nout, P,
                                          nout * P, P, comps, _lib.REDUCE_MEAN, _lib.F32,
                                          0, _lib.ptr(vout), st))
    ms, _ = timed(torch, vmean, 5, 2)
    res['vmean'] = {'ms': ms, 'gb_per_s': (nbytes + n_t * P * 4) / (ms * 1e-3) / 1e9}
    return res


def _epoch_spread(G, fd_ms, n_epochs=8):
    """The C3 Greeks over n_epochs independent Philox epochs: per estimator
    (FP32 adjoint, FP64 adjoint, FP64 FD) the mean over epochs and its
    standard error from their spread; the FP32 and FP64 FD means apart in
    units of that SE; and the FP32-minus-FP64 adjoint difference on the same
    paths, in units of the FP64 run's own per-path SE and of the epoch
    spread."""
    names = ('price',) + tuple(G.HESTON_GREEKS)
    kw = dict(paths=PATHS, steps=NSTEPS, strike=STRIKE, maturity=MATURITY,
              key=(0x5eed, 0xc0ffee), **HPRM)
    rows = {'fp32_adjoint': [], 'fp64_adjoint': [], 'fp64_fd': [],
            'fp32_adjoint_fp64_variance': []}
    diff_se64, diff_v64 = [], []
    for e in range(n_epochs):
        r32 = G.heston_greeks(dtype=np.float32, epoch=e, **kw)
        rv = G.heston_greeks(dtype=np.float32, variance_dtype=np.float64, epoch=e, **kw)
        r64 = G.heston_greeks(dtype=np.float64, epoch=e, **kw)
        fd = G.HestonFD(dtype=np.float64, epoch=e, h=1e-4, **kw)
        fd.launch()
        p64, _, fdg = fd.result()
        del fd
        rows['fp32_adjoint'].append([r32.price] + [r32.greeks[k] for k in G.HESTON_GREEKS])
        rows['fp64_adjoint'].append([r64.price] + [r64.greeks[k] for k in G.HESTON_GREEKS])
        rows['fp64_fd'].append([p64] + [fdg[k] for k in G.HESTON_GREEKS])
        rows['fp32_adjoint_fp64_variance'].append(
            [rv.price] + [rv.greeks[k] for k in G.HESTON_GREEKS])
        diff_se64.append([(r32.greeks[k] - r64.greeks[k]) / r64.greeks_se[k]
                          for k in G.HESTON_GREEKS])
        diff_v64.append([(rv.greeks[k] - r64.greeks[k]) / r64.greeks_se[k]
                         for k in G.HESTON_GREEKS])
    res = {'epochs': n_epochs, 'paths_per_epoch': PATHS}
    stats = {}
    for est, v in rows.items():
        v = np.asarray(v)
        m, sd = v.mean(axis=0), v.std(axis=0, ddof=1) / math.sqrt(n_epochs)
        stats[est] = (m, sd)
        res[est] = {'mean': dict(zip(names, m)), 'se_from_spread': dict(zip(names, sd))}
    m32, s32 = stats['fp32_adjoint']
    mfd, sfd = stats['fp64_fd']
    res['fp32_adjoint_minus_fd_over_spread_se'] = dict(zip(
        names, (m32 - mfd) / np.sqrt(s32 ** 2 + sfd ** 2)))
    d = np.asarray(diff_se64)
    v32 = np.asarray(rows['fp32_adjoint'])[:, 1:]
    v64 = np.asarray(rows['fp64_adjoint'])[:, 1:]
    res['fp32_minus_fp64_adjoint'] = {
        'max_abs_over_se64_per_epoch': dict(zip(G.HESTON_GREEKS, np.abs(d).max(axis=0))),
        'mean_over_spread_se': dict(zip(G.HESTON_GREEKS, (v32 - v64).mean(axis=0) /
                                        stats['fp64_adjoint'][1][1:]))}
    # the FP32 kernel with an FP64 variance state (FP32 normals), on the same
    # paths: how much of the FP32-FP64 difference the variance's rounding makes
    dv = np.asarray(diff_v64)
    vv = np.asarray(rows['fp32_adjoint_fp64_variance'])[:, 1:]
    res['fp32_fp64_variance_minus_fp64_adjoint'] = {
        'max_abs_over_se64_per_epoch': dict(zip(G.HESTON_GREEKS, np.abs(dv).max(axis=0))),
        'mean_over_spread_se': dict(zip(G.HESTON_GREEKS, (vv - v64).mean(axis=0) /
                                        stats['fp64_adjoint'][1][1:]))}
    return res


def _cpu_name():
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    return line.split(':', 1)[1].strip()
    except OSError:
        pass
    return 'unknown'


def _cpu_info():
    """The host CPU as lscpu describes it, and the CPU counts this process
    sees (os.cpu_count and its affinity mask)."""
    import subprocess
    info = {'model': _cpu_name(), 'os_cpu_count': os.cpu_count()}
    try:
        info['affinity_cpus'] = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        pass
    try:
        out = subprocess.run(['lscpu'], capture_output=True, text=True,
                             timeout=10).stdout
    except (OSError, subprocess.SubprocessError):
        return info
    keys = {'Model name': 'lscpu_model', 'Vendor ID': 'vendor',
            'CPU(s)': 'lscpu_cpus', 'Socket(s)': 'sockets',
            'Core(s) per socket': 'cores_per_socket',
            'Thread(s) per core': 'threads_per_core',
            'CPU max MHz': 'max_mhz', 'Hypervisor vendor': 'hypervisor'}
    for line in out.splitlines():
        k, _, v = line.partition(':')
        if k.strip() in keys and v.strip():
            info[keys[k.strip()]] = v.strip()
    return info


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--legs', default='adjoint,tangent,fd,ou,merton,gbm,batched,cpu',
                    help='comma list of: adjoint, tangent (forward-mode Greeks '
                         'on the same workload), fd, ou, merton, gbm, batched, '
                         'cpu, copyref (a 1 GiB device copy used to calibrate '
                         'rocprof byte counters), ou64 (FP64 twin of the OU '
                         'leg for its error figure; on by default with ou), '
                         'summaries (device summaries of the OU output; on by '
                         'default with ou), batched64 (FP64 twins of the '
                         'batched instances; on by default with batched)')
    ap.add_argument('--skip', default='',
                    help='comma list of legs to leave out (e.g. ou64,batched64 '
                         'under a profiler)')
    ap.add_argument('--cpu-paths', type=int, default=64000)
    args = ap.parse_args()
    legs = set(args.legs.split(','))
    if 'ou' in legs:
        legs.add('ou64')
        legs.add('summaries')
    if 'batched' in legs:
        legs.add('batched64')
    legs -= set(args.skip.split(','))
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: the path kernels run on a single '
              'MI355X (a multi-GPU run would shard the paths by range); '
              '--gpus {} is not supported'.format(args.gpus),
              file=sys.stderr)
        sys.exit(2)

    from sdepy_amd import _lib, greeks as G
    torch = _lib.require_gpu()
    torch.cuda.set_device(0)

    if 'copyref' in legs:
        # a plain copy of known size: rocprofv3 FETCH_SIZE / WRITE_SIZE of this
        # kernel read 1 GiB each when the counters are exact
        src = torch.empty(1 << 28, dtype=torch.float32, device='cuda').fill_(1.)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # ---- headline: FP32 forward + adjoint, price + 7 Greeks ----
    if 'adjoint' not in legs:
        args.steps, args.warmup = 1, 0
    run = G.HestonGreeks(paths=PATHS, steps=NSTEPS, strike=STRIKE,
                         maturity=MATURITY, dtype=np.float32,
                         key=(0x5eed, 0xc0ffee), epoch=0, **HPRM)
    ms, wall = timed(torch, run.launch, args.steps, args.warmup)
    res = run.result()
    value = run.path_steps / (ms * 1e-3)
    out = {
        'metric': METRIC, 'value': value, 'unit': 'path-steps/s',
        'n_gpus': 1, 'steps': args.steps, 'warmup': args.warmup,
        'ms_per_step': ms, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f32', 'data': 'synthetic',
        'config': {'workload': 'heston_ft_adjoint_4Mx365_fp32',
                   'paths': PATHS, 'steps_per_path': NSTEPS,
                   'maturity': MATURITY, 'strike': STRIKE, **HPRM,
                   'greeks': list(G.HESTON_GREEKS)},
        'wall_s_timed': wall,
        'price': res.price, 'price_se': res.price_se,
        'greeks': res.greeks, 'greeks_se': res.greeks_se,
        'greeks_per_s': 7 / (ms * 1e-3),
    }
    achieved = ADJ_FLOPS_PER_PATH_STEP * run.path_steps / (ms * 1e-3) / 1e12
    out['roofline'] = {
        'kernel': 'k_heston_adjoint<float>', 'bound': 'valu',
        'achieved': achieved, 'peak': FP32_VECTOR_PEAK_TFLOPS,
        'unit': 'TFLOP/s', 'frac': achieved / FP32_VECTOR_PEAK_TFLOPS,
        'traffic': ADJ_TRAFFIC_BYTES,
        'flops_per_path_step': ADJ_FLOPS_PER_PATH_STEP,
        'frac_of_unpacked_fp32_peak': achieved / (FP32_VECTOR_PEAK_TFLOPS / 2)}
    if ADJ_VALU_WAVE_INSTS:
        # the kernel's true ceiling: non-transcendental VALU wave-instructions
        # issued per CU per clock (rocprof SQ_INSTS_VALU minus
        # SQ_INSTS_VALU_TRANS_F32 per launch, over the measured launch time)
        rate = (ADJ_VALU_WAVE_INSTS - ADJ_VALU_TRANS_WAVE_INSTS) / (ms * 1e-3)
        out['roofline']['valu_issue'] = {
            'wave_insts_per_launch': ADJ_VALU_WAVE_INSTS,
            'transcendental_wave_insts_per_launch': ADJ_VALU_TRANS_WAVE_INSTS,
            'achieved_per_s': rate, 'peak_per_s': VALU_ISSUE_PEAK_PER_S,
            'frac': rate / VALU_ISSUE_PEAK_PER_S}

    # ---- the same set by forward-mode derivatives (method='tangent') ----
    if 'tangent' in legs:
        trun = G.HestonGreeks(paths=PATHS, steps=NSTEPS, strike=STRIKE,
                              maturity=MATURITY, dtype=np.float32,
                              key=(0x5eed, 0xc0ffee), epoch=0, method='tangent',
                              **HPRM)
        t_ms, _ = timed(torch, trun.launch, args.steps, args.warmup)
        tres = trun.result()
        out['tangent'] = {
            'kernel': 'k_heston_tangent<float>', 'ms': t_ms,
            'path_steps_per_s': trun.path_steps / (t_ms * 1e-3),
            'greeks_per_s': 7 / (t_ms * 1e-3),
            'speedup_over_adjoint': ms / t_ms,
            'price': tres.price, 'greeks': tres.greeks,
            'minus_adjoint_over_se': {
                k: (tres.greeks[k] - res.greeks[k]) / res.greeks_se[k]
                for k in G.HESTON_GREEKS},
            'valu_issue': {
                'wave_insts_per_launch': TANGENT_VALU_WAVE_INSTS,
                'achieved_per_s': TANGENT_VALU_WAVE_INSTS / (t_ms * 1e-3),
                'nominal_per_s': VALU_ISSUE_PEAK_PER_S,
                'frac_of_nominal': TANGENT_VALU_WAVE_INSTS / (t_ms * 1e-3)
                / VALU_ISSUE_PEAK_PER_S}}
        del trun
        # the FP64-variance option of both kernels: its cost
        fv = {}
        for m in ('adjoint', 'tangent'):
            vrun = G.HestonGreeks(paths=PATHS, steps=NSTEPS, strike=STRIKE,
                                  maturity=MATURITY, dtype=np.float32,
                                  variance_dtype=np.float64, key=(0x5eed, 0xc0ffee),
                                  epoch=0, method=m, **HPRM)
            v_ms, _ = timed(torch, vrun.launch, args.steps, args.warmup)
            fv[m + '_ms'] = v_ms
            del vrun
        fv['adjoint_cost_over_fp32'] = fv['adjoint_ms'] / ms
        fv['tangent_cost_over_fp32'] = fv['tangent_ms'] / t_ms
        out['fp64_variance'] = fv

    # ---- FP64 forward + CRN finite differences, same paths ----
    if 'fd' in legs:
        fd = G.HestonFD(paths=PATHS, steps=NSTEPS, strike=STRIKE,
                        maturity=MATURITY, dtype=np.float64,
                        key=(0x5eed, 0xc0ffee), epoch=0, h=1e-4, **HPRM)
        fd_ms, _ = timed(torch, fd.launch, 2, 1)
        price64, se64, fdg = fd.result()
        fd_value = PATHS * NSTEPS / (fd_ms * 1e-3)
        acc = {'price': abs(res.price - price64) / res.price_se}
        for k in G.HESTON_GREEKS:
            acc[k] = abs(res.greeks[k] - fdg[k]) / res.greeks_se[k]
        if 'tangent' in out:
            out['tangent']['accuracy_max_se'] = max(
                [abs(out['tangent']['price'] - price64) / res.price_se] +
                [abs(out['tangent']['greeks'][k] - fdg[k]) / res.greeks_se[k]
                 for k in G.HESTON_GREEKS])
            out['tangent']['speedup_over_fp64_fd'] = \
                out['tangent']['path_steps_per_s'] / fd_value
        out['fp64_fd'] = {'value': fd_value, 'unit': 'path-steps/s',
                          'ms_per_set': fd_ms, 'price': price64,
                          'greeks': fdg, 'speedup_of_adjoint': value / fd_value}
        out['accuracy_se'] = acc
        out['accuracy_max_se'] = max(acc.values())
        # the FP64 adjoint on the same paths: the k/theta/xi/rho/v0 pathwise
        # estimators have heavy tails here (the variance reaches 0: Feller
        # fails), so single paths with tiny variance dominate their sums
        r64 = G.heston_greeks(paths=PATHS, steps=NSTEPS, strike=STRIKE,
                              maturity=MATURITY, dtype=np.float64,
                              key=(0x5eed, 0xc0ffee), epoch=0, **HPRM)
        out['fp64_adjoint'] = {
            'price': r64.price, 'greeks': r64.greeks,
            'fp32_minus_fp64_over_se64': {
                k: (res.greeks[k] - r64.greeks[k]) / r64.greeks_se[k]
                for k in G.HESTON_GREEKS},
            'fp64_adjoint_minus_fd_over_se64': {
                k: (r64.greeks[k] - fdg[k]) / r64.greeks_se[k]
                for k in G.HESTON_GREEKS}}
        # 8 independent epochs (fresh paths) of the FP32 adjoint, the FP64
        # adjoint and the FP64 FD set: the spread of each Greek across them
        # estimates the standard error of one run where the pathwise
        # estimator's variance is dominated by a few paths
        out['epochs'] = _epoch_spread(G, fd_ms)
        del fd
        torch.cuda.empty_cache()

    # ---- OU-4 correlated full path (HBM-bound) ----
    if 'ou' in legs:
        import sdepy_amd as sa
        g = np.random.default_rng(1234)
        th, kk, sg = (g.uniform(-1, 1, (4, 1)), g.uniform(.5, 3, (4, 1)),
                      g.uniform(.1, .5, (4, 1)))
        A = g.standard_normal((4, 4))
        S = A @ A.T + 4 * np.eye(4)
        dd = 1 / np.sqrt(np.diag(S))
        C = S * dd[:, None] * dd[None, :]
        p = sa.ornstein_uhlenbeck_process(
            paths=10 ** 6, vshape=4, dtype=np.float32, x0=0., theta=th, k=kk,
            sigma=sg, corr=C, rng=np.random.default_rng(1))
        ou = p.prepare(np.linspace(0., 1., 1001))
        ou_ms, _ = timed(torch, lambda: ou.launch(fresh=True), 10, 5)
        nbytes = 1001 * 4 * 10 ** 6 * 4
        gbs = nbytes / (ou_ms * 1e-3) / 1e9
        out['ou_fullpath'] = {
            'workload': 'ou4_corr_fullpath_1Mx1000_fp32',
            'ms': ou_ms, 'path_steps_per_s': ou.path_steps / (ou_ms * 1e-3),
            'roofline': {'kernel': 'k_euler<OU,float,4>', 'bound': 'hbm',
                         'achieved': gbs, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                         'frac': gbs / HBM_PEAK_GBS,
                         'traffic': OU_TRAFFIC_BYTES,
                         'bytes_per_path_step': 16}}
        # the store ceiling of this output pattern measured in the same run:
        # the same rows written with nothing computed (one thread per path,
        # rows in order, non-temporal), and torch's fill of the buffer
        lib = _lib.lib()
        rows = int(ou.xx_dev.shape[0] * ou.xx_dev.shape[1])

        def fill_rows():
            _lib.check(lib.sdegpu_fill_rows(_lib.F32, float('nan'), _lib.ptr(ou.xx_dev),
                                            rows, 10 ** 6, _lib.stream()))
        fr_ms, _ = timed(torch, fill_rows, 10, 3)
        tf_ms, _ = timed(torch, lambda: ou.xx_dev.fill_(float('nan')), 10, 3)
        store_peak = nbytes / (min(fr_ms, tf_ms) * 1e-3) / 1e9
        out['ou_fullpath']['roofline']['store_ceiling'] = {
            'fill_rows_gb_per_s': nbytes / (fr_ms * 1e-3) / 1e9,
            'torch_fill_gb_per_s': nbytes / (tf_ms * 1e-3) / 1e9,
            'frac_of_measured_store_peak': gbs / store_peak}
        ou.launch(fresh=True)
        # the same run with the row moments reduced by the kernel as it
        # stores (block partials) and merged after it: pmean/pvar of every
        # row without another pass over the 16 GB
        ouf = p.prepare(np.linspace(0., 1., 1001), moments=True)
        f_ms, _ = timed(torch, lambda: ouf.launch(fresh=True), 10, 5)
        out['ou_fullpath']['fused_moments'] = {
            'ms': f_ms, 'over_plain_kernel': f_ms / ou_ms,
            'partials_bytes': int(ouf._row_mom.numel() * 8),
            'separate_pass_ms': None}
        del ouf
        torch.cuda.empty_cache()
        if 'summaries' in legs:
            out['summaries'] = _device_summaries(torch, ou)
            out['ou_fullpath']['fused_moments']['separate_pass_ms'] = \
                ou_ms + out['summaries']['moments']['ms']
        if 'ou64' in legs:
            # the same paths (key and last epoch) stepped in FP64
            p64 = sa.ornstein_uhlenbeck_process(
                paths=10 ** 6, vshape=4, dtype=np.float64, x0=0., theta=th,
                k=kk, sigma=sg, corr=C, rng=np.random.default_rng(1))
            ou64 = p64.prepare(np.linspace(0., 1., 1001))
            ou64.passes[0].prob.epoch_w = ou.passes[0].prob.epoch_w
            ou64.launch()
            torch.cuda.synchronize()
            err = (ou.xx_dev.double() - ou64.xx_dev).abs().max().item()
            scale = ou64.xx_dev.abs().max().item()
            out['ou_fullpath']['max_abs_err_vs_fp64'] = err
            out['ou_fullpath']['max_abs_fp64'] = scale
            del ou64, p64
        del ou, p
        torch.cuda.empty_cache()

    # ---- Merton jump-diffusion adjoint (C4) ----
    if 'merton' in legs:
        mprm = dict(x0=1., mu=.03, sigma=.2, lam=1., a=-.1, b=.15, r=.03,
                    strike=1., maturity=1., steps=252, paths=2_000_000,
                    key_w=(0x3e47, 0x11), key_j=(0x3e47, 0x12), epoch=0)
        mrun = G.LogGreeks(dtype=np.float32, **mprm)
        m_ms, _ = timed(torch, mrun.launch, 20, 10)
        mres = mrun.result()
        m64 = G.LogGreeks(dtype=np.float64, **mprm)
        m64.launch()
        r64 = m64.result()
        err = {'price': abs(mres.price - r64.price) / r64.price_se}
        for k in G.LOG_GREEKS:
            err[k] = abs(mres.greeks[k] - r64.greeks[k]) / r64.greeks_se[k]
        out['merton'] = {
            'workload': 'merton_adjoint_2Mx252_fp32',
            'ms': m_ms, 'path_steps_per_s': mrun.path_steps / (m_ms * 1e-3),
            'greeks_per_s': len(G.LOG_GREEKS) / (m_ms * 1e-3),
            'price': mres.price, 'price_se': mres.price_se,
            'greeks': mres.greeks, 'greeks_se': mres.greeks_se,
            'fp32_vs_fp64_same_paths_se': err,
            'valu_issue': {
                'kernel': 'k_loglin_greeks<float>', 'bound': 'valu',
                'wave_insts_per_launch': MERTON_VALU_WAVE_INSTS,
                'achieved_per_s': MERTON_VALU_WAVE_INSTS / (m_ms * 1e-3),
                'nominal_per_s': VALU_ISSUE_PEAK_PER_S,
                'frac_of_nominal': MERTON_VALU_WAVE_INSTS / (m_ms * 1e-3)
                / VALU_ISSUE_PEAK_PER_S}}
        del mrun, m64
        # the preset forward (k_euler, jumpdiff) on the same configuration:
        # the terminal row only (jump sizes drawn at the stored row,
        # jump_defer) and the full path of 253 rows (jumps summed in place)
        import sdepy_amd as sa
        fwd = {}
        for name, tl in (('terminal', (0., 1.)), ('full_path', np.linspace(0., 1., 253))):
            mp = sa.merton_jumpdiff_process(paths=2_000_000, steps=253, dtype=np.float32,
                                            x0=1., mu=.03, sigma=.2, lam=1., a=-.1, b=.15,
                                            rng=np.random.default_rng(7))
            frun = mp.prepare(tl)
            f_ms, _ = timed(torch, lambda: frun.launch(fresh=True), 10, 3)
            fwd[name] = {'ms': f_ms, 'path_steps_per_s': frun.path_steps / (f_ms * 1e-3),
                         'rows': len(tl), 'jump_defer': int(frun.passes[0].prob.jump_defer)}
            del frun, mp
        out['merton']['forward'] = fwd

    # ---- GBM 10k x 252 FP64 full path + pmean/pvar (C1, latency-bound) ----
    if 'gbm' in legs:
        import sdepy_amd as sa
        tt1 = np.linspace(0., 1., 253)

        def c1_proc():
            return sa.lognorm_process(paths=10_000, x0=1., mu=.05, sigma=.2,
                                      dtype=np.float64,
                                      rng=np.random.default_rng(1234))
        prep = c1_proc().prepare(tt1)
        k_ms, _ = timed(torch, lambda: prep.launch(fresh=True), 20, 10)
        # each variant once untimed first (its kernels' first use, the first
        # pinned result buffer), then 5 timed calls
        x = c1_proc()(tt1)
        x.pmean(), x.pvar()
        t0 = time.perf_counter()
        reps = 5
        for _ in range(reps):
            x = c1_proc()(tt1)
            m, v = x.pmean()[-1], x.pvar()[-1]
        e2e = (time.perf_counter() - t0) / reps
        # the same through a prepared run whose kernel reduces the rows as
        # it stores them: the path array still comes back to the host, the
        # moments without another upload or pass
        run = c1_proc().prepare(tt1, moments=True)
        run.run()
        t0 = time.perf_counter()
        for _ in range(reps):
            run = c1_proc().prepare(tt1, moments=True)
            x2 = run.run()
            m2, v2 = run.pmean()[-1], run.pvar()[-1]
        e2e_fused = (time.perf_counter() - t0) / reps
        out['gbm_c1'] = {
            'workload': 'gbm_fullpath_10kx252_fp64',
            'kernel_ms': k_ms,
            'kernel_path_steps_per_s': 10_000 * 252 / (k_ms * 1e-3),
            'end_to_end_ms': e2e * 1e3,
            'end_to_end_fused_moments_ms': e2e_fused * 1e3,
            'terminal_mean': float(np.asarray(m).ravel()[0]),
            'terminal_var': float(np.asarray(v).ravel()[0]),
            'fused_minus_host_moments': [
                float(np.asarray(m2).ravel()[0] - np.asarray(m).ravel()[0]),
                float(np.asarray(v2).ravel()[0] - np.asarray(v).ravel()[0])]}
        del prep, x

    # ---- 1024 heterogeneous GBM / 2-factor Hull-White instances ----
    if 'batched' in legs:
        from sdepy_amd import batched as B
        specs = B.heterogeneous_specs(1024, seed=1234)
        t0 = time.perf_counter()
        bt = B.BatchedTerminalMoments(specs, autotune=True, graph=True)
        prep = time.perf_counter() - t0
        bms, _ = timed(torch, bt.launch, 3, 1, stream=bt.stream)
        mean, var = bt.result()
        inst_err = None
        if 'batched64' in legs:
            # FP64 twins: same keys, epoch 0, stepped by the ordinary preset
            errs = []
            for (p64, tl), mu32 in zip(B.heterogeneous_specs(1024, seed=1234,
                                                             dtype=np.float64),
                                       mean):
                xt = p64(tl)
                m64 = float(np.asarray(xt.pmean()[-1]).ravel()[0])
                v64 = float(np.asarray(xt.pvar()[-1]).ravel()[0])
                se = math.sqrt(v64 / (p64.paths - 1))
                errs.append((abs(mu32 - m64) / se, abs(mu32 - m64) / max(abs(m64), 1e-300)))
            errs = np.array(errs)
            inst_err = {'max_mean_err_se': float(errs[:, 0].max()),
                        'median_mean_err_se': float(np.median(errs[:, 0]))}
        out['batched'] = {
            'workload': 'c5_1024_gbm_hw2f_heterogeneous_fp32',
            'instances': bt.n, 'tiles': bt.n_items,
            'path_steps': bt.path_steps, 'ms': bms,
            'path_steps_per_s': bt.path_steps / (bms * 1e-3),
            'blocks': bt.blocks, 'cu': bt.cu, 'tile_waves': bt.tile_waves,
            'claim_order': bt.order,
            'autotune_s': {str(k): v for k, v in bt.tuning.items()},
            'prepare_s': prep, 'graph': bt.graph is not None,
            'fp32_vs_fp64_same_paths': inst_err,
            'valu_issue': {
                'kernel': 'k_batch_persistent', 'bound': 'valu',
                'wave_insts_per_launch': BATCH_VALU_WAVE_INSTS,
                'achieved_per_s': BATCH_VALU_WAVE_INSTS / (bms * 1e-3),
                'nominal_per_s': VALU_ISSUE_PEAK_PER_S,
                'frac_of_nominal': BATCH_VALU_WAVE_INSTS / (bms * 1e-3)
                / VALU_ISSUE_PEAK_PER_S}}
        del bt
        torch.cuda.empty_cache()

    if 'cpu' in legs:
        out['cpu_baseline'] = cpu_baseline_subprocess(args.cpu_paths, NSTEPS)
        out['cpu_baseline']['gpu_over_cpu'] = value / out['cpu_baseline']['value']
        # the same restatement path-sharded over the host's cores (at most 16,
        # the CPU share of one GPU)
        workers = max(1, min(16, os.cpu_count() or 1))
        out['cpu_baseline_parallel'] = cpu_baseline_parallel(
            args.cpu_paths // 2, NSTEPS, workers)
        out['cpu_baseline_parallel']['gpu_over_cpu'] = \
            value / out['cpu_baseline_parallel']['value']
    else:
        out['cpu_baseline'] = None
    print(json.dumps(out))


if __name__ == '__main__':
    main()
