"""Benchmark driver (one JSON line on stdout).

Workloads
  mc   : (default) 4096 Monte Carlo 16-bit mirror-adder ripple-carry adders,
         .tran 5p 20n at reltol 1e-3 on the batched engine; a step is one
         whole transient of the batch.  Then, unless --no-guess, the guess
         network is trained on engine waveforms of held-out seeds and the
         same batch runs again from the learned Newton guess.
  eval : 2**20 isolated BSIM4 instances of the 45 nm base card, one
         backward-Euler transient load each plus deterministic assembly of
         their 16 Jacobian entries into a random 2**21-row CSC matrix.  A step
         is one evaluation pass + one assembly pass.
  rings: the uneven batch of 2048 ring oscillators (11..301 stages).
  mesh : the 1024 x 1024 power-grid mesh's Newton matrix (~1.1M unknowns);
         a step is one refactor + one solve by the whole-GPU multifrontal LU.
  mesh_tran: the mesh with its 16384 clocked inverters as a transient
         (ngmi_large_*), timed over a window of its .tran 10p 10n.
  c3   : the rings' transient with the mesh transient running alongside on
         a second stream (configuration 3 together).

Inputs are resident in HBM before the timed region.  Kernel durations come
from HIP events recorded on the stream the kernels are launched on.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, 'ngspice-sf-mirror_amd'))

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# HBM bytes per BSIM4 evaluation that the hardware counters read for the
# default (2-wave, default-card) kernels: rocprofv3 FETCH_SIZE and
# WRITE_SIZE summed over the kernel's dispatches, scaled by what the same
# counters read for a device copy of known size (FETCH_SIZE x 2.00, WRITE_SIZE
# x 1.00), divided by the evaluations the engine counted.  None where not
# measured for the selected variant.
MEASURED_HBM_BYTES_PER_EVAL = {('k_tran_mos', 2): 1764, ('k_bsim4_eval', 2): 1595}
# FP64 FLOP per evaluation from SQ_INSTS_VALU_{ADD,MUL,FMA}_F64 x 64 lanes
# (FMA counted twice), same runs
MEASURED_FLOP_PER_EVAL = {'k_tran_mos': 6617, 'k_bsim4_eval': 6568}
# inverter loads per 1024 x 1024 mesh, scaled with the mesh area
MESH_LOADS_PER_SIDE2 = 16384


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def bench_eval(args):
    import torch
    import ngmi
    import workloads

    n = args.n
    t0 = time.time()
    wl = workloads.device_eval_workload(n=n, seed=1)
    models = [ngmi.Model(workloads.BASE_NMOS), ngmi.Model(workloads.BASE_PMOS)]
    s = ngmi.InstanceSet()
    s.add(models, wl['model_idx'], wl['params'])
    s.finalize()
    icols, mcols, tbytes = s.layout()
    Ap, Ai, e2z = ngmi.csc_from_elements(wl['nrows'], wl['rows_e'], wl['cols_e'])
    nnz = len(Ai)
    gat = ngmi.Gather(nnz, e2z)
    log('setup %.1fs: n=%d inst cols %d model cols %d nnz %d' % (time.time() - t0, n, icols,
                                                                  mcols, nnz))
    st = [torch.from_numpy(x).cuda() for x in workloads.eval_states(wl)]
    st0_init = st[0].clone()
    vt = torch.from_numpy(wl['vterm']).cuda()
    cache = torch.zeros(38 * n, dtype=torch.float64, device='cuda')
    stamp = torch.zeros(16 * n, dtype=torch.float64, device='cuda')
    rhs = torch.zeros(4 * n, dtype=torch.float64, device='cuda')
    nc = torch.zeros(n, dtype=torch.int32, device='cuda')
    ax = torch.zeros(nnz, dtype=torch.float64, device='cuda')
    stream = torch.cuda.current_stream().cuda_stream
    h = workloads.H_STEP
    ag = (1 / h, -1 / h, 0, 0, 0, 0, 0)
    mode = ngmi.MODETRAN | ngmi.MODEINITFLOAT
    ptrs = [x.data_ptr() for x in st]

    def step(t_eval=None, t_gat=None):
        if t_eval:
            t_eval.start(stream)
        s.eval(mode, vt.data_ptr(), ptrs, cache.data_ptr(), stamp.data_ptr(), rhs.data_ptr(),
               nc.data_ptr(), ag=ag, order=1, stream=stream)
        if t_eval:
            t_eval.stop(stream)
        if t_gat:
            t_gat.start(stream)
        gat.run(stamp.data_ptr(), ax.data_ptr(), stream)
        if t_gat:
            t_gat.stop(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    te, tg = ngmi.Timer(), ngmi.Timer()
    ev_ms, ga_ms = [], []
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        step(te, tg)
        ev_ms.append(te.ms())
        ga_ms.append(tg.ms())
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    ev = float(np.mean(ev_ms))
    ga = float(np.mean(ga_ms))
    ms_step = wall / args.steps * 1e3
    # algorithmic bytes of one evaluation launch (per instance): per-instance
    # table columns, 4 terminal voltages, state reads (4 voltages + 3 charges of
    # level 1 + 2 junction charges of level 0), state writes (12 voltage/charge
    # rows + 3 cqX + 4 charges), cache 38 read + 38 written, 16 + 4 outputs,
    # 1 int vote, 1 int model index
    per_inst = 8 * (icols + 4 + 9 + 19 + 76 + 20) + 4 + 4
    achieved = per_inst * n / (ev * 1e-3) / 1e9
    mb = MEASURED_HBM_BYTES_PER_EVAL.get(('k_bsim4_eval', args.eval_waves or 2))
    traffic = mb * n if mb else None
    return {
        'metric': 'BSIM4 device evaluations/s (1M instances, with CSC assembly)',
        'value': n / (ms_step * 1e-3),
        'unit': 'evaluations/s',
        'ms_per_step': ms_step,
        'dtype': 'f64',
        'config': {'workload': 'device-eval: 2^20 BSIM4 instances, 45nm base card, '
                               'random 2^21-row CSC (~8 nnz/row)', 'nnz': nnz},
        'roofline': {'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS, 'traffic': traffic,
                     'kernel': 'k_bsim4_eval', 'kernel_ms': ev, 'bytes_per_instance': per_inst,
                     'fp64_tflops': MEASURED_FLOP_PER_EVAL['k_bsim4_eval'] * n / (ev * 1e-3) / 1e12,
                     'fp64_peak_tflops': FP64_PEAK_TFLOPS},
        'extra': {'gather_ms': ga, 'eval_ms': ev, 'inst_cols': icols, 'model_cols': mcols},
    }


def bench_copy(args):
    """A plain device-to-device copy of a known size: the reference point for
    the HBM byte counters (FETCH_SIZE / WRITE_SIZE) of the profiled kernels."""
    import torch
    n = args.copy_bytes // 8
    a = torch.ones(n, dtype=torch.float64, device='cuda')
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - w0) / args.steps * 1e3
    gbs = 2 * n * 8 / (ms * 1e-3) / 1e9
    return {'metric': 'device copy bandwidth (counter calibration)', 'value': gbs, 'unit': 'GB/s',
            'ms_per_step': ms, 'dtype': 'f64',
            'config': {'workload': 'copy of %d bytes' % (n * 8)},
            'roofline': {'bound': 'hbm', 'achieved': gbs, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                         'frac': gbs / HBM_PEAK_GBS, 'traffic': 2 * n * 8},
            'extra': {'bytes_read': n * 8, 'bytes_written': n * 8}}


def _cpu_worker(job):
    """one host core: adder circuits seeds w, w + P, ... through the oracle's
    DCtran restatement until the time budget is spent; returns (accepted
    points, seconds, circuits, solver)"""
    w, procs, tstop, budget_s = job
    sys.path.insert(0, os.path.join(REPO, 'oracle'))
    import pyoracle as po
    import workloads
    pts, secs, n, solver, seed = 0, 0.0, 0, 'dense', w
    while secs < budget_s and n < 64:
        d = workloads.mc_adder(seed)
        T = po.Transient(d)
        solver = 'klu' if T.use_reference_klu() else 'dense'
        t0 = time.perf_counter()
        rc, t, _ = T.run(d['tstep'], tstop)
        secs += time.perf_counter() - t0
        if rc != 0:
            break
        pts += len(t)
        n += 1
        seed += procs
    return pts, secs, n, solver


def cpu_baseline_mc(tstop, budget_s=12.0, procs=1):
    """The oracle's DCtran restatement on whole adder circuits, solving with
    the reference's own KLU when oracle/_ref was built: `procs` independent
    processes (one per core) each run whole circuits until about budget_s of
    CPU time; throughput = all accepted points / the slowest process's time.
    Runs before the GPU is touched, so the worker processes fork from a
    process without a GPU context."""
    import multiprocessing as mp
    jobs = [(w, procs, tstop, budget_s) for w in range(procs)]
    if procs == 1:
        res = [_cpu_worker(jobs[0])]
    else:
        with mp.get_context('fork').Pool(procs) as pool:
            res = pool.map(_cpu_worker, jobs)
    pts = sum(r[0] for r in res)
    secs = max(r[1] for r in res)
    n = sum(r[2] for r in res)
    solver = res[0][3]
    busy = sum(r[1] for r in res)
    return {'value': pts / secs if secs > 0 else None, 'unit': 'timepoints x circuits/s',
            'cores': procs, 'kind': 'port', 'per_core': pts / busy if busy > 0 else None,
            'sample': '%d adder circuit(s) on %d process(es), one per core, full .tran 5p %gn, '
                      'oracle DCtran restatement with %s linear solve, %d accepted points, '
                      'slowest process %.1f s'
                      % (n, procs, tstop * 1e9,
                         'the reference KLU' if solver == 'klu' else 'dense LU', pts, secs)}


def bench_mc(args):
    import torch
    import ngmi
    import workloads

    C = args.circuits
    tstop = args.tstop
    t0 = time.time()
    descs = [workloads.mc_adder(s, tstop=tstop) for s in range(C)]
    names = descs[0]['names']
    probe = [names['s%d' % i] for i in range(workloads.ADDER_BITS)] + \
        [names['co%d' % (workloads.ADDER_BITS - 1)]]
    circ = [ngmi.Circuit(d) for d in descs]
    del descs
    b = ngmi.Batch(circ, probe=probe, max_points=args.max_points)
    del circ
    info = b.info()
    icols, mcols = b.table_cols()
    log('setup %.1fs: %d circuits, %s, table cols %d/%d' % (time.time() - t0, C, info, icols,
                                                            mcols))
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(args.warmup):
        rc, passes = b.tran(workloads.ADDER_TSTEP, tstop, stream=stream)
        log('warmup rc %d passes %d' % (rc, passes))
    # the timed runs replay captured graphs without per-launch events; the
    # kernels' own durations come from a separate profiled run below
    b.set_profile(False)
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        rc, passes = b.tran(workloads.ADDER_TSTEP, tstop, stream=stream)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - w0)
        if rc != 0:
            raise RuntimeError('transient failed: %d %s' % (rc, ngmi.last_error()))
    wall = float(np.mean(times))
    st = [b.stats(c) for c in range(C)]
    failed = sum(1 for s in st if s['status'] != 0)
    acc = sum(s['accepted'] for s in st)
    rej = sum(s['rejected'] for s in st)
    its = sum(s['iterations'] for s in st)
    _, _, evals, npasses = b.kernel_times()
    # algorithmic bytes of one BSIM4 evaluation in the transient kernel: the
    # per-instance table columns, 4 node voltages, state reads (4 terminal
    # voltages, 3 charges and 3 charge currents of the previous level), 21
    # state writes, 38 cached values read and written, 16 + 4 outputs, the
    # convergence vote and the model index
    b_eval = 8 * (icols + 4 + 4 + 6 + 21 + 38 + 38 + 20) + 4 + 4
    mb = MEASURED_HBM_BYTES_PER_EVAL.get(('k_tran_mos', args.mos_waves or 2))
    log('passes %d, accepted %d, rejected %d, iterations %d, failed %d; wall %.1f ms (%d stream(s), graphs)'
        % (passes, acc, rej, its, failed, wall * 1e3, args.streams or 2))
    single = None
    if not args.no_single_stream:
        # one more run, on one stream and with per-launch events, gives the
        # BSIM4 kernel's own duration: on two streams a half's launch shares
        # the chip with the other half's step kernel
        keep = ngmi.get_tune('streams')
        ngmi.tune('streams', 1)
        b.set_profile(True)
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        rc1, passes1 = b.tran(workloads.ADDER_TSTEP, tstop, stream=stream)
        torch.cuda.synchronize()
        wall1 = time.perf_counter() - w0
        b.set_profile(False)
        ngmi.tune('streams', keep)
        if rc1 != 0:
            raise RuntimeError('single-stream transient failed: %d %s' % (rc1, ngmi.last_error()))
        acc1 = sum(b.stats(c)['accepted'] for c in range(C))
        mos1, step1, evals1, _ = b.kernel_times()
        ach1 = evals1 * b_eval / (mos1 * 1e-3) / 1e9 if mos1 > 0 else None
        single = {'value': acc1 / wall1, 'ms': wall1 * 1e3, 'profiled': True, 'bsim4_kernel_ms_total': mos1,
                  'step_kernel_ms_total': step1,
                  'roofline': {'bound': 'hbm', 'achieved': ach1, 'peak': HBM_PEAK_GBS,
                               'unit': 'GB/s', 'frac': ach1 / HBM_PEAK_GBS if ach1 else None,
                               'traffic': mb * evals1 / passes1 if (mb and passes1) else None,
                               'kernel': 'k_tran_mos', 'launches': passes1,
                               'kernel_ms_total': mos1, 'bytes_per_eval': b_eval, 'evals': evals1,
                               'fp64_tflops': (MEASURED_FLOP_PER_EVAL['k_tran_mos'] * evals1 /
                                               (mos1 * 1e-3) / 1e12) if mos1 > 0 else None,
                               'fp64_peak_tflops': FP64_PEAK_TFLOPS}}
        log('one stream, profiled: %.0f tp*c/s, wall %.1f ms, bsim4 %.1f ms (%.0f GB/s), step %.1f ms'
            % (single['value'], wall1 * 1e3, mos1, ach1 or 0, step1))
    learned = None
    if not args.no_guess:
        learned = learned_guess_run(args, b, probe, stream, acc, its, acc / wall)
    # The headline roofline is priced on launches that do not share the chip
    # (the one-stream profiled run), so the kernel's summed durations stay
    # below that run's wall time.
    if single is not None:
        roof = dict(single['roofline'])
        roof['timed_on'] = 'the one-stream profiled run (kernel time per .tran %.0f ms <= its wall %.0f ms)' \
            % (single['bsim4_kernel_ms_total'], single['ms'])
    else:
        roof = {'bound': 'hbm', 'achieved': None, 'peak': HBM_PEAK_GBS, 'unit': 'GB/s', 'frac': None,
                'traffic': None, 'kernel': 'k_tran_mos', 'timed_on': 'not measured (--no-single-stream)'}
    return {
        'metric': 'MC transient timepoints x circuits/s at reltol 1e-3 (plain vs learned NR guess)',
        'value': acc / wall,
        'unit': 'timepoints x circuits/s',
        'ms_per_step': wall * 1e3,
        'dtype': 'f64',
        'config': {'workload': '%d Monte Carlo 16-bit mirror-adder RCAs (448 BSIM4 each), '
                               '.tran 5p %gn, reltol 1e-3, plain predictor' % (C, tstop * 1e9),
                   'circuits': C, 'unknowns': info['unknowns'],
                   'lu_nonzeros': info['lu_nonzeros']},
        'roofline': roof,
        'extra': {'accepted': acc, 'rejected': rej, 'newton_iterations': its,
                  'max_circuit_iterations': max(s['iterations'] for s in st),
                  'iterations_per_point': its / max(acc, 1), 'passes': passes,
                  'failed_circuits': failed, 'streams': args.streams or 2, 'graphs': True,
                  'one_stream': single, 'learned_guess': learned},
    }


def bench_rings(args):
    """The uneven batch: ring oscillators of 11..301 stages (odd, uniform),
    one topology group per stage count, advanced together; plain predictor."""
    import torch
    import ngmi
    import workloads

    t0 = time.time()
    descs = workloads.uneven_rings(count=args.circuits, tstop=args.tstop)
    b = ngmi.Batch([ngmi.Circuit(d) for d in descs], probe=[1], max_points=args.max_points)
    C = len(descs)
    unknowns = sum(d['nnodes'] + 1 for d in descs)
    mosfets = sum(len(d['mos']) for d in descs)
    del descs
    log('setup %.1fs: %d rings in %d topology groups, %d unknowns, %d MOSFETs'
        % (time.time() - t0, C, b.groups(), unknowns, mosfets))
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(args.warmup):
        rc, passes = b.tran(10e-12, args.tstop, stream=stream)
        log('warmup rc %d passes %d' % (rc, passes))
    # timed on graphs without per-launch events; kernel durations from one
    # more, profiled run
    b.set_profile(False)
    times = []
    for _ in range(args.steps):
        torch.cuda.synchronize()
        w0 = time.perf_counter()
        rc, passes = b.tran(10e-12, args.tstop, stream=stream)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - w0)
        if rc != 0:
            raise RuntimeError('transient failed: %d %s' % (rc, ngmi.last_error()))
    wall = float(np.mean(times))
    st = [b.stats(c) for c in range(C)]
    acc = sum(s['accepted'] for s in st)
    its = sum(s['iterations'] for s in st)
    b.set_profile(True)
    rc2, _ = b.tran(10e-12, args.tstop, stream=stream)
    torch.cuda.synchronize()
    b.set_profile(False)
    if rc2 != 0:
        raise RuntimeError('profiled transient failed: %d %s' % (rc2, ngmi.last_error()))
    mos_ms, step_ms, evals, npasses = b.kernel_times()
    log('rings: passes %d, accepted %d, iterations %d, failed %d; kernels: bsim4 %.1f ms, '
        'step %.1f ms, wall %.1f ms' % (passes, acc, its, sum(1 for s in st if s['status']),
                                        mos_ms, step_ms, wall * 1e3))
    return {
        'metric': 'uneven-batch transient timepoints x circuits/s at reltol 1e-3',
        'value': acc / wall, 'unit': 'timepoints x circuits/s', 'ms_per_step': wall * 1e3,
        'dtype': 'f64',
        'config': {'workload': '%d ring oscillators of 11..301 stages (odd, uniform, seed 3), '
                               '.tran 10p %gn, reltol 1e-3' % (C, args.tstop * 1e9),
                   'circuits': C, 'topology_groups': b.groups(), 'unknowns': unknowns,
                   'mosfets': mosfets},
        'roofline': None,
        'extra': {'accepted': acc, 'newton_iterations': its, 'passes': passes,
                  'iterations_per_point': its / max(acc, 1), 'bsim4_kernel_ms_total': mos_ms,
                  'step_kernel_ms_total': step_ms,
                  'failed_circuits': sum(1 for s in st if s['status'])},
    }


def bench_mesh(args):
    """One large system: the power-grid mesh's Newton matrix (1024 x 1024
    RC mesh, 256 VDD pads, 16384 inverter loads; ~1.1M unknowns) refactored
    and solved by the whole-GPU multifrontal LU.  A step is one refactor
    (klu_refactor) plus one solve (klu_solve) of new values on the fixed
    pattern, as one Newton iteration of the mesh needs."""
    import torch
    import ngmi
    import workloads

    side = args.mesh_side
    loads = MESH_LOADS_PER_SIDE2 * side * side // (1024 * 1024)
    t0 = time.time()
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(side, loads)
    b = workloads.rng(5, side).uniform(-1, 1, n)
    t1 = time.time()
    lu = ngmi.SparseLU(n, Ap, Ai)
    info = lu.info()
    log('mesh %dx%d, %d loads: n %d, nnz %d, matrix %.1fs, analyze %.1fs, %s'
        % (side, side, loads, n, len(Ai), t1 - t0, time.time() - t1, info))
    dA = torch.from_numpy(Ax).cuda()
    db0 = torch.from_numpy(b).cuda()
    db = db0.clone()
    stream = torch.cuda.current_stream().cuda_stream
    for _ in range(args.warmup):
        lu.factor(dA.data_ptr(), stream)
        db.copy_(db0)
        lu.solve(db.data_ptr(), stream)
    torch.cuda.synchronize()
    tf, ts = ngmi.Timer(), ngmi.Timer()
    f_ms = s_ms = 0.0
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for _ in range(args.steps):
        tf.start(stream)
        lu.factor(dA.data_ptr(), stream)
        tf.stop(stream)
        db.copy_(db0)
        ts.start(stream)
        lu.solve(db.data_ptr(), stream)
        ts.stop(stream)
        torch.cuda.synchronize()
        f_ms += tf.ms()
        s_ms += ts.ms()
    wall = (time.perf_counter() - w0) / args.steps
    x = db.cpu().numpy()
    import scipy.sparse as sp
    A = sp.csc_matrix((Ax, Ai, Ap), shape=(n, n))
    res = np.abs(A @ x - b).max() / (abs(A).sum(axis=1).max() * np.abs(x).max() + np.abs(b).max())
    f_ms /= args.steps
    s_ms /= args.steps
    tflops = info['flops'] / (f_ms * 1e-3) / 1e12
    log('mesh: refactor %.3f ms (%.2f TFLOP/s FP64), solve %.3f ms, wall %.3f ms per step, '
        'scaled residual %.2e' % (f_ms, tflops, s_ms, wall * 1e3, res))
    return {
        'metric': 'power-grid mesh refactor + solve per second',
        'value': 1.0 / wall, 'unit': 'refactor+solve/s', 'ms_per_step': wall * 1e3, 'dtype': 'f64',
        'config': {'workload': 'power-grid Newton matrix: %dx%d RC mesh (0.1 ohm, 10 fF), %d VDD pads, '
                               '%d inverter loads; refactor + solve' % (side, side, side * side // 4096,
                                                                        loads),
                   'unknowns': n, 'nonzeros': int(len(Ai)), 'fronts': info['fronts'],
                   'max_front': info['max_front'], 'lu_nonzeros': info['lu_nonzeros']},
        'roofline': {'bound': 'mfma', 'achieved': tflops, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': tflops / FP64_PEAK_TFLOPS, 'traffic': None,
                     'kernel': 'refactor (every k_sp_* launch of one ngmi_splu_factor)',
                     'flops_per_refactor': info['flops'], 'refactor_ms': f_ms},
        'extra': {'solve_ms': s_ms, 'refactor_ms': f_ms, 'scaled_residual': float(res),
                  'front_bytes': info['front_doubles'] * 8},
    }


def bench_c3(args):
    """Configuration 3 together: one window (--mesh-window simulated seconds
    from the operating point) of the 1024 x 1024 power-grid mesh's transient
    runs on a second stream from a host thread, while the uneven batch of
    2048 rings repeats its whole transient on the first stream until the
    mesh window ends.  Reports the rings' rate over that time (beside their
    rate alone) and the mesh's accepted timepoints per second alongside."""
    import threading
    import torch
    import ngmi
    import workloads

    res = bench_rings(args)   # alone first: its warm-up and timed runs
    alone = res['value']
    descs = workloads.uneven_rings(count=args.circuits, tstop=args.tstop)
    b = ngmi.Batch([ngmi.Circuit(d) for d in descs], probe=[1], max_points=args.max_points)
    C = len(descs)
    del descs
    side = args.mesh_side
    d = workloads.power_grid(side=side, tstop=10e-9)
    lg = ngmi.Large(ngmi.Circuit(d), probe=[1], max_points=1000000)
    n = lg.info()['unknowns']
    del d
    ms = torch.cuda.Stream()
    rs = torch.cuda.Stream()
    win = args.mesh_window
    rc, _ = b.tran(10e-12, args.tstop, stream=rs.cuda_stream)   # warm-up of this batch
    torch.cuda.synchronize()
    mesh = {}

    def mesh_window():
        t0 = time.perf_counter()
        mesh['rc'] = lg.tran(10e-12, win, 10e-12, stream=ms.cuda_stream)
        ms.synchronize()
        mesh['s'] = time.perf_counter() - t0
        mesh['accepted'] = lg.stats()['accepted']

    th = threading.Thread(target=mesh_window)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    th.start()
    acc = runs = 0
    while True:
        rc, passes = b.tran(10e-12, args.tstop, stream=rs.cuda_stream)
        rs.synchronize()
        if rc != 0:
            break
        acc += sum(b.stats(c)['accepted'] for c in range(C))
        runs += 1
        if not th.is_alive():
            break
    wall = time.perf_counter() - w0
    th.join()
    if rc != 0:
        raise RuntimeError('transient failed: %d %s' % (rc, ngmi.last_error()))
    if mesh.get('rc', 1) != 0:
        raise RuntimeError('mesh transient failed: %s' % ngmi.last_error())
    log('c3: rings %.0f tp*c/s over %d runs with the mesh transient alongside (%.0f alone); mesh %d accepted '
        'points in %.2f s = %.1f timepoints/s alongside' % (acc / wall, runs, alone, mesh['accepted'], mesh['s'],
                                                           mesh['accepted'] / mesh['s']))
    res['metric'] = 'uneven-batch transient timepoints x circuits/s with the power-grid mesh transient alongside'
    res['value'] = acc / wall
    res['ms_per_step'] = wall * 1e3 / max(runs, 1)
    res['config']['workload'] += ('; alongside: a %g ps window of the %dx%d power-grid mesh transient (%d unknowns, '
                                  '16384 clocked inverters) on a second stream' % (win * 1e12, side, side, n))
    res['extra'].update({'rings_alone': alone, 'rings_runs_alongside': runs,
                         'mesh_timepoints_per_s_alongside': mesh['accepted'] / mesh['s'],
                         'mesh_accepted_points': mesh['accepted'], 'mesh_window_s': mesh['s'],
                         'accepted_alongside': acc})
    return res


MESH_TRAN_WINDOW = 3e-12   # the first 3 ps of configuration 3's .tran 10p 10n


def bench_mesh_tran(args):
    """Configuration 3's mesh as a transient: the 1024 x 1024 power grid
    with 16384 clocked inverter loads (1.1M unknowns), .tran 10p 10n with
    TMAX 10 ps, run over a window (its first --mesh-window seconds of
    simulated time: the 16384 clock phases put a PULSE corner every ~15 fs,
    so the full 10 ns is about 6.5e5 accepted points).  A step is the whole
    window; value = accepted timepoints per second (one circuit)."""
    import torch
    import ngmi
    import workloads

    side = args.mesh_side
    t0 = time.time()
    d = workloads.power_grid(side=side, tstop=10e-9)
    nn = d['nnodes']
    probe = [1, 1 + (side // 2) * side + side // 2, nn]
    lg = ngmi.Large(ngmi.Circuit(d), probe=probe, max_points=1000000)
    info = lg.info()
    del d
    log('mesh transient setup %.1fs: %s' % (time.time() - t0, info))
    stream = torch.cuda.current_stream().cuda_stream
    win = args.mesh_window
    rc = lg.tran(10e-12, win, 10e-12, stream=stream)     # warm-up (code objects, fronts)
    if rc != 0:
        raise RuntimeError('mesh transient failed: %d %s' % (rc, ngmi.last_error()))
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    rc = lg.tran(10e-12, win, 10e-12, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    if rc != 0:
        raise RuntimeError('mesh transient failed: %d %s' % (rc, ngmi.last_error()))
    st = lg.stats()
    tm = lg.times()
    kcl, per = lg.kcl()
    its = st['iterations']
    refactor_ms = tm['factor_ms'] / max(its, 1)
    tflops = info['flops'] / (refactor_ms * 1e-3) / 1e12 if refactor_ms > 0 else None
    log('mesh transient: %d accepted, %d rejected, %d iterations in %.2f s (%.1f ms/iteration: load %.2f, '
        'refactor %.2f, solve %.2f, conv %.2f); KCL %.2e'
        % (st['accepted'], st['rejected'], its, wall, wall * 1e3 / max(its, 1), tm['load_ms'] / max(its, 1),
           refactor_ms, tm['solve_ms'] / max(its, 1), tm['conv_ms'] / max(its, 1), kcl))
    return {
        'metric': 'power-grid mesh transient timepoints/s (windowed)',
        'value': st['accepted'] / wall, 'unit': 'timepoints/s', 'ms_per_step': wall * 1e3, 'dtype': 'f64',
        'config': {'workload': 'power-grid mesh %dx%d, 16384 clocked inverter loads (%d unknowns), .tran 10p '
                               '10n (TMAX 10p) over its first %g ps' % (side, side, info['unknowns'], win * 1e12),
                   'unknowns': info['unknowns'], 'lu_nonzeros': info['lu_nonzeros']},
        'roofline': {'bound': 'mfma', 'achieved': tflops, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': tflops / FP64_PEAK_TFLOPS if tflops else None, 'traffic': None,
                     'kernel': 'refactor (every k_sp_* launch of one ngmi_splu_factor) inside the transient',
                     'flops_per_refactor': info['flops'], 'refactor_ms': refactor_ms},
        'extra': {'accepted': st['accepted'], 'rejected': st['rejected'], 'newton_iterations': its,
                  'ms_per_iteration': wall * 1e3 / max(its, 1),
                  'phase_ms_per_iteration': {k: v / max(its, 1) for k, v in tm.items()},
                  'kcl_scaled_residual_final_point': kcl, 'windowed': True},
    }


def cpu_baseline_mesh(side):
    """The reference KLU (oracle/_ref) on the same matrix: klu_factor once,
    then one timed klu_refactor and the solve, on one host core."""
    sys.path.insert(0, os.path.join(REPO, 'oracle'))
    import pyoracle as po
    import workloads
    loads = MESH_LOADS_PER_SIDE2 * side * side // (1024 * 1024)
    n, Ap, Ai, Ax = workloads.power_grid_jacobian(side, loads)
    b = workloads.rng(5, side).uniform(-1, 1, n)
    got = po.klu_csc(n, Ap, Ai, Ax, b, refactors=1)
    if got is None:
        return None
    _, tf, tr = got
    return {'value': 1.0 / tr, 'unit': 'refactor+solve/s (refactor only)', 'cores': 1, 'kind': 'reference',
            'sample': 'the reference KLU (klu_factor once, then one timed klu_refactor) on the same '
                      '%dx%d mesh matrix; klu_factor took %.1f s' % (side, side, tf)}


def learned_guess_run(args, b, probe, stream, plain_acc, plain_its, plain_value):
    """Train the guess network on engine waveforms of held-out seeds, then
    run the same batch with the learned guess; returns its figures and the
    waveform error of circuit 0 (plain and learned) against the oracle."""
    import torch
    import ngmi
    import workloads
    import guess

    t0 = time.time()
    seeds = range(4096, 4096 + args.train_circuits)
    descs = [workloads.mc_adder(s, tstop=args.tstop) for s in seeds]
    nn = descs[0]['nnodes']
    tb = ngmi.Batch([ngmi.Circuit(d) for d in descs], max_points=args.max_points)
    rc, _ = tb.tran(workloads.ADDER_TSTEP, args.tstop, stream=stream)
    if rc != 0:
        raise RuntimeError('training-data run failed: %s' % ngmi.last_error())
    nbr = tb.neighbours(nn)
    n = tb.info()['unknowns']
    Fs, Ts = [], []
    for c in range(len(descs)):
        t, x = tb.waveform(c, n)
        F, T = guess.features(t, x, nbr, nn, every=2)
        Fs.append(F)
        Ts.append(T)
    del tb
    F = np.concatenate(Fs)
    T = np.concatenate(Ts)
    g = np.random.default_rng(4096)
    sel = g.permutation(len(F))[:min(len(F), 2_000_000)]
    F, T = F[sel], T[sel]
    base_ms = float(np.mean(T.astype(np.float64) ** 2))
    t1 = time.time()
    net = ngmi.GuessNet(seed=4096)
    dF = torch.from_numpy(F).cuda()
    dT = torch.from_numpy(T).cuda()
    M = 8192
    for it in range(args.train_steps):
        k = (it * M) % max(len(F) - M, 1)
        net.train_step(dF[k:].data_ptr(), dT[k:].data_ptr(), M, 1e-3 if it < args.train_steps // 2
                       else 3e-4, want_loss=False, stream=stream)
    loss = net.train_step(dF.data_ptr(), dT.data_ptr(), min(len(F), 262144), 0.0, stream=stream)
    torch.cuda.synchronize()
    t2 = time.time()
    log('guess training: %d rows, baseline %.4e -> loss %.4e, data %.1fs, train %.1fs'
        % (len(F), base_ms, loss, t1 - t0, t2 - t1))
    ora = None if args.no_cpu_baseline else oracle_waveform(args)
    plain_err = waveform_error(b, probe, ora) if ora else None
    b.set_guess(net)
    # untimed warm-up runs, as for the plain run: the learned passes are
    # captured as graphs of their own and the guess buffers are first used
    for _ in range(max(args.warmup, 1)):
        rc, _ = b.tran(workloads.ADDER_TSTEP, args.tstop, stream=stream)
        if rc != 0:
            raise RuntimeError('learned warm-up run failed: %s' % ngmi.last_error())
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    rc, passes = b.tran(workloads.ADDER_TSTEP, args.tstop, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    if rc != 0:
        raise RuntimeError('learned run failed: %s' % ngmi.last_error())
    C = b.C
    st = [b.stats(c) for c in range(C)]
    gs = [b.guess_stats(c) for c in range(C)]
    acc = sum(s['accepted'] for s in st)
    its = sum(s['iterations'] for s in st)
    rows = b.guess_rows()
    out = {'value': acc / wall, 'unit': 'timepoints x circuits/s', 'ms': wall * 1e3,
           'vs_plain': (acc / wall) / plain_value,
           'accepted': acc, 'newton_iterations': its, 'iterations_per_point': its / max(acc, 1),
           'plain_iterations_per_point': plain_its / max(plain_acc, 1),
           'steps_from_guess': sum(x['guessed'] for x in gs),
           'fallbacks_to_predictor': sum(x['fallbacks'] for x in gs),
           'steps_paused': sum(x['paused'] for x in gs),
           'network_rows': rows, 'network_rows_per_pass': rows / max(passes, 1), 'passes': passes,
           'rejected': sum(s['rejected'] for s in st),
           'max_circuit_iterations': max(s['iterations'] for s in st),
           'slowest_circuit': int(np.argmax([s['iterations'] for s in st])),
           'failed_circuits': sum(1 for s in st if s['status']),
           'training': {'seeds': '%d..%d' % (seeds[0], seeds[-1]), 'rows': int(len(F)),
                        'steps': args.train_steps, 'target_mean_square': base_ms,
                        'final_loss': loss, 'seconds': t2 - t0}}
    if not args.no_single_stream:
        # the guess kernels' own durations: the same learned run once more,
        # one stream, with per-launch events
        keep = ngmi.get_tune('streams')
        ngmi.tune('streams', 1)
        b.set_profile(True)
        rc2, passes2 = b.tran(workloads.ADDER_TSTEP, args.tstop, stream=stream)
        torch.cuda.synchronize()
        b.set_profile(False)
        ngmi.tune('streams', keep)
        if rc2 != 0:
            raise RuntimeError('profiled learned run failed: %s' % ngmi.last_error())
        mos_ms, step_ms, _, _ = b.kernel_times()
        g_ms = b.guess_stats(0)['guess_ms']
        out.update({'profiled': {'guess_kernels_ms_total': g_ms, 'guess_us_per_pass': g_ms * 1e3 / max(passes2, 1),
                                 'bsim4_kernel_ms_total': mos_ms, 'step_kernel_ms_total': step_ms,
                                 'passes': passes2}})
    log('learned: %.0f tp*c/s (%.3f x plain), %.3f it/pt (plain %.3f), %d fallbacks, %.0f network rows/pass%s'
        % (out['value'], out['vs_plain'], out['iterations_per_point'], out['plain_iterations_per_point'],
           out['fallbacks_to_predictor'], out['network_rows_per_pass'],
           (', guess kernels %.1f us/pass' % out['profiled']['guess_us_per_pass']) if 'profiled' in out else ''))
    if ora:
        out['waveform_error_vs_oracle'] = {'plain': plain_err,
                                           'learned': waveform_error(b, probe, ora)}
    b.set_guess(None)
    return out


def oracle_waveform(args):
    """the oracle's DCtran restatement (reference KLU) on circuit 0 (seed 0)"""
    sys.path.insert(0, os.path.join(REPO, 'oracle'))
    import pyoracle as po
    import workloads
    d = workloads.mc_adder(0, tstop=args.tstop)
    T = po.Transient(d)
    T.use_reference_klu()
    rc, to, xo = T.run(d['tstep'], args.tstop)
    return (to, xo) if rc == 0 else None


def waveform_error(b, probe, ora):
    """max |v_engine - v_oracle| over the probed outputs of circuit 0 at the
    oracle's timepoints (engine waveform interpolated where the timepoints
    differ)"""
    to, xo = ora
    te, xe = b.waveform(0, len(probe))
    err = max(float(np.abs(np.interp(to, te, xe[:, k]) - xo[:, u]).max())
              for k, u in enumerate(probe))
    same = len(te) == len(to) and bool(np.allclose(te, to, rtol=1e-9, atol=0))
    return {'circuit': 0, 'max_abs_V': err, 'oracle_points': int(len(to)),
            'engine_points': int(len(te)), 'same_timepoints': same}


def host_cores():
    """Cores this process may run on (sched_getaffinity), os.cpu_count(), and
    the count the CPU baseline uses: the affinity count, capped by
    OMP_NUM_THREADS where the host sets it (its CPU share per GPU)."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else (os.cpu_count() or 1)
    omp = os.environ.get('OMP_NUM_THREADS')
    used = aff
    if omp and omp.isdigit() and int(omp) > 0:
        used = min(aff, int(omp))
    return {'affinity_cores': aff, 'cpu_count': os.cpu_count(), 'omp_num_threads': omp, 'used': used}


C0_TSTOP = 200e-9
C0_CPU_SAMPLE_TSTOP = 20e-9


def cpu_baseline_c0():
    """The oracle's DCtran restatement (with the reference KLU when built) on
    the 101-stage ring at one thread, over the first 20 ns of its .tran 10p
    200n (a bounded sample: about 2000 of its 20000 accepted points)."""
    sys.path.insert(0, os.path.join(REPO, 'oracle'))
    import pyoracle as po
    import workloads
    d = workloads.ring_oscillator()
    T = po.Transient(d)
    klu = T.use_reference_klu()
    t0 = time.perf_counter()
    rc, t, _ = T.run(d['tstep'], C0_CPU_SAMPLE_TSTOP)
    secs = time.perf_counter() - t0
    if rc != 0:
        return None
    return {'value': len(t) / secs, 'unit': 'timepoints/s', 'cores': 1, 'kind': 'port',
            'sample': '101-stage ring, first %g ns of .tran 10p 200n: %d accepted points in %.2f s, '
                      'oracle DCtran restatement with %s' % (C0_CPU_SAMPLE_TSTOP * 1e9, len(t), secs,
                                                             'the reference KLU' if klu else 'dense LU')}


def bench_c0(cpu):
    """The reference's CPU-sized case on the GPU: one 101-stage ring,
    .ic v(1)=0.51, .tran 10p 200n, reltol 1e-3, trapezoidal; one circuit, so
    every Newton pass is launch-bound.  Timed once after one warm-up run."""
    import torch
    import ngmi
    import workloads
    d = workloads.ring_oscillator(tstop=C0_TSTOP)
    b = ngmi.Batch([ngmi.Circuit(d)], probe=[1], max_points=32768)
    stream = torch.cuda.current_stream().cuda_stream
    b.tran(d['tstep'], C0_TSTOP, stream=stream)
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    rc, passes = b.tran(d['tstep'], C0_TSTOP, stream=stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    if rc != 0:
        raise RuntimeError('101-stage ring transient failed: %d %s' % (rc, ngmi.last_error()))
    st = b.stats(0)
    out = {'workload': '1 circuit: 101-stage ring (202 BSIM4), .tran 10p 200n, reltol 1e-3',
           'value': st['accepted'] / wall, 'unit': 'timepoints/s', 'ms': wall * 1e3,
           'accepted': st['accepted'], 'newton_iterations': st['iterations'], 'passes': passes,
           'us_per_pass': wall * 1e6 / max(passes, 1), 'cpu_baseline': cpu}
    if cpu and cpu.get('value'):
        out['gpu_over_cpu'] = out['value'] / cpu['value']
    log('c0: %.0f timepoints/s on the GPU (%d points, %d passes, %.1f us per pass)%s'
        % (out['value'], st['accepted'], passes, out['us_per_pass'],
           '; CPU %.0f/s at 1 thread' % cpu['value'] if cpu and cpu.get('value') else ''))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='timed steps (mc: whole .tran runs of the batch; default 1)')
    ap.add_argument('--warmup', type=int, default=None,
                    help='untimed steps before (mc default 1)')
    ap.add_argument('--workload', default='mc',
                    choices=['mc', 'eval', 'copy', 'rings', 'mesh', 'mesh_tran', 'c3', 'c0'])
    ap.add_argument('--mesh-window', type=float, default=MESH_TRAN_WINDOW,
                    help='mesh_tran / c3: simulated seconds of the mesh transient timed')
    ap.add_argument('--mesh-side', type=int, default=1024, help='mesh: nodes per side')
    ap.add_argument('--copy-bytes', type=int, default=1 << 30)
    ap.add_argument('--n', type=int, default=1 << 20)
    ap.add_argument('--circuits', type=int, default=None,
                    help='adders (mc, default 4096) or rings (rings, default 2048)')
    ap.add_argument('--tstop', type=float, default=None,
                    help='mc default 20 ns, rings default 10 ns')
    ap.add_argument('--max-points', type=int, default=8192)
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--cpu-procs', type=int, default=None,
                    help='host processes for the CPU baseline (one per core); default: the cores '
                         'this process may run on, capped by OMP_NUM_THREADS where that is set')
    ap.add_argument('--no-guess', action='store_true', help='skip the learned-guess run')
    ap.add_argument('--no-c0', action='store_true',
                    help='mc: skip the single-circuit 101-stage ring run reported in extra.c0')
    ap.add_argument('--no-single-stream', action='store_true',
                    help='skip the extra one-stream run of the default two-stream mode')
    ap.add_argument('--eval-waves', type=int, default=None, help='BSIM4 eval kernel occupancy')
    ap.add_argument('--mos-waves', type=int, default=None, help='transient BSIM4 kernel occupancy')
    ap.add_argument('--guess-fused', type=int, default=None,
                    help='learned guess network: 0 layer by layer, 1 fused (weights in registers), '
                         '2 fused with weights streamed per tile')
    ap.add_argument('--guess-grid', type=int, default=None,
                    help='workgroups of the fused guess network kernel (0 = one per compute unit)')
    ap.add_argument('--specialise', type=int, default=None,
                    help='0: BSIM4 kernels read the selector word per instance')
    ap.add_argument('--streams', type=int, default=None,
                    help='2: the plain transient alternates circuit halves on two streams')
    ap.add_argument('--order-slack', type=int, default=None,
                    help='pivot order: degree slack traded for a shorter elimination tree')
    ap.add_argument('--train-circuits', type=int, default=32)
    ap.add_argument('--train-steps', type=int, default=1500)
    args = ap.parse_args()
    if args.circuits is None:
        args.circuits = 2048 if args.workload in ('rings', 'c3') else 4096
    if args.tstop is None:
        args.tstop = 10e-9 if args.workload in ('rings', 'c3') else 20e-9
    if args.steps is None:
        args.steps = 1 if args.workload in ('mc', 'rings', 'c3', 'c0') else 20
    if args.warmup is None:
        args.warmup = 1 if args.workload in ('mc', 'rings', 'c3', 'c0') else 3
    host = host_cores()
    if args.cpu_procs is None:
        args.cpu_procs = host['used']
    if args.gpus != 1:
        print('bench.py runs on one GPU only (the engine has no multi-GPU path); '
              'got --gpus %d' % args.gpus, file=sys.stderr)
        sys.exit(2)
    cpu = None
    c0_cpu = None
    if args.workload == 'mc' and not args.no_cpu_baseline:
        cpu = cpu_baseline_mc(args.tstop, procs=args.cpu_procs)
        cpu['host'] = host
        log('cpu baseline: %s' % cpu)
        if not args.no_c0:
            c0_cpu = cpu_baseline_c0()
            log('cpu baseline, 101-stage ring: %s' % c0_cpu)
    if args.workload == 'mesh' and not args.no_cpu_baseline:
        cpu = cpu_baseline_mesh(args.mesh_side)
        log('cpu baseline: %s' % cpu)
    if args.workload == 'c0' and not args.no_cpu_baseline:
        c0_cpu = cpu_baseline_c0()
        log('cpu baseline, 101-stage ring: %s' % c0_cpu)
    import torch
    if not torch.cuda.is_available():
        print('bench.py needs a GPU', file=sys.stderr)
        sys.exit(2)
    import ngmi
    if args.eval_waves:
        ngmi.tune('eval_waves', args.eval_waves)
    if args.mos_waves:
        ngmi.tune('mos_waves', args.mos_waves)
    if args.guess_fused is not None:
        ngmi.tune('guess_fused', args.guess_fused)
    if args.guess_grid is not None:
        ngmi.tune('guess_grid', args.guess_grid)
    if args.streams:
        ngmi.tune('streams', args.streams)
    if args.specialise is not None:
        ngmi.tune('specialise', args.specialise)
    if args.order_slack is not None:
        ngmi.tune('order_slack', args.order_slack)
    if args.workload == 'mc':
        res = bench_mc(args)
        if not args.no_c0:
            res['extra']['c0'] = bench_c0(c0_cpu)
    elif args.workload == 'copy':
        res = bench_copy(args)
        cpu = None
    elif args.workload == 'rings':
        res = bench_rings(args)
        cpu = None
    elif args.workload == 'mesh':
        res = bench_mesh(args)
    elif args.workload == 'mesh_tran':
        res = bench_mesh_tran(args)
        cpu = None
    elif args.workload == 'c3':
        res = bench_c3(args)
        cpu = None
    elif args.workload == 'c0':
        # the reference's CPU-sized case alone: one 101-stage ring
        c0 = bench_c0(c0_cpu)
        res = {'metric': 'single-circuit transient timepoints/s', 'value': c0['value'], 'unit': 'timepoints/s',
               'ms_per_step': c0['ms'], 'dtype': 'f64', 'config': {'workload': c0['workload']},
               'roofline': None, 'extra': c0}
        cpu = c0_cpu
    else:
        res = bench_eval(args)
        cpu = None
    out = {'metric': res['metric'], 'value': res['value'], 'unit': res['unit'], 'n_gpus': 1,
           'steps': args.steps, 'warmup': args.warmup, 'ms_per_step': res['ms_per_step'],
           'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
           'dtype': res['dtype'], 'data': 'synthetic (seeded Philox)', 'config': res['config'],
           'roofline': res['roofline'], 'cpu_baseline': cpu, 'extra': res['extra']}
    print(json.dumps(out))


if __name__ == '__main__':
    main()
