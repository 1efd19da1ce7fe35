"""The learned Newton initial guess inside the batched engine.

Trained on waveforms the engine generates for held-out adder seeds, the
network must (1) learn: its loss on the training rows falls well below the
linear-extrapolation baseline it corrects (target 0, the loss is in units
of each node's recent motion); (2) generalise: on the evaluation seeds'
waveforms its guesses are closer to the next accepted voltages than the
linear extrapolation, in volts (RMS over the active nodes); (3) leave
every circuit converging: all circuits finish, steps whose Newton run from
the guess fails are redone from the predictor (counted as fallbacks), and
the Newton iterations per point stay within 10 % of the plain predictor's
(NIiter never stops before its second iteration, so 2 per point is the
floor both runs share); (4) keep the waveforms: the learned run computes
the same logic values and its node voltages at the plain run's timepoints
stay within 50 mV (the converged points of both runs satisfy the same
reltol 1e-3 / vntol 1e-6 Newton test; timepoints differ, and interpolating
steep edges between them is the residual)."""
import numpy as np
import pytest

from engine_path import ngmi, workloads
import guess
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(__file__), '..', 'oracle'))
import pyoracle as po  # noqa: E402

torch = pytest.importorskip('torch')
pytestmark = pytest.mark.gpu

TSTOP = 2.5e-9


def run(seeds, net=None, probe=None):
    descs = [workloads.mc_adder(s, tstop=TSTOP) for s in seeds]
    circ = [ngmi.Circuit(d) for d in descs]
    b = ngmi.Batch(circ, probe=probe, max_points=4096)
    if net is not None:
        b.set_guess(net)
    rc, passes = b.tran(workloads.ADDER_TSTEP, TSTOP)
    assert rc == 0, ngmi.last_error()
    return b, descs


@pytest.fixture(scope='module')
def trained():
    """the guess network trained on engine waveforms of held-out seeds
    4096..4103 (.tran 5p 2.5n); returns (net, training rows, loss, baseline)"""
    print('case: guess network training, seeds 4096..4103', flush=True)
    train_seeds = list(range(4096, 4104))
    b, descs = run(train_seeds)
    nn = descs[0]['nnodes']
    nbr = b.neighbours(nn)
    n = b.info()['unknowns']
    Fs, Ts = [], []
    for c in range(len(train_seeds)):
        t, x = b.waveform(c, n)
        F, T = guess.features(t, x, nbr, nn)
        Fs.append(F)
        Ts.append(T)
    F = np.concatenate(Fs)
    T = np.concatenate(Ts)
    rng = np.random.default_rng(0)
    sel = rng.choice(len(F), size=min(len(F), 200000), replace=False)
    F, T = F[sel], T[sel]
    base_loss = float(np.mean(T.astype(np.float64) ** 2))
    net = ngmi.GuessNet(seed=1)
    dF = torch.from_numpy(F).cuda()
    dT = torch.from_numpy(T).cuda()
    M = 8192
    for it in range(1200):
        k = (it * M) % (len(F) - M)
        net.train_step(dF[k:].data_ptr(), dT[k:].data_ptr(), M, 1e-3 if it < 600 else 3e-4,
                       want_loss=False)
    loss = net.train_step(dF.data_ptr(), dT.data_ptr(), min(len(F), 65536), 0.0)
    print('training rows %d, target mean square %.4e, trained loss %.4e' % (len(F), base_loss, loss),
          flush=True)
    return net, len(F), loss, base_loss


def test_learned_guess_end_to_end(trained):
    print('case: learned guess, train seeds 4096..4103, eval seeds 0..3', flush=True)
    net, _, loss, base_loss = trained
    assert loss < 0.5 * base_loss
    nn = workloads.mc_adder(0)['nnodes']

    eval_seeds = [0, 1, 2, 3]
    bp, dp = run(eval_seeds)
    nbr = bp.neighbours(nn)
    n = bp.info()['unknowns']
    Fs, Ts, As = [], [], []
    for c in range(len(eval_seeds)):
        t, x = bp.waveform(c, n)
        F, T, A = guess.features(t, x, nbr, nn, with_scale=True)
        Fs.append(F)
        Ts.append(T)
        As.append(A)
    Fe, Te, Ae = np.concatenate(Fs), np.concatenate(Ts), np.concatenate(As)
    dFe = torch.from_numpy(Fe).cuda()
    dye = torch.zeros(len(Fe), dtype=torch.float32, device='cuda')
    net.forward(dFe.data_ptr(), dye.data_ptr(), len(Fe))
    torch.cuda.synchronize()
    ye = np.clip(dye.cpu().numpy(), -guess.CLAMP, guess.CLAMP).astype(np.float64)
    rms_lin = float(np.sqrt(np.mean((Te * Ae) ** 2)))
    rms_net = float(np.sqrt(np.mean(((Te - ye) * Ae) ** 2)))
    print('evaluation seeds: %d active rows, RMS error of the next accepted voltage: linear '
          'extrapolation %.3e V, learned guess %.3e V' % (len(Fe), rms_lin, rms_net), flush=True)
    assert rms_net < rms_lin
    bl, _ = run(eval_seeds, net=net)
    its_plain = its_learned = 0
    for c in range(len(eval_seeds)):
        sp, sl, g = bp.stats(c), bl.stats(c), bl.guess_stats(c)
        print('circuit %d: plain %d pts %d iters (%.3f/pt); learned %d pts %d iters (%.3f/pt), '
              '%d steps from the guess, %d fallbacks'
              % (c, sp['accepted'], sp['iterations'], sp['iterations'] / sp['accepted'],
                 sl['accepted'], sl['iterations'], sl['iterations'] / sl['accepted'],
                 g['guessed'], g['fallbacks']), flush=True)
        assert sl['status'] == 0
        assert g['guessed'] > 0
        its_plain += sp['iterations'] / sp['accepted']
        its_learned += sl['iterations'] / sl['accepted']
        tp, xp = bp.waveform(c, n)
        tl, xl = bl.waveform(c, n)
        err = max(np.abs(np.interp(tp, tl, xl[:, k]) - xp[:, k]).max() for k in range(nn))
        print('  max |v_learned - v_plain| at plain timepoints: %.3e V' % err, flush=True)
        assert err < 0.05
        names = dp[c]['names']
        for i in range(workloads.ADDER_BITS):
            u = names['s%d' % i] - 1
            assert (xp[-1, u] > 0.5) == (xl[-1, u] > 0.5)
    print('Newton iterations per point: plain %.3f, learned %.3f'
          % (its_plain / len(eval_seeds), its_learned / len(eval_seeds)), flush=True)
    assert its_learned <= 1.10 * its_plain


def test_features_match_the_engine_bit_for_bit():
    """guess.features (the host side that builds training rows) and the
    engine's features kernel (through ngmi_guess_features, the device
    function the in-loop kernel calls) give the same float32 rows, bit for
    bit, on the windows of real adder waveforms; the rule that drops rows
    the network cannot matter for (activity below 10 uV, or 4 x activity
    within the Newton tolerance) selects the same rows."""
    print('case: guess feature rows, host vs engine, adder seeds 0..1 .tran 5p 2.5n', flush=True)
    b, descs = run([0, 1])
    nn = descs[0]['nnodes']
    nbr = b.neighbours(nn).reshape(-1, 8)
    n = b.info()['unknowns']
    checked = 0
    for c in range(2):
        t, x = b.waveform(c, n)
        F_host, T_host, A_host = guess.features(t, x, nbr.reshape(-1), nn, with_scale=True)
        P = len(t)
        steps = np.arange(3, P - 1)
        V = x[:, :nn]
        S = len(steps)
        # every (step, node) window, step-major then node, the order guess.features keeps
        t4 = np.stack([t[steps], t[steps - 1], t[steps - 2], t[steps - 3]], axis=1)     # [S, 4]
        t4 = np.repeat(t4[:, None, :], nn, axis=1).reshape(-1, 4)
        h = np.repeat((t[steps + 1] - t[steps])[:, None], nn, axis=1).reshape(-1)
        v4 = np.stack([V[steps], V[steps - 1], V[steps - 2], V[steps - 3]], axis=2).reshape(-1, 4)
        nb = nbr[1:nn + 1]                                                                # [nn, 8]
        mask = ((nb >= 1).astype(np.uint32) << np.arange(8, dtype=np.uint32)).sum(axis=1).astype(np.uint32)
        nbv = np.where(nb[None, :, :] >= 1, V[steps][:, np.clip(nb - 1, 0, nn - 1)], 0.0).reshape(-1, 8)
        dev = {k: torch.from_numpy(np.ascontiguousarray(a)).cuda()
               for k, a in (('t4', t4), ('h', h), ('v4', v4), ('nbv', nbv),
                            ('mask', np.tile(mask, S).astype(np.int32)))}
        F_dev, act_dev = ngmi.guess_features(dev['t4'], dev['h'], dev['v4'], dev['nbv'], dev['mask'])
        torch.cuda.synchronize()
        F_dev = F_dev.cpu().numpy()
        act_dev = act_dev.cpu().numpy()
        keep = act_dev > 0
        assert keep.sum() == len(F_host), (keep.sum(), len(F_host))
        F_eng = F_dev[keep]
        same = F_eng.view(np.uint32) == F_host.view(np.uint32)
        print('circuit %d: %d windows, %d active rows, columns not bit-equal: %s'
              % (c, len(h), len(F_host), sorted(set(np.nonzero(~same)[1].tolist()))), flush=True)
        assert same.all(), np.argwhere(~same)[:5]
        assert np.array_equal(act_dev[keep], A_host)
        checked += len(F_host)
    assert checked > 1000


@pytest.mark.parametrize('seed', [0, 1])
def test_learned_guess_full_length_against_oracle(trained, seed):
    """A learned run of one adder at full length (.tran 5p 20n) against the
    oracle.  The learned run's Newton iterates start elsewhere, so its
    converged points differ from the oracle's within the Newton tolerance
    (reltol 1e-3 of the value + vntol 1e-6 V, niconv.c:50-60), and its LTE
    control may place timepoints differently; the bound is that tolerance
    at every oracle timepoint (the learned waveform interpolated there), and
    the error is printed per probed output (the 16 sums and the carry)."""
    net = trained[0]
    d = workloads.mc_adder(seed)
    tstop = workloads.ADDER_TSTOP
    names = d['names']
    outs = ['s%d' % i for i in range(workloads.ADDER_BITS)] + ['co%d' % (workloads.ADDER_BITS - 1)]
    probe = [names[k] for k in outs]
    print('case: learned guess, adder seed %d, .tran 5p %gn, against the oracle' % (seed, tstop * 1e9),
          flush=True)
    b = ngmi.Batch([ngmi.Circuit(d)], probe=probe, max_points=8192)
    b.set_guess(net)
    rc, _ = b.tran(workloads.ADDER_TSTEP, tstop)
    assert rc == 0, ngmi.last_error()
    st, g = b.stats(0), b.guess_stats(0)
    assert st['status'] == 0 and g['guessed'] > 0
    te, xe = b.waveform(0, len(probe))
    T = po.Transient(d)
    T.use_reference_klu()
    rc, to, xo = T.run(d['tstep'], tstop)
    assert rc == 0
    same = len(te) == len(to) and bool(np.allclose(te, to, rtol=1e-9, atol=0))
    worst = 0.0
    per = []
    for k, u in enumerate(probe):
        ref = xo[:, u]
        got = np.interp(to, te, xe[:, k])
        err = np.abs(got - ref)
        tol = 1e-3 * np.abs(ref) + 1e-6
        per.append('%s %.1e' % (outs[k], err.max()))
        worst = max(worst, float((err / tol).max()))
    print('learned: %d points (oracle %d, same timepoints: %s), %d steps from the guess, %d fallbacks; '
          'max |v - v_oracle| per output: %s; worst error / Newton tolerance %.2e'
          % (len(te), len(to), same, g['guessed'], g['fallbacks'], ', '.join(per), worst), flush=True)
    assert worst <= 1.0


def test_fused_network_kernel_matches_layered_path(trained):
    """The transient's fused network kernel (features -> three layers ->
    head -> starting values in one launch, weights in registers) against the
    layer-by-layer kernels it replaces: the same learned runs, bit for bit
    (waveforms, Newton statistics and the number of network rows), since
    both issue the same matrix-core steps and the same element-wise code;
    and the fused kernel with weights streamed per tile (guess_fused 2)
    against both."""
    net = trained[0]
    seeds = [0, 1, 2, 3]
    print('case: fused vs layered vs streamed guess network, adder seeds 0..3 .tran 5p %gn' % (TSTOP * 1e9),
          flush=True)
    out = {}
    keep = ngmi.get_tune('guess_fused')
    try:
        for fused in (1, 0, 2):
            ngmi.tune('guess_fused', fused)
            b, descs = run(seeds, net=net)
            n = b.info()['unknowns']
            out[fused] = (b.guess_rows(), [b.waveform(c, n) for c in range(len(seeds))],
                          [b.stats(c) for c in range(len(seeds))],
                          [b.guess_stats(c)['guessed'] for c in range(len(seeds))])
    finally:
        ngmi.tune('guess_fused', keep)
    rows_f, wf_f, st_f, g_f = out[1]
    rows_l, wf_l, st_l, g_l = out[0]
    print('network rows: fused %d, layered %d; steps from the guess %s / %s'
          % (rows_f, rows_l, g_f, g_l), flush=True)
    assert rows_f == rows_l and rows_f > 0
    assert g_f == g_l
    for c in range(len(seeds)):
        assert st_f[c]['accepted'] == st_l[c]['accepted']
        assert st_f[c]['iterations'] == st_l[c]['iterations']
        assert np.array_equal(wf_f[c][0], wf_l[c][0])
        assert np.array_equal(wf_f[c][1], wf_l[c][1])
    rows_s, wf_s, st_s, g_s = out[2]
    print('streamed weights: network rows %d, steps from the guess %s' % (rows_s, g_s), flush=True)
    assert rows_s == rows_f and g_s == g_f
    for c in range(len(seeds)):
        assert st_s[c]['iterations'] == st_f[c]['iterations']
        assert np.array_equal(wf_s[c][0], wf_f[c][0])
        assert np.array_equal(wf_s[c][1], wf_f[c][1])
