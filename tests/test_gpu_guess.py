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


def test_learned_guess_end_to_end():
    print('case: learned guess, train seeds 4096..4103, eval seeds 0..3', flush=True)
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
    assert loss < 0.5 * base_loss

    eval_seeds = [0, 1, 2, 3]
    bp, dp = run(eval_seeds)
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
