"""Training side of the learned Newton initial guess.

Feature rows are built from waveforms the engine itself produced (every
accepted point of every node), with exactly the formula of the engine's
k_guess_features kernel: for node u and the step t_n -> t_{n+1} = t_n + h,
from the accepted values v0..v3 at t_n > t_{n-1} > t_{n-2} > t_{n-3} and the
node's activity a = |v0 - v1| + |v1 - v2| + |v2 - v3|:

  [v0, (v0 - v1)/a, (v1 - v2)/a, (v2 - v3)/a, r0, r1, r2,
   nb_1 - v0 .. nb_8 - v0, log10(a)/5]

with r0 = h / (t_n - t_{n-1}), r1 = (t_n - t_{n-1}) / (t_{n-1} - t_{n-2}),
r2 = (t_{n-1} - t_{n-2}) / (t_{n-2} - t_{n-3}), each clamped to [0, 4], and
nb_k the accepted voltage of the k-th neighbour node (0 where absent).
Only nodes the network can matter for get a row: a >= 10 uV, and the
largest correction the network may make, 4 a, above the node's Newton
tolerance reltol |v0| + vntol (niconv.c:50-60); the engine gives the
other nodes the linear extrapolation.  The target is the linear extrapolation's error in
units of a, (v_{n+1} - (v0 + r0 (v0 - v1))) / a, clipped to [-4, 4] as the
engine clamps the network output.  In training h is the step actually taken
to t_{n+1}.
"""
import numpy as np

RATIO_MAX = 4.0
MIN_ACTIVITY = 1e-5
CLAMP = 4.0


def _ratio(a, b):
    with np.errstate(divide='ignore', invalid='ignore'):
        r = a / b
    r = np.where(np.isnan(r), 0.0, r)
    return np.clip(r, 0.0, RATIO_MAX)


def features(t, x, nbr, nnodes, every=1, with_scale=False, reltol=1e-3, vntol=1e-6):
    """t[P], x[P, N] (unknowns 1..N in columns 0..N-1), nbr[(nnodes+1)*8].
    Returns (F [rows, 16] float32, target [rows] float32) over steps
    n = 3 .. P-2 (every `every`-th) and the active nodes 1..nnodes; with
    with_scale also the rows' activity a (float64, volts)."""
    P = len(t)
    if P < 5:
        z = (np.zeros((0, 16), np.float32), np.zeros(0, np.float32))
        return z + (np.zeros(0),) if with_scale else z
    nb = np.asarray(nbr).reshape(-1, 8)[1:nnodes + 1]          # [NN, 8], node numbers
    steps = np.arange(3, P - 1, every)
    V = x[:, :nnodes]                                             # node voltages
    v0, v1, v2, v3 = V[steps], V[steps - 1], V[steps - 2], V[steps - 3]
    t0, t1, t2, t3 = t[steps], t[steps - 1], t[steps - 2], t[steps - 3]
    h = t[steps + 1] - t0
    r0 = _ratio(h, t0 - t1)[:, None]
    r1 = _ratio(t0 - t1, t1 - t2)[:, None]
    r2 = _ratio(t1 - t2, t2 - t3)[:, None]
    d0, d1, d2 = v0 - v1, v1 - v2, v2 - v3
    act = np.abs(d0) + np.abs(d1) + np.abs(d2)
    keep = (act >= MIN_ACTIVITY) & (CLAMP * act > reltol * np.abs(v0) + vntol)
    a = np.where(keep, act, 1.0)
    S, NN = v0.shape
    F = np.zeros((S, NN, 16))
    F[:, :, 0] = v0
    F[:, :, 1] = d0 / a
    F[:, :, 2] = d1 / a
    F[:, :, 3] = d2 / a
    F[:, :, 4] = r0
    F[:, :, 5] = r1
    F[:, :, 6] = r2
    for k in range(8):
        idx = nb[:, k]
        valid = idx >= 1
        vals = np.where(valid[None, :], V[steps][:, np.clip(idx - 1, 0, NN - 1)], 0.0)
        F[:, :, 7 + k] = np.where(valid[None, :], vals - v0, 0.0)
    F[:, :, 15] = np.log10(a) / 5.0
    lin = v0 + r0 * d0
    target = np.clip((V[steps + 1] - lin) / a, -CLAMP, CLAMP)
    out = (F[keep].astype(np.float32), target[keep].astype(np.float32))
    return out + (act[keep],) if with_scale else out
