"""Learned per-stage schedule: an MLP proposes (gamma_i, t_i), the simulator scores it.

The heuristic of MultiQW.cpp:386-423 picks gamma_i and the stage times from
closed-form moments.  Here a 32 -> 128 -> 128 -> 10 GELU network (trained on
the GPU by libmqw's own kernels) maps instance features to per-stage offsets
(log gamma_i, log t_i) from that heuristic.  A proposal keeps the heuristic's
total evolution time (the stage times are rescaled to the same sum), keeps
the per-sample U(t_i, 2 t_i) randomisation, and is scored by the simulator
itself; when it scores below the heuristic the heuristic is kept.

Training data: for seeded SK instances, candidate offset vectors are drawn
around zero, every candidate is simulated (mqw_solve_batch with a schedule
override, sharing the Monte-Carlo draws so comparisons are paired) and the
best-scoring candidate of each instance is the regression target.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L
from .api import Context, generate_sk, schedule_heuristic

M_MAX = 5
OFFSET_CLIP = 1.5


class MLP:
    """Handle of an mqw_mlp (parameters and Adam state on the device)."""

    def __init__(self, ctx: Context, seed: int = 1):
        self._ctx = ctx
        self._lib = L.load()
        h = C.c_void_p()
        ctx._check(self._lib.mqw_mlp_create(ctx._h, seed, C.byref(h)))
        self._h = h
        self.step = 0

    def close(self):
        if getattr(self, "_h", None):
            self._lib.mqw_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def params(self) -> np.ndarray:
        out = np.zeros(L.MLP_PARAMS, np.float32)
        self._ctx._check(self._lib.mqw_mlp_get_params(self._h, L.f32p(out), out.size))
        return out

    def set_params(self, p) -> None:
        p = np.ascontiguousarray(p, np.float32)
        self._ctx._check(self._lib.mqw_mlp_set_params(self._h, L.f32p(p), p.size))
        self.step = 0

    def forward(self, X) -> np.ndarray:
        X = np.ascontiguousarray(X, np.float32)
        Y = np.zeros((X.shape[0], L.MLP_OUT), np.float32)
        self._ctx._check(self._lib.mqw_mlp_forward(self._h, L.f32p(X), X.shape[0], L.f32p(Y)))
        return Y

    def loss_grad(self, X, T, Mk, want_grads: bool = False):
        X = np.ascontiguousarray(X, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        Mk = np.ascontiguousarray(Mk, np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(L.MLP_PARAMS, np.float32) if want_grads else None
        self._ctx._check(self._lib.mqw_mlp_loss_grad(self._h, L.f32p(X), L.f32p(T), L.f32p(Mk), X.shape[0],
                                                     L.f32p(loss), L.f32p(g) if g is not None else None))
        return (float(loss[0]), g) if want_grads else float(loss[0])

    def adam(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        self.step += 1
        self._ctx._check(self._lib.mqw_mlp_adam_step(self._h, lr, beta1, beta2, eps, self.step))


def features(ctx: Context, instances: Sequence[Tuple[int, int, np.ndarray]]) -> np.ndarray:
    """mqw_features: 32 per instance (n, m, coupling/field moments, 24-bin diagonal histogram)."""
    arr = (L.Instance * max(1, len(instances)))()
    keep = []
    for k, (n, m, rec) in enumerate(instances):
        r = np.ascontiguousarray(rec, np.float64)
        keep.append(r)
        arr[k].n, arr[k].m, arr[k].rec = n, m, L.f64p(r)
    out = np.zeros((len(instances), L.FEATURES), np.float32)
    ctx._check(L.load().mqw_features(ctx._h, arr, len(instances), L.f32p(out)))
    return out


def loss_mask(ms: Sequence[int]) -> np.ndarray:
    Mk = np.zeros((len(ms), L.MLP_OUT), np.float32)
    for k, m in enumerate(ms):
        Mk[k, : 2 * m] = 1.0
    return Mk


def apply_offsets(gamma_h, t_h, offsets) -> Tuple[np.ndarray, np.ndarray]:
    """Per-stage (log gamma, log t) offsets -> schedule with the heuristic's total time."""
    m = len(gamma_h)
    d = np.clip(np.asarray(offsets, np.float64)[: 2 * m].reshape(m, 2), -OFFSET_CLIP, OFFSET_CLIP)
    g = np.asarray(gamma_h, np.float64) * np.exp(d[:, 0])
    t = np.asarray(t_h, np.float64) * np.exp(d[:, 1])
    t *= float(np.sum(np.asarray(t_h, np.float64))) / float(np.sum(t))
    t32 = t.astype(np.float32)
    # The simulator's total time is the sequential float32 sum of the stage
    # times (as the heuristic's total_t, MultiQW.cpp:420-423).  Rounding to
    # float32 can leave that sum an ulp or two away from the heuristic's, so
    # one stage, the largest that can, absorbs the difference and the totals
    # are equal.  Ties in the last addition can skip a total when one stage
    # moves, so the stages are tried in turn, by up to 8 ulps either way.
    want = _f32_total(t_h)
    if _f32_total(t32) != want:
        for i in np.argsort(-t32):
            base = t32[i]
            for k in range(1, 17):  # +1, -1, +2, -2, ... ulps of stage i
                step = (k + 1) // 2 * (1 if k % 2 else -1)
                v = base
                for _ in range(abs(step)):
                    v = np.nextafter(v, np.float32(np.inf if step > 0 else 0.0), dtype=np.float32)
                t32[i] = v
                if _f32_total(t32) == want:
                    return g.astype(np.float32), t32
            t32[i] = base
    return g.astype(np.float32), t32


def _f32_total(t) -> np.float32:
    s = np.float32(0.0)
    for x in np.asarray(t, np.float32):
        s = np.float32(s + x)
    return s


def score(ctx: Context, instances, schedules, samples: int, with_total_t: bool = False):
    """Mean ground-state probability of each (instance, schedule) on the simulator
    (and, with with_total_t, the total evolution time the simulator ran)."""
    batch = [(n, m, rec, g, t) for (n, m, rec), (g, t) in zip(instances, schedules)]
    p, tt = ctx.solve_batch(batch, samples=samples)
    return (p, tt) if with_total_t else p


@dataclass
class Dataset:
    X: np.ndarray       # features (count x 32)
    T: np.ndarray       # targets: best offsets (count x 10)
    Mk: np.ndarray      # loss mask (count x 10)
    p_heur: np.ndarray  # heuristic score
    p_best: np.ndarray  # best candidate score
    scored: int         # schedules simulated


def make_dataset(ctx: Context, count: int, seed: int, n_range=(12, 18), m: int = 3, candidates: int = 10,
                 sigma: float = 0.35, samples: int = 20, chunk: int = 256, log=None) -> Dataset:
    rng = np.random.default_rng(seed)
    X, T, P0, PB = [], [], [], []
    scored = 0
    for c0 in range(0, count, chunk):
        cnt = min(chunk, count - c0)
        inst = []
        for k in range(cnt):
            n = int(rng.integers(n_range[0], n_range[1] + 1))
            inst.append((n, m, generate_sk(n, 1, seed * 1000003 + c0 + k)[0]))
        X.append(features(ctx, inst))
        offs = rng.normal(0.0, sigma, (cnt, candidates, L.MLP_OUT)).astype(np.float32)
        offs[:, 0, :] = 0.0  # candidate 0 is the heuristic
        offs[:, :, 2 * m:] = 0.0
        heur = [schedule_heuristic(n, mm, rec) for (n, mm, rec) in inst]
        all_inst, all_sched = [], []
        for k in range(cnt):
            for c in range(candidates):
                all_inst.append(inst[k])
                all_sched.append(apply_offsets(heur[k].gamma, heur[k].short_t, offs[k, c]))
        p = score(ctx, all_inst, all_sched, samples).reshape(cnt, candidates)
        scored += cnt * candidates
        best = np.argmax(p, axis=1)
        T.append(offs[np.arange(cnt), best])
        P0.append(p[:, 0])
        PB.append(p[np.arange(cnt), best])
        if log:
            log(dict(stage="dataset", done=c0 + cnt, of=count, scored=scored))
    X = np.concatenate(X)
    return Dataset(X, np.concatenate(T), loss_mask([m] * X.shape[0]), np.concatenate(P0), np.concatenate(PB), scored)


def train(mlp: MLP, data: Dataset, epochs: int = 200, batch: int = 256, lr: float = 1e-3, seed: int = 0,
          val_frac: float = 0.1) -> dict:
    rng = np.random.default_rng(seed)
    N = data.X.shape[0]
    idx = rng.permutation(N)
    nv = max(1, int(N * val_frac))
    val, tr = idx[:nv], idx[nv:]
    hist = []
    for ep in range(epochs):
        order = rng.permutation(tr)
        for b0 in range(0, len(order), batch):
            bi = order[b0:b0 + batch]
            mlp.loss_grad(data.X[bi], data.T[bi], data.Mk[bi])
            mlp.adam(lr=lr)
        if ep % 10 == 0 or ep == epochs - 1:
            hist.append((ep, mlp.loss_grad(data.X[tr], data.T[tr], data.Mk[tr]),
                         mlp.loss_grad(data.X[val], data.T[val], data.Mk[val])))
    return dict(history=hist, train=len(tr), val=len(val))


def evaluate(ctx: Context, mlp: MLP, count: int, seed: int, n: int = 20, m: int = 3, samples: int = 100,
             chunk: int = 64, log=None) -> dict:
    """Mean ground-state probability: heuristic vs MLP proposal vs proposal-with-fallback."""
    ph, pm = [], []
    equal_time = True
    t0 = time.time()
    for c0 in range(0, count, chunk):
        cnt = min(chunk, count - c0)
        inst = [(n, m, generate_sk(n, 1, seed * 1000003 + c0 + k)[0]) for k in range(cnt)]
        heur = [schedule_heuristic(nn, mm, rec) for (nn, mm, rec) in inst]
        Y = mlp.forward(features(ctx, inst))
        prop = [apply_offsets(h.gamma, h.short_t, Y[k]) for k, h in enumerate(heur)]
        p_h, tt_h = score(ctx, inst, [(h.gamma, h.short_t) for h in heur], samples, with_total_t=True)
        p_m, tt_m = score(ctx, inst, prop, samples, with_total_t=True)
        equal_time &= bool(np.array_equal(tt_h, tt_m)) and bool(
            np.array_equal(tt_h, np.array([h.total_t for h in heur], np.float32)))
        ph.append(p_h)
        pm.append(p_m)
        if log:
            log(dict(stage="evaluate", done=c0 + cnt, of=count, p_heuristic=float(np.concatenate(ph).mean()),
                     p_mlp=float(np.concatenate(pm).mean())))
    ph, pm = np.concatenate(ph), np.concatenate(pm)
    pf = np.maximum(ph, pm)
    return dict(count=count, n=n, m=m, samples=samples, p_heuristic=float(ph.mean()), p_mlp=float(pm.mean()),
                p_fallback=float(pf.mean()), gain_fallback=float(pf.mean() - ph.mean()),
                gain_raw=float(pm.mean() - ph.mean()), win_rate=float(np.mean(pm > ph)), equal_time=equal_time,
                seconds=time.time() - t0)


def main(argv=None) -> int:
    """Train on simulator-scored schedules and evaluate against the heuristic.

    python -m multistageqw_amd.learned --train 2000 --candidates 10 --eval 2048 --weights w.bin
    prints one JSON line per progress step and a final summary line.
    """
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=2000, help="training instances (n uniform in n-min..n-max)")
    ap.add_argument("--candidates", type=int, default=10, help="scored schedules per training instance")
    ap.add_argument("--n-min", type=int, default=12)
    ap.add_argument("--n-max", type=int, default=18)
    ap.add_argument("--m", type=int, default=3)
    ap.add_argument("--train-samples", type=int, default=20)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--eval", type=int, default=2048, help="fresh evaluation instances")
    ap.add_argument("--eval-n", type=int, default=20)
    ap.add_argument("--eval-samples", type=int, default=100)
    ap.add_argument("--weights", default=None, help="save (and with --load, read) the parameters here")
    ap.add_argument("--load", action="store_true")
    a = ap.parse_args(argv)

    def log(d):
        print(json.dumps(d), flush=True)

    ctx = Context(0)
    mlp = MLP(ctx, seed=1)
    summary = dict(m=a.m)
    if a.load and a.weights:
        mlp.set_params(np.fromfile(a.weights, "<f4"))
    else:
        t0 = time.time()
        data = make_dataset(ctx, a.train, seed=1, n_range=(a.n_min, a.n_max), m=a.m, candidates=a.candidates,
                            samples=a.train_samples, log=log)
        t1 = time.time()
        info = train(mlp, data, epochs=a.epochs, batch=256, lr=1e-3)
        summary.update(scored_schedules=data.scored, dataset_s=t1 - t0, train_s=time.time() - t1,
                       loss_history=info["history"][-3:], train_p_heuristic=float(data.p_heur.mean()),
                       train_p_best=float(data.p_best.mean()))
        if a.weights:
            mlp.params().astype("<f4").tofile(a.weights)
    if a.eval > 0:
        summary["evaluation"] = evaluate(ctx, mlp, a.eval, seed=2, n=a.eval_n, m=a.m, samples=a.eval_samples,
                                         log=log)
    log(dict(stage="summary", **summary))
    mlp.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    import sys
    sys.exit(main())
