"""The learned troubled-face indicator: training data from the plain solver,
training on the GPU, threshold calibration, and the hybrid form's error.

All device work runs in libcbns (cbns_face_samples, cbns_mlp_*,
cbns_gather_rows); numpy only chooses indices and keeps host copies of the
targets.  The reference has no learned form; the label it is trained on is
the reference's own WENO5-JS nonlinear weights (src/weno.cu:13-44) on the
Roe-averaged characteristic fields (src/calc_eigenvectors_roe.cu).

  collect    seeded shock-vortex runs of the plain FP64 solver; at snapshot
             steps every face's features and label max |omega_k - d_k|; all
             troubled faces (label > tau) and a seeded sample of the smooth
             ones, capped per snapshot and axis
  train      weighted logistic loss, Adam, minibatches cut from a
             device-side shuffled copy of the data
  calibrate  the logit threshold reaching a target recall of troubled faces
  compare    a case run with the plain FP64 form and with another form from
             the same initial state: L2/Linf of rho and E, troubled fraction

Usage on a GPU machine:
  python -m cbns3d_cuda_amd.learned --out indicator.npz   (train and save)
"""
import argparse
import json
import time

import numpy as np

from . import cases
from .block import (BUF_Q, F_Q, FORM_FP32, FORM_HYBRID, MLP_NPARAM, Block, CbnsError,
                    DeviceArray, Mlp, gather_rows)

ERR_NAN = 3  # CBNS_ERR_NAN

ROW = 32


def init_params(seed):
    """He-normal hidden layers, small output layer, zero biases."""
    rng = np.random.Generator(np.random.PCG64(seed))
    P = np.zeros(MLP_NPARAM, dtype=np.float32)
    W1 = rng.standard_normal((32, 32)) * np.sqrt(2.0 / 30.0)
    W1[:, 30:] = 0.0
    P[0:1024] = W1.ravel()
    P[1056:2080] = (rng.standard_normal((32, 32)) * np.sqrt(2.0 / 32.0)).ravel()
    P[2112:2144] = rng.standard_normal(32) * 0.1
    return P


def start_block(case):
    b = Block(case.params)
    b.calc_metrics(case.x, case.y, case.z, case.period)
    b.set_conservative(cases.q_ref_aos(case))
    b.set_bc_calc_primitive(BUF_Q)
    return b


class Dataset:
    """Features on the device, targets and labels on the host."""

    def __init__(self, capacity):
        self.capacity = int(capacity)
        self.X = DeviceArray((self.capacity, ROW))
        self.n = 0
        self.labels = np.zeros(self.capacity, dtype=np.float32)

    def targets(self, tau):
        return (self.labels[:self.n] > tau).astype(np.float32)

    def add(self, d_feat, labels, idx):
        k = min(len(idx), self.capacity - self.n)
        if k <= 0:
            return 0
        idx = np.ascontiguousarray(idx[:k], dtype=np.int64)
        d_idx = DeviceArray.from_host(idx)
        gather_rows(d_feat.ptr, d_idx.ptr, k, ROW, self.X.at(self.n * ROW))
        self.labels[self.n:self.n + k] = labels[idx]
        self.n += k
        return k


def collect(seeds, size=(128, 64, 64), steps=200, snap_every=40, tau=0.1, cap=10000,
            capacity=8_000_000, log=print):
    """Samples from plain FP64 shock-vortex runs, one per seed."""
    ds = Dataset(capacity)
    feat = lab = None
    for seed in seeds:
        case = cases.shock_vortex(*size, seed=seed)
        b = start_block(case)
        rng = np.random.Generator(np.random.PCG64(10_000 + seed))
        t = 0.0
        try:
            for step in range(steps + 1):
                if step % snap_every == 0 and step > 0:
                    for axis in range(3):
                        n = b.face_count(axis)
                        if feat is None or feat.shape[0] < n:
                            feat, lab = DeviceArray((n, ROW)), DeviceArray((n,))
                        b.face_samples(axis, feat.ptr, lab.ptr)
                        labels = lab.download(n)
                        pos = np.flatnonzero(labels > tau)
                        neg = np.flatnonzero(labels <= tau)
                        pos = rng.permutation(pos)[:cap]
                        neg = rng.permutation(neg)[:max(cap, len(pos))]
                        ds.add(feat, labels, np.concatenate([pos, neg]))
                if step < steps:
                    t, _ = b.rk3_step(1e9, t)
        finally:
            b.free()
        log("collect seed %d: t=%.3f samples %d" % (seed, t, ds.n))
    return ds


def train(ds, tau, *, epochs=30, batch=65536, lr=3e-3, w_pos=2.0, seed=0, log=print,
          params=None):
    """Adam on the GPU; returns the trained parameters."""
    n = ds.n
    y = ds.targets(tau)
    m = Mlp(init_params(seed) if params is None else params)
    rng = np.random.Generator(np.random.PCG64(seed))
    Xs = DeviceArray((n, ROW))
    Ys = DeviceArray((n,))
    steps_per_epoch = max(1, n // batch)
    total = epochs * steps_per_epoch
    it = 0
    for ep in range(epochs):
        perm = rng.permutation(n).astype(np.int64)
        d_perm = DeviceArray.from_host(perm)
        gather_rows(ds.X.ptr, d_perm.ptr, n, ROW, Xs.ptr)
        Ys.upload(y[perm])
        d_perm.free()
        losses = []
        for s in range(steps_per_epoch):
            lo = s * batch
            k = min(batch, n - lo)
            rate = lr * 0.5 * (1.0 + np.cos(np.pi * it / total))  # cosine decay
            losses.append(m.loss_grad(Xs.at(lo * ROW), Ys.at(lo), k, w_pos=w_pos))
            m.adam_step(max(rate, 1e-5))
            it += 1
        log("epoch %d loss %.5f" % (ep, float(np.mean(losses))))
    return m.get_params()


def logits(params, d_X, n):
    m = Mlp(params)
    out = DeviceArray((max(n, 1),))
    m.forward(d_X, n, out.ptr)
    return out.download(n)


def calibrate(s, target, recall=0.995):
    """Largest logit threshold whose recall of troubled faces is >= recall;
    with the precision and troubled fraction it gives."""
    pos = np.sort(s[target > 0.5])
    if len(pos) == 0:
        return float(np.max(s)), {}
    thr = float(pos[int(np.floor((1.0 - recall) * len(pos)))]) - 1e-6
    return thr, stats_at(s, target, thr)


def stats_at(s, target, thr):
    pred = s > thr
    pos = target > 0.5
    tp = float(np.sum(pred & pos))
    return {"recall": tp / max(float(pos.sum()), 1.0),
            "precision": tp / max(float(pred.sum()), 1.0),
            "flagged_fraction": float(pred.mean()), "troubled_fraction": float(pos.mean())}


def compare(case, steps, form, params=None, thresh=0.0, per_stage=False, log=print, reuse=False,
            region=None):
    """Run `steps` RK3 steps with the plain FP64 form and with `form` from the
    same state; L2/Linf of rho and E (updated nodes, or the node box
    region = ((k0, k1), (j0, j1), (i0, i1)) of them) and troubled counts.  If
    `form` produces a non-physical state (the time step is NaN) its errors
    are inf and "diverged_step" is the step that failed."""
    out = {}
    qs = {}
    for name, f in (("fp64", 0), ("form", form)):
        b = start_block(case)
        try:
            if f & FORM_HYBRID:
                b.set_indicator(params, thresh, per_stage=per_stage, count=True, reuse=reuse)
            b.set_form(f)
            t = 0.0
            flagged = []
            for n in range(steps):
                try:
                    t, _ = b.rk3_step(1e9, t)
                except CbnsError as e:
                    if name == "fp64" or e.code != ERR_NAN:
                        raise
                    out["diverged_step"] = n + 1
                    break
                if f & FORM_HYBRID:
                    flagged.append(sum(b.flag_count()))
            qs[name] = b.get_field(F_Q, 5)
            out["t_" + name] = t
            if flagged:
                faces = sum(b.face_count(a) for a in range(3))
                out["troubled_fraction_mean"] = float(np.mean(flagged)) / faces
                out["troubled_fraction_max"] = float(np.max(flagged)) / faces
        finally:
            b.free()
    p = case.params
    G = 4
    sl = [slice(None)]
    for r, (a, n) in enumerate(((2, p.KM), (1, p.JM), (0, p.IM))):
        per = p.bc[a] == 1
        lo, hi = (0, n - 1) if per else (1, n - 2)
        if region is not None:
            lo, hi = max(lo, region[r][0]), min(hi, region[r][1] - 1)
        sl.append(slice(G + lo, G + hi + 1))
    a_, b_ = qs["form"][tuple(sl)], qs["fp64"][tuple(sl)]
    for c, nm in ((0, "rho"), (4, "E")):
        d = a_[c] - b_[c]
        diverged = "diverged_step" in out
        out["L2_" + nm] = float("inf") if diverged else float(np.sqrt(np.mean(d * d)))
        out["Linf_" + nm] = float("inf") if diverged else float(np.abs(d).max())
    log("compare %s form %d: %s" % (case.name, form, json.dumps(out)))
    return out


def save(path, params, thresh, meta):
    np.savez(path, params=np.asarray(params, dtype=np.float32), threshold=np.float32(thresh),
             meta=np.array(json.dumps(meta)))


def load(path):
    with np.load(path, allow_pickle=False) as z:
        return z["params"].astype(np.float32), float(z["threshold"]), json.loads(str(z["meta"]))


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--out", required=True)
    ap.add_argument("--train-seeds", type=int, default=20)
    ap.add_argument("--size", default="128,64,64")
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--snap-every", type=int, default=40)
    ap.add_argument("--tau", type=float, default=0.1)
    ap.add_argument("--cap", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=65536)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--w-pos", type=float, default=2.0)
    ap.add_argument("--recall", type=float, default=0.995)
    a = ap.parse_args()
    size = tuple(int(v) for v in a.size.split(","))
    t0 = time.time()
    log = lambda s: print("[%7.1fs] %s" % (time.time() - t0, s), flush=True)
    ds = collect(range(a.train_seeds), size, a.steps, a.snap_every, a.tau, a.cap, log=log)
    held = collect([1000, 1001], size, a.steps, a.snap_every, a.tau, a.cap, log=log)
    params = train(ds, a.tau, epochs=a.epochs, batch=a.batch, lr=a.lr, w_pos=a.w_pos, log=log)
    s = logits(params, held.X.ptr, held.n)
    thr, stats = calibrate(s, held.targets(a.tau), a.recall)
    stats_tr = stats_at(logits(params, ds.X.ptr, ds.n), ds.targets(a.tau), thr)
    meta = {"tau": a.tau, "size": size, "train_seeds": a.train_seeds, "held_out_seeds": [1000, 1001],
            "steps": a.steps, "snap_every": a.snap_every, "cap": a.cap, "epochs": a.epochs,
            "batch": a.batch, "lr": a.lr, "w_pos": a.w_pos, "samples": ds.n,
            "recall_target": a.recall, "held_out": stats,
            "train_at_threshold": stats_tr}
    log("threshold %.4f held-out %s" % (thr, json.dumps(stats)))
    save(a.out, params, thr, meta)
    log("saved " + a.out)


if __name__ == "__main__":
    main()
