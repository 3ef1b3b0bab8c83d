"""Learned ghost-track filter: features, training and inference on the GPU.

The network (10 -> 32 -> 32 -> 1, ReLU, ReLU, sigmoid; FP16 weights, FP32
accumulation) and its kernels are described in include/velo_ghost.h.  This
module drives them:

* ``track_features``: the ten features of every consolidated track of a
  sequence run (velo_track_features_launch);
* ``Trainer``: minibatch Adam on the GPU (velo_ghost_grad_launch +
  velo_ghost_adam_launch), FP32 master weights, FP16 model copy;
* ``choose_threshold``: the largest p(ghost) cut whose loss of
  reconstructible-track efficiency stays within a budget, measured with the
  checker on the training events;
* ``GhostModel.to_device``: the velo_ghost_model a sequence run uses to
  filter inline in consolidation (velo_sequence_input.dev_ghost_model).

Labels come from the checker: a track is a ghost when no generated particle
holds at least 70 % of its matched hits.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import native as nat

N_FEATURES = 10
HIDDEN = 32
N_PARAMS = 1441
FEATURE_NAMES = ("n_hits", "first_module", "last_module", "tx", "ty", "chi2_ndof", "mean_residual",
                 "max_residual", "missing_pairs", "seed_phi_spread")
# parameter blocks: W1[32][10], b1, W2[32][32], b2, w3, b3
OFFSETS = {"W1": 0, "b1": 320, "W2": 352, "b2": 1376, "w3": 1408, "b3": 1440}

P = C.c_void_p
U32 = C.c_uint32


class ModelStruct(C.Structure):
    _fields_ = [("mean", C.c_float * N_FEATURES), ("scale", C.c_float * N_FEATURES), ("threshold", C.c_float),
                ("reserved", U32), ("weights", C.c_uint16 * (N_PARAMS + 1))]


class TrackFeaturesArgs(C.Structure):
    _fields_ = [("dev_atomics_velo", P), ("dev_velo_track_hit_number", P), ("dev_velo_track_hits", P),
                ("dev_velo_states", P), ("dev_features", P), ("n_events", U32), ("tuning", nat.Tuning)]


class InferArgs(C.Structure):
    _fields_ = [("dev_features", P), ("dev_model", P), ("dev_score", P), ("n", U32), ("tuning", nat.Tuning)]


class GradArgs(C.Structure):
    _fields_ = [("dev_features", P), ("dev_labels", P), ("dev_batch", P), ("dev_model", P), ("dev_partials", P),
                ("dev_grad", P), ("batch_size", U32), ("pos_weight", C.c_float), ("tuning", nat.Tuning)]


class AdamArgs(C.Structure):
    _fields_ = [("dev_grad", P), ("dev_master", P), ("dev_m", P), ("dev_v", P), ("dev_model", P), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("step", U32)]


_bound = False


def _lib():
    global _bound
    L = nat.lib()
    if not _bound:
        for name, st in (("velo_track_features_launch", TrackFeaturesArgs), ("velo_ghost_infer_launch", InferArgs),
                         ("velo_ghost_grad_launch", GradArgs), ("velo_ghost_adam_launch", AdamArgs)):
            f = getattr(L, name)
            f.argtypes = [C.POINTER(st), P]
            f.restype = C.c_int
        L.velo_ghost_partials_size.argtypes = [U32]
        L.velo_ghost_partials_size.restype = C.c_size_t
        nat.check_mirrors(L.velo_abi_struct_size, {
            ModelStruct: "velo_ghost_model", TrackFeaturesArgs: "velo_track_features_args",
            InferArgs: "velo_ghost_infer_args", GradArgs: "velo_ghost_grad_args", AdamArgs: "velo_ghost_adam_args"},
            nat.LIB_PATH)
        _bound = True
    return L


def _call(name, args, stream=None):
    nat.check(getattr(_lib(), name)(C.byref(args), stream), name)


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised weights, zero biases (float32, the parameter order)."""
    rng = np.random.default_rng(seed)
    w = np.zeros(N_PARAMS, np.float32)
    w[0:320] = rng.normal(0, np.sqrt(2.0 / N_FEATURES), 320)
    w[352:1376] = rng.normal(0, np.sqrt(2.0 / HIDDEN), 1024)
    w[1408:1440] = rng.normal(0, np.sqrt(1.0 / HIDDEN), 32)
    return w


@dataclass
class GhostModel:
    params: np.ndarray                      # float32[1441] (FP32 master values)
    mean: np.ndarray = field(default_factory=lambda: np.zeros(N_FEATURES, np.float32))
    scale: np.ndarray = field(default_factory=lambda: np.ones(N_FEATURES, np.float32))
    threshold: float = 0.5

    @property
    def fp16(self) -> np.ndarray:
        return self.params.astype(np.float16)

    def struct(self) -> ModelStruct:
        m = ModelStruct()
        m.mean[:] = [float(v) for v in self.mean]
        m.scale[:] = [float(v) for v in self.scale]
        m.threshold = float(self.threshold)
        bits = np.zeros(N_PARAMS + 1, np.uint16)
        bits[:N_PARAMS] = self.fp16.view(np.uint16)
        C.memmove(m.weights, bits.ctypes.data, bits.nbytes)
        return m

    def to_device(self):
        """A device copy of the velo_ghost_model (a torch uint8 tensor)."""
        import torch
        m = self.struct()
        return torch.frombuffer(bytearray(C.string_at(C.addressof(m), C.sizeof(m))), dtype=torch.uint8).cuda()

    def save(self, path):
        np.savez(path, params=self.params, mean=self.mean, scale=self.scale, threshold=np.float32(self.threshold))

    @classmethod
    def load(cls, path):
        d = np.load(path)
        return cls(d["params"].astype(np.float32), d["mean"].astype(np.float32), d["scale"].astype(np.float32),
                   float(d["threshold"]))


def track_features(out, n_events: int, stream=None):
    """Features of the consolidated tracks of a sequence output (device
    pointers of velo_sequence_output); returns a torch float32 (T, 10)."""
    import torch
    T = int(out.total_tracks)
    feats = torch.zeros((max(T, 1), N_FEATURES), dtype=torch.float32, device="cuda")
    a = TrackFeaturesArgs(out.dev_atomics_velo, out.dev_track_hit_number, out.dev_track_hits, out.dev_states,
                          feats.data_ptr(), n_events, nat.Tuning(0, 0))
    _call("velo_track_features_launch", a, stream)
    torch.cuda.synchronize()
    return feats[:T]


def infer(model_dev, feats) -> "torch.Tensor":
    import torch
    n = feats.shape[0]
    score = torch.zeros(max(n, 1), dtype=torch.float32, device="cuda")
    _call("velo_ghost_infer_launch", InferArgs(feats.data_ptr(), model_dev.data_ptr(), score.data_ptr(), n,
                                               nat.Tuning(0, 0)))
    torch.cuda.synchronize()
    return score[:n]


class Trainer:
    """Minibatch Adam on the GPU over resident features and labels."""

    def __init__(self, feats, labels, model: GhostModel, pos_weight=1.0, max_batch=4096):
        import torch
        self.feats = feats.contiguous()
        self.labels = labels.to(torch.float32).contiguous()
        self.model = model
        self.pos_weight = float(pos_weight)
        self.master = torch.from_numpy(model.params.astype(np.float32)).cuda()
        self.m = torch.zeros_like(self.master)
        self.v = torch.zeros_like(self.master)
        self.grad = torch.zeros(N_PARAMS + 1, dtype=torch.float32, device="cuda")
        self.partials = torch.zeros(int(_lib().velo_ghost_partials_size(max_batch)), dtype=torch.float32,
                                    device="cuda")
        self.max_batch = max_batch
        self.dev_model = model.to_device()
        self.steps = 0

    def gradient(self, batch) -> np.ndarray:
        """Summed gradient and loss (index N_PARAMS) of one minibatch, with
        the current FP16 weights (no update)."""
        import torch
        b = int(batch.numel())
        if b > self.max_batch:
            raise ValueError("minibatch larger than max_batch")
        a = GradArgs(self.feats.data_ptr(), self.labels.data_ptr(), batch.data_ptr(), self.dev_model.data_ptr(),
                     self.partials.data_ptr(), self.grad.data_ptr(), b, self.pos_weight, nat.Tuning(0, 0))
        _call("velo_ghost_grad_launch", a)
        torch.cuda.synchronize()
        return self.grad.cpu().numpy()

    def step(self, batch, lr=3e-3, beta1=0.9, beta2=0.999, eps=1e-8) -> None:
        b = int(batch.numel())
        a = GradArgs(self.feats.data_ptr(), self.labels.data_ptr(), batch.data_ptr(), self.dev_model.data_ptr(),
                     self.partials.data_ptr(), self.grad.data_ptr(), b, self.pos_weight, nat.Tuning(0, 0))
        _call("velo_ghost_grad_launch", a)
        self.steps += 1
        _call("velo_ghost_adam_launch", AdamArgs(self.grad.data_ptr(), self.master.data_ptr(), self.m.data_ptr(),
                                                 self.v.data_ptr(), self.dev_model.data_ptr(), lr, beta1, beta2, eps,
                                                 self.steps))

    def train(self, epochs=20, batch_size=2048, lr=3e-3, seed=0, log=None):
        """Shuffled minibatches; returns the mean loss per epoch."""
        import torch
        n = self.feats.shape[0]
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n).astype(np.int32)).cuda()
            loss = 0.0
            k = 0
            for s in range(0, n - batch_size + 1, batch_size):
                self.step(perm[s: s + batch_size], lr=lr)
                loss += float(self.grad[N_PARAMS])  # the loss of the minibatch before the update
                k += 1
            losses.append(loss / max(k, 1))
            if log:
                log(f"[ghost] epoch {ep}: loss {losses[-1]:.5f}")
        torch.cuda.synchronize()
        self.model.params = self.master.cpu().numpy().astype(np.float32)
        return losses


def normalisation(feats: np.ndarray):
    mean = feats.mean(axis=0).astype(np.float32)
    std = feats.std(axis=0)
    scale = np.where(std > 0, 1.0 / np.maximum(std, 1e-12), 1.0).astype(np.float32)
    return mean, scale


def choose_threshold(scores: np.ndarray, efficiency_at, budget: float, lo=0.05, hi=0.999, iters=18):
    """Smallest threshold t in [lo, hi] whose efficiency loss
    efficiency_at(None) - efficiency_at(scores < t) stays within budget,
    found by bisection (efficiency grows with t)."""
    base = efficiency_at(None)
    if base - efficiency_at(scores < lo) <= budget:
        return lo
    for _ in range(iters):
        mid = 0.5 * (lo + hi)
        if base - efficiency_at(scores < mid) <= budget:
            hi = mid
        else:
            lo = mid
    return hi


def labelled_tracks(seq, ev, ghost_model=None):
    """Plain sequence run on ev's host events: (features (T, 10) on the GPU,
    checker labels (T,) as float32 with 1 = ghost, host tracks)."""
    from . import checker
    seq.run(ev.n_events, host_events=ev.events, host_offsets=ev.offsets)
    host = seq.host_tracks()
    feats = track_features(seq.output(), ev.n_events)
    labels = checker.check(ev, host).ghost_labels.astype(np.float32)
    return feats, labels, host


def train_and_evaluate(seq, train_ev, test_ev, epochs=20, batch_size=2048, lr=3e-3, budget=0.004, seed=0,
                       log=None):
    """Trains on train_ev's tracks, picks the threshold on train_ev within
    the efficiency budget, and reports ghost rate and efficiency on the
    held-out test_ev against the unfiltered output and against the best
    single cut on chi2/ndof at the same budget (the filter without the
    network).  Returns a report dict; the model is report["model"]."""
    import torch
    from . import checker
    ftr, ytr, htr = labelled_tracks(seq, train_ev)
    mean, scale = normalisation(ftr.cpu().numpy())
    model = GhostModel(init_params(seed), mean, scale)
    n_pos = float(ytr.sum())
    pos_weight = float(np.sqrt(max(len(ytr) - n_pos, 1.0) / max(n_pos, 1.0)))
    bs = min(batch_size, max(64, len(ytr) // 4))
    tr = Trainer(ftr, torch.from_numpy(ytr).cuda(), model, pos_weight=pos_weight, max_batch=bs)
    losses = tr.train(epochs, bs, lr, seed, log)
    s_train = infer(model.to_device(), ftr).cpu().numpy()
    model.threshold = choose_threshold(
        s_train, lambda acc: checker.check(train_ev, htr, accepted=acc).efficiency(), budget)

    fte, yte, hte = labelled_tracks(seq, test_ev)
    s_test = infer(model.to_device(), fte).cpu().numpy()
    plain = checker.check(test_ev, hte)
    filt = checker.check(test_ev, hte, accepted=s_test < model.threshold)

    # the filter without the network: one cut on chi2/ndof, chosen the same way
    chi_train = ftr[:, 5].cpu().numpy()
    chi_max = float(chi_train.max()) + 1.0
    cut = choose_threshold(-chi_train / chi_max, lambda acc: checker.check(train_ev, htr, accepted=acc).efficiency(),
                           budget, lo=-1.0, hi=0.0)
    chi_test = fte[:, 5].cpu().numpy()
    base = checker.check(test_ev, hte, accepted=(-chi_test / chi_max) < cut)

    def summary(r):
        return {"efficiency": r.efficiency(), "ghost_rate": r.ghost_rate, "tracks": r.tracks, "ghosts": r.ghosts}
    return {
        "model": model,
        "train": {"events": train_ev.n_events, "tracks": int(len(ytr)), "ghosts": int(n_pos), "loss": losses,
                  "pos_weight": pos_weight, "threshold": model.threshold, "batch_size": bs, "epochs": epochs},
        "test": {"events": test_ev.n_events, "tracks": int(len(yte)),
                 "plain": summary(plain), "filtered": summary(filt), "chi2_cut": summary(base),
                 "efficiency_loss": plain.efficiency() - filt.efficiency(),
                 "ghost_rate_plain": plain.ghost_rate, "ghost_rate_filtered": filt.ghost_rate,
                 "ghost_rate_chi2_cut": base.ghost_rate},
    }


def main():
    import argparse
    import json
    import sys
    import time
    from . import gen
    from .stream import VeloSequence
    ap = argparse.ArgumentParser(description="train the ghost filter on the GPU and evaluate it on held-out events")
    ap.add_argument("--train-events", type=int, default=20000)
    ap.add_argument("--test-events", type=int, default=10000)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--batch-size", type=int, default=2048)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--budget", type=float, default=0.004, help="allowed absolute efficiency loss (training events)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out-model", default="velo_ghost.npz")
    ap.add_argument("--out-report", default=None)
    a = ap.parse_args()
    say = lambda *x: print(*x, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.time()
    train_ev = gen.minbias(a.train_events, seed=1000 + a.seed)
    test_ev = gen.minbias(a.test_events, seed=2000 + a.seed)
    seq = VeloSequence(gen.geometry(), gen.module_zs(), max_events=max(a.train_events, a.test_events),
                       reserve_mb=16384)
    say(f"[ghost] events generated in {time.time() - t0:.1f} s")
    rep = train_and_evaluate(seq, train_ev, test_ev, a.epochs, a.batch_size, a.lr, a.budget, a.seed, log=say)
    rep["model"].save(a.out_model)
    out = {k: v for k, v in rep.items() if k != "model"}
    out["seconds"] = round(time.time() - t0, 1)
    out["train"]["loss"] = [round(x, 6) for x in out["train"]["loss"]]
    text = json.dumps(out)
    if a.out_report:
        with open(a.out_report, "w") as f:
            f.write(text + "\n")
    print(text)
    seq.close()


if __name__ == "__main__":
    main()
