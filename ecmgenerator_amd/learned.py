"""Learned ORCA proposal (include/ecm_mlp.h): the network handle, sample
collection from simulator runs, GPU training and weight files.

The network predicts which ORCA constraints are active at the LP optimum;
the step verifies every proposal exactly (restricted LP + containment of
all other constraints) and falls back to the plain 2D / 3D LP
(ORCA.cpp:428-669) when the check fails, so the plain LP has the last word.

    python -m ecmgenerator_amd.learned train --out weights.npz

collects samples from the c1 and c2 workloads, trains on the GPU with the
library's own BF16 MFMA kernels and writes the parameters (numpy .npz,
arrays only).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import _native as N

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_WEIGHTS = os.path.join(HERE, "weights", "orca_proposal.npz")

N_IN, N_IN_PAD, N_OUT, N_SLOTS = 61, 64, 16, 14
POS_SCALE, VEL_SCALE, RAD_SCALE, FEAT_CLAMP = 0.25, 0.5, 4.0, 8.0


class ProposalNet:
    """An ecm_mlp handle: parameters, dataset and training on the GPU."""

    def __init__(self, seed: int = 0, params: np.ndarray | None = None):
        self._L = N.lib()
        h = C.c_void_p()
        N.check(self._L.ecm_mlp_create(C.c_uint64(seed), C.byref(h)))
        self._h = h
        self.n_params = int(self._L.ecm_mlp_n_params())
        if params is not None:
            self.set_params(params)

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            self._L.ecm_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # parameters
    def get_params(self) -> np.ndarray:
        p = np.empty(self.n_params, np.float32)
        N.check(self._L.ecm_mlp_get_params(self._h, N.fptr(p), self.n_params))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        if p.shape != (self.n_params,):
            raise ValueError(f"expected {self.n_params} parameters, got {p.shape}")
        N.check(self._L.ecm_mlp_set_params(self._h, N.fptr(p), self.n_params))

    def save(self, path: str, meta: dict | None = None) -> None:
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        np.savez(path, params=self.get_params(), meta=np.array(json.dumps(meta or {})))

    @classmethod
    def load(cls, path: str = DEFAULT_WEIGHTS) -> "ProposalNet":
        with np.load(path, allow_pickle=False) as z:
            p = z["params"]
        return cls(params=p)

    # dataset
    def reserve(self, capacity: int) -> None:
        N.check(self._L.ecm_mlp_dataset_reserve(self._h, C.c_int64(capacity)))

    def dataset_size(self) -> int:
        n = C.c_int64()
        N.check(self._L.ecm_mlp_dataset_size(self._h, C.byref(n)))
        return n.value

    def clear(self) -> None:
        N.check(self._L.ecm_mlp_dataset_clear(self._h))

    def read(self, first: int, n: int):
        x = np.empty((n, N_IN_PAD), np.float32)
        lab = np.empty(n, np.uint32)
        v = np.empty((n, 2), np.float32)
        N.check(self._L.ecm_mlp_dataset_read(self._h, C.c_int64(first), n, N.fptr(x),
                                             lab.ctypes.data_as(C.POINTER(C.c_uint32)), N.fptr(v)))
        return x, lab, v

    # training / evaluation
    def train(self, epochs: int, batch: int = 8192, lr: float = 1e-3, pos_weight: float = 2.0,
              vel_weight: float = 1.0, seed: int = 0) -> np.ndarray:
        losses = np.zeros(max(epochs, 1), np.float32)
        N.check(self._L.ecm_mlp_train(self._h, epochs, batch, C.c_float(lr), C.c_float(pos_weight),
                                      C.c_float(vel_weight), C.c_uint64(seed), N.fptr(losses)))
        return losses[:epochs]

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        assert x.ndim == 2 and x.shape[1] == N_IN_PAD
        y = np.empty((x.shape[0], N_OUT), np.float32)
        N.check(self._L.ecm_mlp_forward(self._h, N.fptr(x), x.shape[0], N.fptr(y)))
        return y

    def forward_device(self, x_ptr: int, n: int, y_ptr: int) -> None:
        """Forward pass on device buffers without any allocation
        (ecm_mlp_forward_noalloc): x_ptr holds n*64 FP32 features, y_ptr
        receives n*16 raw outputs, both device addresses on this network's
        GPU."""
        N.check(self._L.ecm_mlp_forward_noalloc(self._h, C.c_void_p(x_ptr), int(n), C.c_void_p(y_ptr)))

    def loss_grad(self, x, label, v, pos_weight=2.0, vel_weight=1.0):
        x = np.ascontiguousarray(x, np.float32)
        label = np.ascontiguousarray(label, np.uint32)
        v = np.ascontiguousarray(v, np.float32)
        loss = C.c_float()
        g = np.empty(self.n_params, np.float32)
        N.check(self._L.ecm_mlp_loss_grad(self._h, N.fptr(x), label.ctypes.data_as(C.POINTER(C.c_uint32)),
                                          N.fptr(v), x.shape[0], C.c_float(pos_weight), C.c_float(vel_weight),
                                          C.byref(loss), N.fptr(g)))
        return float(loss.value), g


def collect(net: ProposalNet, workload: str, steps: int, every: int, seed: int = 1, warmup: int = 0,
            log=print) -> int:
    """Run `workload` for `steps` steps and log a sample set every `every`
    steps (after `warmup` plain steps).  Returns the samples appended."""
    from . import workloads as W
    from .simulator import Simulator
    wl = W.build(workload, seed=seed)
    sim = Simulator(wl.ecm, wl.obstacles, wl.box, wl.params, planner=wl.scene)
    sim.SpawnAgents(wl.agents)
    added = 0
    for _ in range(warmup):
        sim.Update()
    for i in range(steps):
        if i % every == 0:
            added += sim.log_step(net)
        else:
            sim.Update()
    sim.close()
    log(f"[collect] {workload}: {added} samples ({net.dataset_size()} total)")
    return added


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--out", default=DEFAULT_WEIGHTS)
    t.add_argument("--samples", type=int, default=20_000_000)
    t.add_argument("--epochs", type=int, default=4)
    t.add_argument("--batch", type=int, default=8192)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--pos-weight", type=float, default=2.0)
    t.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    t0 = time.time()
    net = ProposalNet(seed=a.seed)
    net.reserve(a.samples)
    # c1: 100k agents, a sample set every 6 steps; c2: 1M agents every 30
    collect(net, "c1", steps=600, every=6, seed=11)
    collect(net, "c2", steps=300, every=30, seed=12)
    n = net.dataset_size()
    print(f"[train] {n} samples collected in {time.time() - t0:.1f}s", flush=True)
    t1 = time.time()
    losses = net.train(a.epochs, a.batch, a.lr, a.pos_weight, seed=a.seed)
    print(f"[train] epochs {a.epochs} losses {losses.tolist()} in {time.time() - t1:.1f}s", flush=True)
    net.clear()
    sweep = pick_theta(net, "c2", seed=13)
    theta = min(sweep, key=lambda r: r["ms_per_step"])["theta"]
    print(f"[train] theta sweep {sweep}; chose {theta}", flush=True)
    net.save(a.out, meta={"samples": n, "epochs": a.epochs, "batch": a.batch, "lr": a.lr,
                          "pos_weight": a.pos_weight, "seed": a.seed, "losses": losses.tolist(),
                          "theta": theta, "theta_sweep": sweep,
                          "sources": {"c1": {"seed": 11, "steps": 600, "every": 6},
                                      "c2": {"seed": 12, "steps": 300, "every": 30}}})
    print(f"[train] wrote {a.out}", flush=True)


def pick_theta(net: ProposalNet, workload: str, seed: int, thetas=(-3.0, -2.0, -1.5, -1.0, -0.5, 0.0),
               warmup: int = 20, steps: int = 10) -> list:
    """Step time and acceptance of the learned step for each threshold, on a
    held-out run of `workload` (seed not used for training)."""
    from . import workloads as W
    from .simulator import Simulator
    wl = W.build(workload, seed=seed)
    sim = Simulator(wl.ecm, wl.obstacles, wl.box, wl.params, planner=wl.scene)
    sim.SpawnAgents(wl.agents)
    for _ in range(warmup):
        sim.Update()
    out = []
    for th in thetas:
        sim.set_learned(net, th)
        sim.Update()
        sim.sync()
        s0 = sim.stats()
        t0 = time.perf_counter()
        for _ in range(steps):
            sim.Update()
        sim.sync()
        t1 = time.perf_counter()
        s1 = sim.stats()
        tried = s1["learned_tried"] - s0["learned_tried"]
        acc = s1["learned_accepted"] - s0["learned_accepted"]
        out.append({"theta": th, "ms_per_step": round(1000 * (t1 - t0) / steps, 4),
                    "acceptance": round(acc / max(tried, 1), 4)})
    sim.set_learned(None)
    sim.close()
    return out


if __name__ == "__main__":
    main()
