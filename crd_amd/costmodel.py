"""The learned plan cost model: a 3-layer MLP (32 plan features -> 64 -> 64 -> 1, FP32) that predicts log(time) of a
plan, trained and evaluated in libcrdfft's own HIP kernels, used to rank autotuner candidates so only the top k are
timed. Exhaustive timing (crdfft_autotune_plan) stays the fallback and the measure the ranking is scored against.

The reference has nothing learned: its planner is a checked-in table measured offline
(engine/numerics/hesap-fft/include/crd/hesap/fft/detail/level_plans_gen.hpp, scripts/gen_fft_level_plans.py).

Run the experiment (on a GPU):  python -m crd_amd.costmodel --out report.json
"""
import argparse
import ctypes
import json
import time

import numpy as np

from . import _native as N
from ._native import PlanRecord, Tuning, check

FEATURES = N.FEATURES
HIDDEN = 64


class Mlp:
    """The cost model on the device (crdfft_mlp_*)."""

    def __init__(self, inputs=FEATURES, hidden=HIDDEN, seed=1):
        self.inputs, self.hidden = int(inputs), int(hidden)
        self.handle = ctypes.c_void_p()
        check(N.lib().crdfft_mlp_create(ctypes.byref(self.handle), self.inputs, self.hidden, seed),
              "crdfft_mlp_create")

    def param_count(self):
        c = ctypes.c_uint64()
        check(N.lib().crdfft_mlp_param_count(self.handle, ctypes.byref(c)), "crdfft_mlp_param_count")
        return c.value

    def params(self):
        p = np.empty(self.param_count(), dtype=np.float32)
        check(N.lib().crdfft_mlp_get_params(self.handle, _fp(p)), "crdfft_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.param_count()
        check(N.lib().crdfft_mlp_set_params(self.handle, _fp(p)), "crdfft_mlp_set_params")

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty(x.shape[0], dtype=np.float32)
        check(N.lib().crdfft_mlp_predict(self.handle, _fp(x), x.shape[0], _fp(y)), "crdfft_mlp_predict")
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(self.param_count(), dtype=np.float32)
        loss = ctypes.c_float()
        check(N.lib().crdfft_mlp_loss_grad(self.handle, _fp(x), _fp(t), x.shape[0], ctypes.byref(loss), _fp(g)),
              "crdfft_mlp_loss_grad")
        return loss.value, g

    def train(self, x, t, epochs=200, batch=256, lr=2e-3, seed=7):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = ctypes.c_float()
        check(N.lib().crdfft_mlp_train(self.handle, _fp(x), _fp(t), x.shape[0], epochs, batch, lr, seed,
                                       ctypes.byref(loss)), "crdfft_mlp_train")
        return loss.value

    def __del__(self):
        try:
            if self.handle:
                N.lib().crdfft_mlp_destroy(self.handle)
        except Exception:
            pass


def _fp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def features(records):
    """CRDFFT_FEATURES floats per record, computed by the library (crdfft_plan_features)."""
    arr = (PlanRecord * len(records))(*records)
    out = np.empty((len(records), FEATURES), dtype=np.float32)
    check(N.lib().crdfft_plan_features(arr, len(records), _fp(out)), "crdfft_plan_features")
    return out


def target(records):
    """log(milliseconds) of each record: what the network regresses."""
    return np.log(np.array([max(r.ms, 1e-6) for r in records], dtype=np.float64)).astype(np.float32)


_TYPES = {"c2c": N.C2C, "r2c": N.R2C, "c2r": N.C2R}


def candidates(n, batch, type="c2c"):
    cnt = ctypes.c_uint32()
    buf = (Tuning * 256)()
    check(N.lib().crdfft_plan_candidates(n, batch, _TYPES[type], buf, 256, ctypes.byref(cnt)),
          "crdfft_plan_candidates")
    return [buf[i] for i in range(min(cnt.value, 256))]


def time_case(case, reps=3):
    """Every candidate of one case, timed on the device: a list of PlanRecord (candidate order)."""
    recs = []
    if case["kind"] == "nd":
        dims = (ctypes.c_uint64 * len(case["dims"]))(*case["dims"])
        for flags in (0, 1):
            r = PlanRecord()
            check(N.lib().crdfft_time_nd(dims, len(case["dims"]), case["batch"], flags, reps, ctypes.byref(r)),
                  "crdfft_time_nd")
            recs.append(r)
        return recs
    for t in candidates(case["n"], case["batch"], case["type"]):
        r = PlanRecord()
        st = N.lib().crdfft_time_plan(case["n"], case["batch"], _TYPES[case["type"]], ctypes.byref(t), reps,
                                      ctypes.byref(r))
        if st == N.OUT_OF_MEMORY:
            continue
        check(st, "crdfft_time_plan")
        recs.append(r)
    return recs


def draw_cases(count, seed, lo=6, hi=22, nd_fraction=0.08):
    """Seeded size/batch mixes: n = 2^U{lo..hi}, n*batch log-uniform in [2^16, 2^25], C2C (80%) / R2C / C2R; a
    fraction of 3D C2C volumes with power-of-two sides 16..256 (at most 256^3)."""
    rng = np.random.default_rng(seed)
    cases = []
    while len(cases) < count:
        if rng.random() < nd_fraction:
            dims = [int(2 ** rng.integers(4, 9)) for _ in range(3)]
            if np.prod(dims) > 256 ** 3:
                continue
            batch = int(max(1, 2 ** rng.integers(0, 3)))
            if np.prod(dims) * batch > (1 << 25):
                batch = 1
            cases.append({"kind": "nd", "dims": dims, "batch": batch})
            continue
        lg = int(rng.integers(lo, hi + 1))
        n = 1 << lg
        total = 2.0 ** rng.uniform(16, 25)
        batch = int(max(1, round(total / n)))
        u = rng.random()
        typ = "c2c" if u < 0.8 else ("r2c" if u < 0.9 else "c2r")
        cases.append({"kind": "1d", "n": n, "batch": batch, "type": typ})
    return cases


def evaluate(model, timed_cases, k_values=(1, 2, 3), seed=0):
    """Regret of timing only the model's top k against the exhaustive best, beside the planner's default
    (candidate 0) and k candidates picked at random."""
    rng = np.random.default_rng(seed)
    out = {}
    regrets = {k: [] for k in k_values}
    random_regrets = {k: [] for k in k_values}
    default_regret = []
    timed_fraction = {k: [] for k in k_values}
    for recs in timed_cases:
        if len(recs) < 2:
            continue
        ms = np.array([r.ms for r in recs])
        best = ms.min()
        pred = model.predict(features(recs))
        order = np.argsort(pred, kind="stable")
        default_regret.append(ms[0] / best - 1.0)
        for k in k_values:
            kk = min(k, len(recs))
            regrets[k].append(ms[order[:kk]].min() / best - 1.0)
            draws = [ms[rng.choice(len(recs), kk, replace=False)].min() / best - 1.0 for _ in range(16)]
            random_regrets[k].append(float(np.mean(draws)))
            timed_fraction[k].append(kk / len(recs))
    for k in k_values:
        r = np.array(regrets[k])
        out[f"top{k}"] = {
            "mean_regret": float(r.mean()), "median_regret": float(np.median(r)),
            "p90_regret": float(np.quantile(r, 0.9)), "max_regret": float(r.max()),
            "found_best": float(np.mean(r == 0.0)),
            "random_mean_regret": float(np.mean(random_regrets[k])),
            "timed_fraction": float(np.mean(timed_fraction[k])),
        }
    out["default_mean_regret"] = float(np.mean(default_regret))
    out["cases"] = len(default_regret)
    return out


def main():
    ap = argparse.ArgumentParser(description="train and evaluate the plan cost model")
    ap.add_argument("--train-cases", type=int, default=1400)
    ap.add_argument("--eval-cases", type=int, default=2000)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--epochs", type=int, default=300)
    ap.add_argument("--out", default="costmodel_report.json")
    ap.add_argument("--save-records", default=None)
    args = ap.parse_args()

    t0 = time.time()
    train_cases = draw_cases(args.train_cases, seed=101)
    eval_cases = draw_cases(args.eval_cases, seed=202)
    train_timed = []
    for i, c in enumerate(train_cases):
        train_timed.append(time_case(c, args.reps))
        if i % 100 == 0:
            print(f"[costmodel] train case {i}/{len(train_cases)} ({time.time() - t0:.0f} s)", flush=True)
    eval_timed = []
    for i, c in enumerate(eval_cases):
        eval_timed.append(time_case(c, args.reps))
        if i % 100 == 0:
            print(f"[costmodel] eval case {i}/{len(eval_cases)} ({time.time() - t0:.0f} s)", flush=True)
    t_collect = time.time() - t0
    records = [r for recs in train_timed for r in recs]
    x, y = features(records), target(records)
    mu, sd = float(y.mean()), float(y.std() + 1e-6)
    model = Mlp(seed=1)
    t1 = time.time()
    loss = model.train(x, (y - mu) / sd, epochs=args.epochs, batch=256, lr=2e-3, seed=7)
    t_train = time.time() - t1
    report = {
        "train_records": len(records), "train_cases": len(train_cases), "eval_cases": len(eval_cases),
        "eval_records": sum(len(r) for r in eval_timed), "collect_seconds": t_collect, "train_seconds": t_train,
        "final_train_loss_normalized": loss,
        "exhaustive_timing_ms_eval": float(sum(r.ms for recs in eval_timed for r in recs)),
    }
    report.update(evaluate(model, eval_timed))
    # tuning time saved: the measured time of timing only the top k against timing every candidate
    for k in (1, 2, 3):
        spent = 0.0
        for recs in eval_timed:
            if not recs:
                continue
            pred = model.predict(features(recs))
            order = np.argsort(pred, kind="stable")[:k]
            spent += sum(recs[i].ms for i in order)
        report[f"top{k}"]["tuning_time_fraction"] = spent / max(report["exhaustive_timing_ms_eval"], 1e-9)
    if args.save_records:
        np.savez_compressed(args.save_records, x=x, y=y)
    with open(args.out, "w") as f:
        json.dump(report, f, indent=1)
    print(json.dumps(report), flush=True)


if __name__ == "__main__":
    main()
