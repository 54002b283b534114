"""Plans, the autotuner and its learned cost model (Python side).

A plan says how the transform of one size class (length 2^L) is cut into
passes: ``last`` levels in the last pass (``last == L``: one tile), the
widest register chunk ``bm`` (radix 2^bm) and whether the first forward
pass runs one prime per block (``split``). The library keeps one plan per
(L, square?, bucket of the class's job count); ``autotune`` measures
candidates on real jobs and stores the fastest.

The cost model is a small MLP (16 -> 32 -> 32 -> 1, ReLU) over plan features
that predicts a plan's log launch time relative to the other plans of the
same class (only the ranking within a class is used). ``train_gpu`` fits it to
measured plan timings on the GPU with the library's own kernels
(sba_cm_backward: forward, back-propagated errors, parameter gradients;
sba_cm_adam: Adam step). Installed with ``CostModel.install``, it lets the
autotuner time only the plans it ranks best (``mode=MODEL``).

Timing data and the trained model for MI355X live in ``tuning_data/``.
"""
from __future__ import annotations

import ctypes
import json
import os
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

from . import SBA_OK, SbaJob, lib

DEFAULT, EXHAUSTIVE, MODEL = 0, 1, 2
N_FEAT = 16
DATA_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning_data")
TIMINGS_PATH = os.path.join(DATA_DIR, "mi355x_plan_timings.json")
MODEL_PATH = os.path.join(DATA_DIR, "mi355x_costmodel.json")

Plan = Tuple[int, int, int]            # (last, bm, split)


# ---------------------------------------------------------------------------
# plan table
# ---------------------------------------------------------------------------
def plan_candidates(L: int, kind: int = 0) -> List[Plan]:
    """Valid plans of class (L, kind); kind = square? | radix-3? << 1."""
    buf = (ctypes.c_int * (3 * 64))()
    n = lib().sba_plan_candidates(L, kind, buf, 64)
    return [(buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]) for i in range(min(n, 64))]


def plan_heuristic(L: int) -> Plan:
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if lib().sba_plan_heuristic(L, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)) != SBA_OK:
        raise ValueError("no plan for L=%d" % L)
    return (a.value, b.value, c.value)


def plan_get(L: int, kind: int, bucket: int) -> Plan:
    a, b, c = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    if lib().sba_plan_get(L, kind, bucket, ctypes.byref(a), ctypes.byref(b),
                          ctypes.byref(c)) != SBA_OK:
        raise ValueError("bad plan key")
    return (a.value, b.value, c.value)


def plan_set(L: int, kind: int, bucket: int, plan: Plan) -> None:
    if lib().sba_plan_set(L, kind, bucket, *plan) != SBA_OK:
        raise ValueError("invalid plan %r for L=%d" % (plan, L))


def plan_reset() -> None:
    lib().sba_plan_reset()


def plan_bucket(njobs: int) -> int:
    return int(lib().sba_plan_bucket(njobs))


# ---------------------------------------------------------------------------
# measurement
# ---------------------------------------------------------------------------
def _jobs_arr(jobs):
    from . import job_array
    return job_array(jobs)


def _ws_for(arr, n, dev):
    import torch
    nb = int(lib().sba_batch_ws_bytes(ctypes.cast(arr, ctypes.c_void_p), n))
    if nb == 0:
        raise ValueError("jobs outside the band")
    return torch.empty(nb, dtype=torch.uint8, device=dev), nb


def class_time(jobs, plan: Plan, reps: int = 3, stream=None, ws=None) -> float:
    """Mean launch time (ms) of one size class of (r, a, b) device jobs
    under `plan`."""
    arr = _jobs_arr(jobs)
    dev = jobs[0][1].device
    if ws is None:
        ws, nb = _ws_for(arr, len(jobs), dev)
    else:
        nb = ws.numel()
    ms = ctypes.c_float()
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    rc = lib().sba_class_time(ctypes.cast(arr, ctypes.c_void_p), len(jobs), plan[0], plan[1],
                              plan[2], reps, ctypes.c_void_p(ws.data_ptr()), nb, s,
                              ctypes.byref(ms))
    if rc != SBA_OK:
        raise RuntimeError("sba_class_time returned %d for plan %r" % (rc, plan))
    return float(ms.value)


def autotune(jobs, mode: int = EXHAUSTIVE, topk: int = 2, reps: int = 3, stream=None) -> int:
    """Tune every size class of `jobs` ((r, a, b) device tensors); returns the
    number of plan timings made."""
    arr = _jobs_arr(jobs)
    ws, nb = _ws_for(arr, len(jobs), jobs[0][1].device)
    meas = ctypes.c_int()
    s = None if stream is None else ctypes.c_void_p(int(stream.cuda_stream))
    rc = lib().sba_autotune(ctypes.cast(arr, ctypes.c_void_p), len(jobs), mode, topk, reps,
                            ctypes.c_void_p(ws.data_ptr()), nb, s, ctypes.byref(meas))
    if rc != SBA_OK:
        raise RuntimeError("sba_autotune returned %d" % rc)
    return int(meas.value)


# ---------------------------------------------------------------------------
# features and shapes
# ---------------------------------------------------------------------------
def features(L: int, sqr: int, njobs: int, fill: float, plan: Plan) -> np.ndarray:
    f = (ctypes.c_float * N_FEAT)()
    if lib().sba_cm_features(L, sqr, njobs, fill, plan[0], plan[1], plan[2], f) != SBA_OK:
        raise ValueError("bad feature key")
    return np.array(f[:], dtype=np.float32)


def shape_operands(L: int, sqr: int, fill: float) -> Tuple[int, int]:
    """Operand lengths (an, bn) of a job in class L whose product has about
    fill * 2^L limbs (balanced; bn = an for a square)."""
    N = 1 << L
    Z = max(int(round(fill * N)), N // 2 + 2)
    Z = min(Z, N + 1)
    an = (Z + 1) // 2
    bn = an if sqr else Z - an
    while an + bn - 1 > N:
        an -= 1
        if sqr:
            bn = an
    return an, bn


def shape_operands3(L: int, sqr: int, fill: float) -> Tuple[int, int]:
    """Operand lengths (an, bn) of a job in the radix-3 class of row length
    2^L, whose product has about fill * 3 * 2^L limbs: 2^(L+1) < an+bn-1 <=
    3 * 2^L (balanced; bn = an for a square)."""
    N3 = 3 << L
    Z = min(max(int(round(fill * N3)), (2 << L) + 4), N3 + 1)
    an = (Z + 1) // 2
    bn = an if sqr else Z - an
    while an + bn - 1 > N3:
        an -= 1
        if sqr:
            bn = an
    return an, bn


def make_class_jobs(L: int, sqr: int, njobs: int, fill: float, pool, dev, r3: bool = False):
    """njobs device jobs of one size class (radix-3 with r3), operands
    sliced from `pool`."""
    import torch
    an, bn = shape_operands3(L, sqr, fill) if r3 else shape_operands(L, sqr, fill)
    P = pool.numel()
    res = torch.empty(njobs * (an + bn), dtype=torch.int64, device=dev)
    rng = np.random.default_rng(L * 1000 + njobs * 7 + sqr + (2 if r3 else 0))
    jobs = []
    for i in range(njobs):
        oa = int(rng.integers(0, P - an))
        ob = int(rng.integers(0, P - bn))
        r = res[i * (an + bn):(i + 1) * (an + bn)]
        jobs.append((r, pool[oa:oa + an], None if sqr else pool[ob:ob + bn]))
    return jobs


def default_shapes(max_elems: int = 1 << 24) -> List[Tuple[int, int, int, float]]:
    """(L, sqr, njobs, fill) grid of the timing dataset."""
    out = []
    for L in range(10, 25):
        for sqr in (0, 1):
            for nj in (1, 3, 8, 24, 64, 192, 512, 1024):
                if nj * (1 << L) > max_elems:
                    continue
                # fills of power-of-two classes (at most 0.75 of 2^L runs as
                # a radix-3 class of 3 * 2^(L-2) for L >= 15)
                for fill in (0.8, 0.95):
                    out.append((L, sqr, nj, fill))
    return out


def default_shapes3(max_elems: int = 1 << 24) -> List[Tuple[int, int, int, float]]:
    """(L, sqr, njobs, fill) grid of the radix-3 classes (row length 2^L,
    transform 3 * 2^L; fill relative to 3 * 2^L, above 2/3)."""
    out = []
    for L in range(13, 23):
        for sqr in (0, 1):
            for nj in (1, 3, 8, 24, 64, 192, 512, 1024):
                if nj * (3 << L) > max_elems:
                    continue
                for fill in (0.8, 0.95):
                    out.append((L, sqr, nj, fill))
    return out


def collect_timings(shapes, reps: int = 3, log=None, r3: bool = False) -> List[dict]:
    """Time every candidate plan of every shape on the GPU (radix-3 classes
    with r3: rows of row length 2^L, fill relative to 3 * 2^L)."""
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device="cpu").manual_seed(11)
    pool = torch.randint(-2**63, 2**63 - 1, ((1 << 23) + (1 << 20),), dtype=torch.int64,
                         generator=g).to(dev)
    rows = []
    t0 = time.time()
    for k, (L, sqr, nj, fill) in enumerate(shapes):
        jobs = make_class_jobs(L, sqr, nj, fill, pool, dev, r3=r3)
        arr = _jobs_arr(jobs)
        ws, _ = _ws_for(arr, nj, dev)
        an, bn = shape_operands3(L, sqr, fill) if r3 else shape_operands(L, sqr, fill)
        zfill = (an + bn) / float((3 if r3 else 1) << L)
        for plan in plan_candidates(L, 2 if r3 else 0):
            ms = class_time(jobs, plan, reps=reps, ws=ws)
            rows.append({"L": L, "sqr": sqr, "njobs": nj, "fill": round(zfill, 6),
                         "plan": list(plan), "ms": round(ms, 6)})
        del jobs, ws
        if log and (k % 20 == 0 or k == len(shapes) - 1):
            log("timings: %d/%d shapes, %.1f s" % (k + 1, len(shapes), time.time() - t0))
    return rows


def save_timings(rows, path: str = TIMINGS_PATH, meta: Optional[dict] = None,
                 rows3: Optional[List[dict]] = None) -> None:
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        json.dump({"meta": meta or {}, "rows": rows, "rows3": rows3 or []}, f,
                  separators=(",", ":"))


def load_timings(path: str = TIMINGS_PATH) -> List[dict]:
    """The power-of-two classes' rows (the cost model's dataset)."""
    with open(path) as f:
        return json.load(f)["rows"]


def load_timings3(path: str = TIMINGS_PATH) -> List[dict]:
    """The radix-3 classes' rows (plan table only)."""
    with open(path) as f:
        return json.load(f).get("rows3", [])


def dataset(rows) -> Tuple[np.ndarray, np.ndarray]:
    """Raw features X [n, 16] and targets y [n]: log(ms) minus the mean
    log(ms) of the row's shape. The model ranks the plans of one class, so
    it learns the differences between plans rather than the (much larger)
    differences between shapes; on held-out shapes this halves the gap to
    the best plan against absolute log-time targets."""
    X = np.stack([features(r["L"], r["sqr"], r["njobs"], r["fill"], tuple(r["plan"]))
                  for r in rows]).astype(np.float32)
    y = np.log(np.array([r["ms"] for r in rows], dtype=np.float64))
    groups: Dict[Tuple[int, int, int, float], List[int]] = {}
    for i, r in enumerate(rows):
        groups.setdefault(shape_key(r), []).append(i)
    for idx in groups.values():
        y[idx] -= y[idx].mean()
    return X, y.astype(np.float32)


def shape_key(r) -> Tuple[int, int, int, float]:
    return (r["L"], r["sqr"], r["njobs"], r["fill"])


def split_shapes(rows, held_frac: float = 0.3, seed: int = 3):
    """Split rows by shape (all plans of a shape on one side)."""
    keys = sorted({shape_key(r) for r in rows})
    rng = np.random.default_rng(seed)
    held = set(k for k in keys if rng.random() < held_frac)
    tr = [r for r in rows if shape_key(r) not in held]
    te = [r for r in rows if shape_key(r) in held]
    return tr, te


# ---------------------------------------------------------------------------
# the cost model
# ---------------------------------------------------------------------------
H1 = H2 = 32


def init_params(seed: int = 0) -> np.ndarray:
    """He-initialised parameters in the library's layout (W1, b1, W2, b2, w3, b3)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2.0 / N_FEAT), (H1, N_FEAT))
    W2 = rng.normal(0, np.sqrt(2.0 / H1), (H2, H1))
    w3 = rng.normal(0, np.sqrt(1.0 / H2), H2)
    p = np.concatenate([W1.ravel(), np.zeros(H1), W2.ravel(), np.zeros(H2), w3, [0.0]])
    return p.astype(np.float32)


@dataclass
class CostModel:
    params: np.ndarray
    fmean: np.ndarray
    fstd: np.ndarray
    ymean: float
    ystd: float
    history: List[float] = field(default_factory=list)

    def standardize(self, X: np.ndarray) -> np.ndarray:
        return ((X - self.fmean) / self.fstd).astype(np.float32)

    def save(self, path: str = MODEL_PATH, meta: Optional[dict] = None) -> None:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump({"meta": meta or {}, "params": [float(v) for v in self.params],
                       "fmean": [float(v) for v in self.fmean],
                       "fstd": [float(v) for v in self.fstd],
                       "ymean": float(self.ymean), "ystd": float(self.ystd),
                       "history": [float(v) for v in self.history]}, f)

    @staticmethod
    def load(path: str = MODEL_PATH) -> "CostModel":
        with open(path) as f:
            d = json.load(f)
        return CostModel(np.array(d["params"], np.float32), np.array(d["fmean"], np.float32),
                         np.array(d["fstd"], np.float32), d["ymean"], d["ystd"],
                         d.get("history", []))

    def install(self) -> None:
        """Make this the autotuner's model (sba_cm_load)."""
        fp = ctypes.POINTER(ctypes.c_float)
        p = np.ascontiguousarray(self.params, np.float32)
        m = np.ascontiguousarray(self.fmean, np.float32)
        s = np.ascontiguousarray(self.fstd, np.float32)
        rc = lib().sba_cm_load(p.ctypes.data_as(fp), m.ctypes.data_as(fp), s.ctypes.data_as(fp),
                               float(self.ymean), float(self.ystd))
        if rc != SBA_OK:
            raise RuntimeError("sba_cm_load returned %d" % rc)

    def predict_log_ms_gpu(self, X: np.ndarray) -> np.ndarray:
        """Relative log-time predictions of raw feature rows (comparable
        within one class), by the forward kernel."""
        import torch
        dev = torch.device("cuda:0")
        xs = torch.from_numpy(self.standardize(X)).to(dev)
        ps = torch.from_numpy(np.ascontiguousarray(self.params)).to(dev)
        y = torch.empty(X.shape[0], dtype=torch.float32, device=dev)
        rc = lib().sba_cm_forward(ps.data_ptr(), xs.data_ptr(), X.shape[0], y.data_ptr(), None)
        if rc != SBA_OK:
            raise RuntimeError("sba_cm_forward returned %d" % rc)
        torch.cuda.synchronize()
        return y.cpu().numpy().astype(np.float64) * self.ystd + self.ymean


def train_gpu(X: np.ndarray, y: np.ndarray, steps: int = 3000, lr: float = 3e-3,
              seed: int = 0, log_every: int = 100) -> CostModel:
    """Fit the cost model to (raw features, log ms) on the GPU: full-batch
    Adam on the mean squared error of standardised targets, with the
    library's backward and Adam kernels."""
    import torch
    dev = torch.device("cuda:0")
    fmean = X.mean(axis=0).astype(np.float32)
    fstd = X.std(axis=0).astype(np.float32)
    fstd[fstd < 1e-6] = 1.0
    ymean, ystd = float(y.mean()), float(y.std() or 1.0)
    cm = CostModel(init_params(seed), fmean, fstd, ymean, ystd)
    n = X.shape[0]
    xs = torch.from_numpy(cm.standardize(X)).to(dev)
    ts = torch.from_numpy(((y - ymean) / ystd).astype(np.float32)).to(dev)
    P = torch.from_numpy(cm.params.copy()).to(dev)
    G = torch.zeros_like(P)
    M = torch.zeros_like(P)
    V = torch.zeros_like(P)
    ws = torch.empty(int(lib().sba_cm_ws_floats(n)), dtype=torch.float32, device=dev)
    L = lib()
    for step in range(1, steps + 1):
        if L.sba_cm_backward(P.data_ptr(), xs.data_ptr(), ts.data_ptr(), n, G.data_ptr(),
                             ws.data_ptr(), None) != SBA_OK:
            raise RuntimeError("sba_cm_backward failed")
        if step == 1 or step % log_every == 0 or step == steps:
            pred = ws[:n]
            cm.history.append(float(((pred - ts) ** 2).mean().item()))
        if L.sba_cm_adam(P.data_ptr(), G.data_ptr(), M.data_ptr(), V.data_ptr(), step, lr,
                         None) != SBA_OK:
            raise RuntimeError("sba_cm_adam failed")
    torch.cuda.synchronize()
    cm.params = P.cpu().numpy()
    return cm


# ---------------------------------------------------------------------------
# evaluation: how good are the plans a method picks?
# ---------------------------------------------------------------------------
def pick_quality(rows, choose) -> Dict[str, float]:
    """For every shape in rows, time of the plan `choose(shape, plans)` picks
    over the best measured plan. Returns the geometric-mean slowdown, the
    worst one and the summed time of the picks vs the summed best times."""
    by = {}
    for r in rows:
        by.setdefault(shape_key(r), []).append(r)
    ratios, tsum, bsum = [], 0.0, 0.0
    for key, rs in by.items():
        best = min(r["ms"] for r in rs)
        plans = [tuple(r["plan"]) for r in rs]
        pick = choose(key, plans)
        t = [r["ms"] for r in rs if tuple(r["plan"]) == pick][0]
        ratios.append(t / best)
        tsum += t
        bsum += best
    ratios = np.array(ratios)
    return {"shapes": len(ratios), "geomean_slowdown": float(np.exp(np.log(ratios).mean())),
            "worst_slowdown": float(ratios.max()), "time_over_best": float(tsum / bsum)}


def heuristic_chooser(key, plans):
    return plan_heuristic(key[0])


def model_chooser(predict_log_ms):
    """predict_log_ms(X raw) -> relative log time; picks the plan with the
    lowest estimate."""
    def choose(key, plans):
        L, sqr, nj, fill = key
        X = np.stack([features(L, sqr, nj, fill, p) for p in plans])
        return plans[int(np.argmin(predict_log_ms(X)))]
    return choose


PLAN_HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "plan_table.h")


def shipped_table(rows, rows3=None) -> Dict[Tuple[int, int, int], Plan]:
    """(kind, L, bucket) -> the plan with the lowest mean slowdown (time over
    the shape's best) over the measured shapes of that key, kind = square? |
    radix-3? << 1 (rows3: the radix-3 classes); keys without measurements
    take the nearest measured bucket of the same (kind, L)."""
    tagged = [(r["sqr"], r) for r in rows] + [(2 | r["sqr"], r) for r in (rows3 or [])]
    best = {}
    for kind, r in tagged:
        k = (kind,) + shape_key(r)
        best[k] = min(best.get(k, float("inf")), r["ms"])
    agg: Dict[Tuple[int, int, int], Dict[Plan, List[float]]] = {}
    for kind, r in tagged:
        key = (kind, r["L"], plan_bucket(r["njobs"]))
        agg.setdefault(key, {}).setdefault(tuple(r["plan"]), []).append(
            r["ms"] / best[(kind,) + shape_key(r)])
    table = {}
    for key, plans in agg.items():
        table[key] = min(plans, key=lambda p: (float(np.mean(plans[p])), p))
    nb = int(lib().sba_plan_bucket(1 << 30)) + 1
    for kind in range(4):
        for L in range(10, 26):
            have = [b for b in range(nb) if (kind, L, b) in table]
            if not have:
                continue
            for b in range(nb):
                if (kind, L, b) not in table:
                    near = min(have, key=lambda h: (abs(h - b), h))
                    table[(kind, L, b)] = table[(kind, L, near)]
    return table


def emit_plan_header(table, path: str = PLAN_HEADER) -> None:
    nb = int(lib().sba_plan_bucket(1 << 30)) + 1
    out = ["// plan_table.h -- the shipped plan table: per (kind, L, job-count",
           "// bucket) the pass plan {last, bm, split} with the lowest mean slowdown",
           "// over the shapes of that key in the exhaustive MI355X (gfx950) sweep",
           "// tuning_data/mi355x_plan_timings.json; kind = square? | radix-3? << 1",
           "// (radix-3: L is the row length's log); keys the sweep did not reach take",
           "// the nearest measured bucket. {0, 0, 0}: not measured (the hand-written",
           "// rule applies). Generated by: python -m simd_bigint_amd.tuning table",
           "#pragma once",
           "#define SBA_PT_NL 26",
           "static const signed char kPlanTable[4][SBA_PT_NL][SBA_PLAN_BUCKETS][3] = {"]
    names = ("multiplies", "squares", "radix-3 multiplies", "radix-3 squares")
    for kind in range(4):
        out.append("  {  // %s" % names[kind])
        for L in range(26):
            ents = []
            for b in range(nb):
                p = table.get((kind, L, b), (0, 0, 0))
                ents.append("{%d, %d, %d}" % p)
            out.append("    {%s},  // L = %d" % (", ".join(ents), L))
        out.append("  },")
    out.append("};")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def main(argv=None) -> int:
    """python -m simd_bigint_amd.tuning {collect,train,report}

    collect  time every candidate plan of the default shape grid on GPU 0
             and write the timing dataset
    table    write csrc/plan_table.h (the shipped plan table) from the
             dataset
    train    fit the cost model to the dataset's training shapes on the GPU
             and write it
    report   pick quality of the default plans and of the model on the
             held-out shapes"""
    import argparse
    import sys
    ap = argparse.ArgumentParser(prog="python -m simd_bigint_amd.tuning")
    ap.add_argument("cmd", choices=["collect", "train", "report", "table"])
    ap.add_argument("--timings", default=TIMINGS_PATH)
    ap.add_argument("--model", default=MODEL_PATH)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--steps", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    a = ap.parse_args(argv)

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    if a.cmd == "collect":
        import torch
        shapes = default_shapes()
        log("collect: %d shapes" % len(shapes))
        rows = collect_timings(shapes, reps=a.reps, log=log)
        shapes3 = default_shapes3()
        log("collect: %d radix-3 shapes" % len(shapes3))
        rows3 = collect_timings(shapes3, reps=a.reps, log=log, r3=True)
        save_timings(rows, a.timings, {"device": torch.cuda.get_device_name(0),
                                       "reps": a.reps, "shapes": len(shapes),
                                       "shapes3": len(shapes3)}, rows3=rows3)
        log("collect: %d + %d radix-3 rows -> %s" % (len(rows), len(rows3), a.timings))
    elif a.cmd == "train":
        rows = load_timings(a.timings)
        tr, te = split_shapes(rows)
        X, y = dataset(tr)
        t0 = time.time()
        cm = train_gpu(X, y, steps=a.steps, lr=a.lr)
        log("train: %d rows, %d steps, %.1f s, loss %.4f -> %.4f" %
            (len(tr), a.steps, time.time() - t0, cm.history[0], cm.history[-1]))
        cm.save(a.model, {"train_rows": len(tr), "steps": a.steps, "lr": a.lr,
                          "split": "split_shapes(held_frac=0.3, seed=3)"})
    elif a.cmd == "table":
        emit_plan_header(shipped_table(load_timings(a.timings), load_timings3(a.timings)))
        log("table: wrote %s" % PLAN_HEADER)
    else:
        rows = load_timings(a.timings)
        tr, te = split_shapes(rows)
        cm = CostModel.load(a.model)
        out = {"held_out_heuristic": pick_quality(te, heuristic_chooser),
               "held_out_model": pick_quality(te, model_chooser(cm.predict_log_ms_gpu)),
               "train_heuristic": pick_quality(tr, heuristic_chooser),
               "train_model": pick_quality(tr, model_chooser(cm.predict_log_ms_gpu))}
        print(json.dumps(out, indent=1))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
