"""Autotuner sweep, selector training and regret evaluation.

  python -m numc_amd.autotune sweep --out db.csv [--start i --stop j]
  python -m numc_amd.autotune train --db db.csv --out numc_amd/data/selector.bin
  python -m numc_amd.autotune eval  --db numc_amd/data/autotune_db.csv.gz --model numc_amd/data/selector.bin

sweep: shapes 0..19999 (training) and 20000..21999 (held out), drawn from a
seeded generator: M, N log-uniform integers in [1, 16384] (15% skinny,
M log-uniform in [1, 32]), K log-uniform in [1, 65536], a kernel family
uniform over the seven families, and layout flags (20% A stored transposed,
20% B stored transposed, 10% padded rows).  Each shape's every configuration
is timed by numc_autotune_shape (whole calls on device scratch operands).

train: the MLP cost model (24 features -> 64 -> 64 -> log2 ms) is trained by
the library's HIP kernels (numc_mlp_train) on the training shapes.

eval: regret = t(chosen) / t(best measured) - 1 over the held-out shapes, for
the learned selector and for the analytic model it replaces.
"""
import argparse
import csv
import ctypes as C
import os
import struct
import sys
import time

import numpy as np

from . import (FAMILY_NAMES, Ctx, check, config_count, config_get, lib)

N_TRAIN, N_HOLDOUT = 20000, 2000
N_FAMILIES = 7   # every kernel family (FAMILY_F32 .. FAMILY_F16)
MODEL_VERSION = 2
HIDDEN = 64
HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_DB = os.path.join(HERE, "data", "autotune_db.csv.gz")
DEFAULT_MODEL = os.path.join(HERE, "data", "selector.bin")


def _logu(rng, lo, hi, n):
    return np.clip(np.round(np.exp(rng.uniform(np.log(lo), np.log(hi), n))), lo, hi).astype(np.int64)


def sample_shapes(seed=1, n=N_TRAIN + N_HOLDOUT):
    """(family, M, N, K, flags) for every shape index, deterministic in seed."""
    rng = np.random.default_rng(seed)
    fam = rng.integers(0, N_FAMILIES, n)
    M = _logu(rng, 1, 16384, n)
    N = _logu(rng, 1, 16384, n)
    K = _logu(rng, 1, 65536, n)
    skinny = rng.random(n) < 0.15
    M = np.where(skinny, _logu(rng, 1, 32, n), M)
    flags = ((rng.random(n) < 0.2) * 1 | (rng.random(n) < 0.2) * 2 | (rng.random(n) < 0.1) * 4).astype(np.int64)
    return [(int(f), int(m), int(nn), int(k), int(fl)) for f, m, nn, k, fl in zip(fam, M, N, K, flags)]


def _setup():
    L = lib()
    L.numc_autotune_shape.argtypes = [C.c_void_p] + [C.c_int] * 6 + [C.POINTER(C.c_float), C.c_int]
    L.numc_autotune_shape.restype = C.c_int
    L.numc_selector_features.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_float)]
    L.numc_selector_features.restype = C.c_int
    L.numc_selector_feature_count.restype = C.c_int
    L.numc_selector_predict.argtypes = [C.c_int] * 5 + [C.POINTER(C.c_float), C.c_int]
    L.numc_selector_load.argtypes = [C.c_char_p]
    L.numc_selector_heuristic.argtypes = [C.c_int] * 4
    L.numc_selector_heuristic.restype = C.c_int
    L.numc_selector_choice.argtypes = [C.c_int] * 5
    L.numc_selector_choice.restype = C.c_int
    L.numc_selector_candidates.argtypes = [C.c_int] * 6 + [C.POINTER(C.c_int), C.c_int]
    L.numc_selector_candidates.restype = C.c_int
    L.numc_mlp_create.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_uint64]
    L.numc_mlp_create.restype = C.c_void_p
    L.numc_mlp_free.argtypes = [C.c_void_p]
    L.numc_mlp_param_count.argtypes = [C.c_void_p]
    L.numc_mlp_param_count.restype = C.c_size_t
    L.numc_mlp_set_params.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_size_t]
    L.numc_mlp_get_params.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_size_t]
    L.numc_mlp_forward.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_longlong, C.POINTER(C.c_float)]
    L.numc_mlp_loss_grad.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                     C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_double),
                                     C.POINTER(C.c_float)]
    L.numc_mlp_train.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.c_longlong, C.c_int, C.c_int, C.c_float, C.c_float,
                                 C.c_uint64, C.POINTER(C.c_double)]
    return L


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def features(fam, M, N, K, flags, cfg):
    L = _setup()
    out = np.zeros(L.numc_selector_feature_count(), np.float32)
    check(L.numc_selector_features(fam, M, N, K, flags, cfg, fptr(out)), "numc_selector_features")
    return out


# ---- sweep -----------------------------------------------------------------------

def sweep(out, start, stop, seed=1, log=print):
    L = _setup()
    shapes = sample_shapes(seed)
    done = set()
    if os.path.exists(out):
        with open(out) as f:
            for row in csv.DictReader(f):
                done.add(int(row["idx"]))
    new = not os.path.exists(out)
    fh = open(out, "a", newline="")
    w = csv.writer(fh)
    if new:
        w.writerow(["idx", "family", "M", "N", "K", "flags", "cfg", "ms"])
    ctx = Ctx()
    buf = (C.c_float * 16)()
    t0 = last = time.time()
    n = 0
    for idx in range(start, min(stop, len(shapes))):
        if idx in done:
            continue
        fam, M, N, K, flags = shapes[idx]
        flop = 2.0 * M * N * K
        reps = 5 if flop < 1e9 else (3 if flop < 1e11 else 1)
        check(L.numc_autotune_shape(ctx.h, fam, M, N, K, flags, reps, buf, 16), "numc_autotune_shape")
        for c in range(config_count(fam)):
            w.writerow([idx, fam, M, N, K, flags, c, "%.6g" % buf[c]])
        n += 1
        if time.time() - last > 20:
            fh.flush()
            last = time.time()
            log("sweep: %d shapes in %.0f s (idx %d)" % (n, last - t0, idx), flush=True)
    fh.close()
    lib().numc_autotune_release(ctx.h)
    ctx.close()
    log("sweep: done %d shapes in %.0f s" % (n, time.time() - t0), flush=True)


def load_db(path):
    import gzip
    rows = {}
    with (gzip.open(path, "rt") if path.endswith(".gz") else open(path)) as f:
        for r in csv.DictReader(f):
            idx = int(r["idx"])
            e = rows.setdefault(idx, {"family": int(r["family"]), "M": int(r["M"]), "N": int(r["N"]),
                                      "K": int(r["K"]), "flags": int(r["flags"]), "ms": {}})
            e["ms"][int(r["cfg"])] = float(r["ms"])
    return rows


def design(db, idxs):
    X, T = [], []
    for i in idxs:
        e = db[i]
        for c, ms in sorted(e["ms"].items()):
            X.append(features(e["family"], e["M"], e["N"], e["K"], e["flags"], c))
            T.append(np.log2(max(ms, 1e-6)))
    return np.array(X, np.float32), np.array(T, np.float32)[:, None]


# ---- selector file -------------------------------------------------------------------

def save_model(path, params, din, hidden):
    with open(path, "wb") as f:
        f.write(b"NMLS")
        f.write(struct.pack("<5i", MODEL_VERSION, din, hidden, 1, N_FAMILIES))
        for fam in range(N_FAMILIES):
            n = config_count(fam)
            f.write(struct.pack("<i", n))
            for i in range(n):
                c = config_get(fam, i)
                f.write(struct.pack("<8i", c["bm"], c["bn"], c["kc"], c["waves_m"], c["waves_n"], c["split_k"],
                                    c["waves_k"], c["direct"]))
        f.write(np.asarray(params, np.float32).tobytes())


def load_model_params(path):
    with open(path, "rb") as f:
        data = f.read()
    assert data[:4] == b"NMLS"
    ver, din, hidden, dout, nfam = struct.unpack("<5i", data[4:24])
    o = 24
    for _ in range(nfam):
        (n,) = struct.unpack("<i", data[o:o + 4])
        o += 4 + 32 * n
    return np.frombuffer(data[o:], np.float32).copy(), din, hidden, dout


# ---- training and evaluation ------------------------------------------------------------

def train(db_path, out, epochs=80, batch=512, seed=7, log=print):
    L = _setup()
    db = load_db(db_path)
    train_idx = [i for i in sorted(db) if i < N_TRAIN]
    X, T = design(db, train_idx)
    din = X.shape[1]
    log("train: %d rows from %d shapes, %d features" % (len(X), len(train_idx), din), flush=True)
    ctx = Ctx()
    net = L.numc_mlp_create(ctx.h, din, HIDDEN, 1, seed)
    if not net:
        raise RuntimeError("numc_mlp_create failed")
    losses = np.zeros(epochs, np.float64)
    check(L.numc_mlp_train(net, fptr(X), fptr(T), None, len(X), epochs, batch, 3e-3, 2e-5, seed,
                           losses.ctypes.data_as(C.POINTER(C.c_double))), "numc_mlp_train")
    n = L.numc_mlp_param_count(net)
    params = np.zeros(n, np.float32)
    check(L.numc_mlp_get_params(net, fptr(params), n), "numc_mlp_get_params")
    L.numc_mlp_free(net)
    ctx.close()
    save_model(out, params, din, HIDDEN)
    log("train: loss first %.4f last %.4f -> %s" % (losses[0], losses[-1], out), flush=True)
    return losses


def evaluate(db_path, model_path, idxs=None, use_gpu=False):
    """regret statistics of the learned selector and of the analytic model"""
    L = _setup()
    db = load_db(db_path)
    if idxs is None:
        idxs = [i for i in sorted(db) if i >= N_TRAIN]
    check(L.numc_selector_load(model_path.encode()), "numc_selector_load")
    pred_buf = (C.c_float * 16)()
    reg_mlp, reg_heur, reg_fb, t_best, t_mlp, t_heur, t_fb = [], [], [], 0.0, 0.0, 0.0, 0.0
    n_timed = 0
    cand = (C.c_int * 32)()
    by_family = {}
    for i in idxs:
        e = db[i]
        fam, M, N, K, fl = e["family"], e["M"], e["N"], e["K"], e["flags"]
        ms = e["ms"]
        best = min(ms.values())
        check(L.numc_selector_predict(fam, M, N, K, fl, pred_buf, 16), "numc_selector_predict")
        choice = int(np.argmin([pred_buf[c] for c in range(config_count(fam))]))
        h = L.numc_selector_heuristic(fam, M, N, K)
        # the call-time rule: where the model's best few are close (or the call
        # is tiny), numc_matmul times them once and keeps the fastest
        nc_ = L.numc_selector_candidates(fam, M, N, K, fl, 1, cand, 32)
        fb = min((cand[i] for i in range(nc_)), key=lambda c: ms[c]) if nc_ > 0 else choice
        n_timed += nc_ > 1
        reg_mlp.append(ms[choice] / best - 1)
        reg_heur.append(ms[h] / best - 1)
        reg_fb.append(ms[fb] / best - 1)
        t_best += best
        t_mlp += ms[choice]
        t_heur += ms[h]
        t_fb += ms[fb]
        by_family.setdefault(FAMILY_NAMES[fam], []).append(ms[choice] / best - 1)
    rm, rh, rf = np.array(reg_mlp), np.array(reg_heur), np.array(reg_fb)
    return {"shapes": len(idxs),
            "mlp_with_call_timing": {"mean_regret": float(rf.mean()), "p95_regret": float(np.percentile(rf, 95)),
                                     "max_regret": float(rf.max()), "total_time_ratio": t_fb / t_best,
                                     "exact_best_frac": float(np.mean(rf == 0)),
                                     "shapes_timed": int(n_timed)},
            "mlp": {"mean_regret": float(rm.mean()), "p95_regret": float(np.percentile(rm, 95)),
                    "max_regret": float(rm.max()), "total_time_ratio": t_mlp / t_best,
                    "exact_best_frac": float(np.mean(rm == 0))},
            "analytic": {"mean_regret": float(rh.mean()), "p95_regret": float(np.percentile(rh, 95)),
                         "max_regret": float(rh.max()), "total_time_ratio": t_heur / t_best,
                         "exact_best_frac": float(np.mean(rh == 0))},
            "mlp_mean_regret_by_family": {k: float(np.mean(v)) for k, v in by_family.items()}}


def pin(specs, out, reps=10, rounds=3, log=print):
    """Exhaustively autotune exact shapes and record the winners in the
    exact-shape table (lines family,M,N,K,flags,cfg): the per-configuration
    time is the minimum over `rounds` rounds of `reps` calls each."""
    L = _setup()
    ctx = Ctx()
    buf = (C.c_float * 16)()
    rows = {}
    if os.path.exists(out):
        for line in open(out):
            if line.strip() and not line.startswith("#"):
                f = line.strip().split(",")
                rows[tuple(int(x) for x in f[:5])] = int(f[5])
    for fam, M, N, K, fl in specs:
        n = config_count(fam)
        best = [float("inf")] * n
        for _ in range(rounds):
            check(L.numc_autotune_shape(ctx.h, fam, M, N, K, fl, reps, buf, 16), "numc_autotune_shape")
            best = [min(best[c], buf[c]) for c in range(n)]
        cfg = int(np.argmin(best))
        rows[(fam, M, N, K, fl)] = cfg
        log("pin: family %d %dx%dx%d flags %d -> config %d (%s)" % (
            fam, M, N, K, fl, cfg, " ".join("%.4f" % t for t in best)), flush=True)
    lib().numc_autotune_release(ctx.h)
    ctx.close()
    with open(out, "w") as f:
        f.write("# exact-shape configuration table: family,M,N,K,flags,cfg (python -m numc_amd.autotune pin)\n")
        for k in sorted(rows):
            f.write("%d,%d,%d,%d,%d,%d\n" % (k + (rows[k],)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["sweep", "train", "eval", "pin"])
    ap.add_argument("--shapes", default="1:8192:8192:8192:0,0:8192:8192:8192:0,2:8192:8192:8192:0",
                    help="pin: family:M:N:K:flags,...")
    ap.add_argument("--out", default=None)
    ap.add_argument("--db", default=DEFAULT_DB)
    ap.add_argument("--model", default=DEFAULT_MODEL)
    ap.add_argument("--start", type=int, default=0)
    ap.add_argument("--stop", type=int, default=N_TRAIN + N_HOLDOUT)
    ap.add_argument("--epochs", type=int, default=80)
    a = ap.parse_args()
    if a.cmd == "sweep":
        sweep(a.out or a.db, a.start, a.stop)
    elif a.cmd == "train":
        train(a.db, a.out or a.model, epochs=a.epochs)
    elif a.cmd == "pin":
        specs = [tuple(int(x) for x in s.split(":")) for s in a.shapes.split(",")]
        pin(specs, a.out or os.path.join(os.path.dirname(DEFAULT_MODEL), "tuned.csv"))
    else:
        import json
        print(json.dumps(evaluate(a.db, a.model), indent=1))


if __name__ == "__main__":
    main()
