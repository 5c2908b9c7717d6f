"""Train the per-front update-precision network on solver-generated labels and evaluate it.

Training matrices: seeded indefinite Helmholtz systems (7-point, 5 % negative eigenvalues) on
N^3 grids and seeded KKT saddle-point systems (MC64 scaling + matching-based ordering), both as
the workloads module builds them.  Every front with update work gets the label
precision.label_fronts assigns (cheapest precision whose factorization still refines to 1e-12
within 2 steps; the learned solve then allows 3).  The network is trained on the GPU (rv_mlp_train) and evaluated on held-out
matrices against the plain FP64 factorization and the uniform FP32 / BF16 forms.

    python rivrs-linalg_amd/train_precision.py --out model.json --report report.json [--quick]
"""
import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import precision as pr  # noqa: E402
import rivrs_sparse as rs  # noqa: E402
import workloads  # noqa: E402


def say(*a):
    print(*a, flush=True)


def system(kind, size, seed):
    """(name, matrix, rhs, analysed solver) of one workload."""
    if kind == "helmholtz":
        A, b, info = workloads.helmholtz(size, seed=seed)
        s = rs.SparseLDLT.analyze(A, rs.AnalyzeOptions(rs.OrderingStrategy.Grid, grid=info["grid"]))
    else:
        A, b, _x, info = workloads.kkt(size, seed=seed)
        s = rs.SparseLDLT.analyze_with_matrix(A, rs.AnalyzeOptions(rs.OrderingStrategy.MatchOrderMetis))
    return "%s%d-s%d" % (kind, size, seed), A, b, s


def schur_split(s, prec):
    """Fraction of contribution flops per precision class for an assignment."""
    per = s.per_supernode_stats()
    fl = np.array([q.num_eliminated * (q.front_size - q.num_fully_summed) *
                   (q.front_size - q.num_fully_summed + 1.0) for q in per])
    tot = max(fl.sum(), 1.0)
    return {name: float(fl[prec == code].sum() / tot) for name, code in
            (("bf16", pr.PREC_BF16), ("fp32", pr.PREC_FP32), ("fp64", pr.PREC_FP64))}


def timed(fn, reps=3):
    best = None
    out = None
    for _ in range(reps):
        t = time.perf_counter()
        out = fn()
        dt = time.perf_counter() - t
        best = dt if best is None else min(best, dt)
    return best, out


def evaluate(model, kind, size, seed, reps=3):
    name, A, b, s = system(kind, size, seed)
    s.factor(A)
    inert = s.inertia()
    iref = (inert.positive, inert.negative, inert.zero)
    n = s.num_supernodes()
    rec = {"matrix": name, "n": A.n, "inertia": iref}

    def run(prec, max_iter):
        s.set_front_precision(prec)
        s.refactor(A)
        x, it, be = s.solve_refine(b, tol=1e-12, max_iter=max_iter)
        i = s.inertia()
        return it, be, (i.positive, i.negative, i.zero) == iref

    for label, code in (("fp64", 0), ("fp32", 1), ("bf16", 2)):
        t, (it, be, ok) = timed(lambda: run(np.full(n, code, dtype=np.int8), 3 if code else 0), reps)
        rec[label] = {"sec": t, "corrections": it, "backward_error": be, "converged": bool(be <= 1e-12 and ok)}
    s.set_front_precision(None)
    s.refactor(A)
    prec = pr.assign(s, model)
    t, res = timed(lambda: pr.solve_learned(s, A, b, model, prec=prec, inertia_ref=iref), reps)
    rec["learned"] = {"sec": t, "corrections": res.corrections, "backward_error": res.backward_error,
                      "converged": bool(res.converged), "fell_back": bool(res.fell_back),
                      "schur_flops_split": schur_split(s, prec)}
    rec["learned"]["speedup_vs_fp64"] = rec["fp64"]["sec"] / t
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--report", required=True)
    ap.add_argument("--quick", action="store_true", help="small training and evaluation set")
    ap.add_argument("--epochs", type=int, default=3000)
    args = ap.parse_args()
    if args.quick:
        train = [("helmholtz", N, 0) for N in (12, 16, 20)] + [("kkt", m, 0) for m in (24, 32)]
        held = [("helmholtz", 24, 7), ("kkt", 40, 7)]
    else:
        train = ([("helmholtz", N, sd) for N in (16, 20, 24, 28, 32, 36, 40) for sd in (0, 1)] +
                 [("kkt", m, sd) for m in (32, 48, 64, 80, 96) for sd in (0, 1)])
        held = [("helmholtz", 64, 7), ("helmholtz", 48, 8), ("kkt", 128, 7), ("kkt", 112, 8)]
    Fs, ys, infos = [], [], []
    t0 = time.time()
    for kind, size, seed in train:
        name, A, b, s = system(kind, size, seed)
        s.factor(A)
        # labels demand convergence within 2 refinement steps, one fewer than the solve allows:
        # the margin absorbs the spread between training and unseen matrices
        F, y, info = pr.label_fronts(s, A, b, max_iter=2)
        info["matrix"] = name
        info["label_counts"] = np.bincount(y, minlength=3).tolist()
        say("labels", json.dumps(info))
        Fs.append(F)
        ys.append(y)
        infos.append(info)
    F = np.concatenate(Fs)
    y = np.concatenate(ys)
    counts = np.bincount(y, minlength=3)
    cw = np.where(counts > 0, len(y) / (3.0 * np.maximum(counts, 1)), 0.0)
    model = pr.PrecisionMLP.initial(seed=0)
    t1 = time.time()
    loss = model.fit(F, y, class_weight=cw, epochs=args.epochs, lr=3e-3)
    acc = float(np.mean(model.predict(F) == y))
    say("trained on %d fronts from %d matrices: loss %.4f, accuracy %.3f, %.1f s (labels %.1f s)" % (
        len(y), len(train), loss, acc, time.time() - t1, t1 - t0))
    evals = [evaluate(model, *h) for h in held]
    for e in evals:
        say("eval", json.dumps(e))
    train_info = {"matrices": [i["matrix"] for i in infos], "fronts": int(len(y)),
                  "label_counts": counts.tolist(), "class_weight": cw.tolist(), "epochs": args.epochs,
                  "final_loss": loss, "train_accuracy": acc, "label_seconds": t1 - t0}
    model.save(args.out, train_info)
    with open(args.report, "w") as f:
        json.dump({"training": train_info, "labels": infos, "evaluation": evals}, f, indent=1)
    say("done")


if __name__ == "__main__":
    main()
