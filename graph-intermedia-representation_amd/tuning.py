"""Learned tuning of the GIR path: record collection, training, model files.

The method generates its own training data: for each graph of a family set
(RMAT with varied skew, percolated 2D lattices, Erdos-Renyi graphs) and a few
query-batch sizes k, the GPU times SSSP for a ladder of bucket widths delta
and buildSCC_DAG for a ladder of trim-round counts; each timing is one
record (features, knob, value, ms).  The MLP (16 -> 64 -> 64 -> 2, trained
on the GPU by the library's own kernels) learns
    y0 = log2(fastest delta)                           for (graph, k)
    y1 = index of the fastest SCC schedule in SCHEDULES for the graph
where a schedule is (pivot rule, trim rounds): the remainder after trimming
goes straight to colouring (pivot 0) or first through one forward-backward
sweep from the vertex of largest out * in degree (pivot 1).
At run time girhip_learned_tune applies it with the reference's defaults as
fallback and judge.  Knobs only change work order; results never depend on
them (tests/test_gpu_tuning.py checks that).

    python tuning.py collect OUT.npz [--graphs N] [--seed S]   (GPU)
    python tuning.py merge OUT.npz RECORDS.npz [...]
    python tuning.py train RECORDS.npz [...] --out MODEL.npz  (GPU)
    python tuning.py evaluate [--model M] [--graphs rmat23,grid4800]  (GPU)
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import girhip  # noqa: E402

DELTAS = [2.0 ** i for i in range(0, 17)]  # 1 .. 65536
# (pivot rule, trim rounds), the table girhip_learned_tune indexes
# (csrc/tune.hip); the reference's default is (1, 3)
SCHEDULES = [(0, 1), (0, 3), (1, 0), (1, 1), (1, 2), (1, 3), (1, 4), (1, 6), (1, 8)]
DEFAULT_SCHEDULE = SCHEDULES.index((1, 3))
KS = [1, 8, 32]
MODEL_PATH = os.path.join(HERE, "tuned_mlp.npz")


def log(*a):
    print("[tuning]", *a, file=sys.stderr, flush=True)


def family_graph(rng):
    """One training graph drawn from the families (name, V, fb, fa, bb, ba,
    weight spec)."""
    kind = rng.choice(["rmat", "rmat", "grid", "er"])
    if kind == "rmat":
        scale = int(rng.choice([16, 17, 18, 18, 19, 19, 20, 20, 21, 22]))
        a = float(rng.uniform(0.45, 0.65))
        d = 0.05
        b = c = (1 - a - d) / 2
        ef = int(rng.choice([8, 16, 16, 32]))
        g = girhip.rmat_graph(scale, ef, a, b, c, seed=int(rng.integers(1 << 30)))
        name = f"rmat{scale}_a{a:.2f}_ef{ef}"
    elif kind == "grid":
        side = int(rng.choice([200, 400, 700, 1000, 1400, 2000]) * rng.uniform(0.8, 1.2))
        pk = float(rng.uniform(0.7, 0.95))
        pb = float(rng.uniform(0.3, 0.9))
        g = girhip.grid_graph(side, side, pk, pb, seed=int(rng.integers(1 << 30)))
        name = f"grid{side}_k{pk:.2f}_b{pb:.2f}"
    else:
        V = int(2 ** rng.integers(15, 22))
        deg = float(rng.uniform(4, 32))
        s, dd = girhip.uniform_edges(V, int(V * deg), seed=int(rng.integers(1 << 30)))
        g = (V,) + girhip.csr_from_edges(V, s, dd)
        name = f"er{V}_d{deg:.1f}"
    wmax = int(rng.choice([0, 16, 64, 256, 1000, 10000]))  # 0: the reference's hash weights
    return name, g, wmax, kind


def time_sssp(G, src, delta, reps=1):
    best = 1e300
    for _ in range(reps):
        G.sssp(src, delta)
        best = min(best, G.last_ms("sssp"))
    return best


def collect(n_graphs, seed, out_path, budget_s=None):
    rng = np.random.default_rng(seed)
    feats, kinds, vals, ms, gids, names = [], [], [], [], [], []
    t0 = time.time()
    for gi in range(n_graphs):
        if budget_s and time.time() - t0 > budget_s:
            log(f"time budget reached after {gi} graphs")
            break
        name, (V, fb, fa, bb, ba), wmax, kind = family_graph(rng)
        with girhip.Graph(fb, fa, bb, ba) as G:
            if wmax:
                G.set_weights(girhip.weights(len(fa), 1, wmax, int(rng.integers(1 << 30))))
            for k in KS:
                src = girhip.sources(fb, k, int(rng.integers(1 << 30)))
                f = G.features(k)
                # delta ladder from the top; stop once far slower than the best
                # or slow in absolute terms (the slope below is then known)
                best = 1e300
                for delta in reversed(DELTAS):
                    t = time_sssp(G, src, delta)
                    feats.append(f)
                    kinds.append(0)
                    vals.append(delta)
                    ms.append(t)
                    gids.append(gi * 10 + k)
                    best = min(best, t)
                    if (t > 8 * best and delta < 64) or (t > 3 * best and t > 1500):
                        break
            f = G.features(KS[0])
            # colouring alone takes one propagation per hop of the longest
            # path inside an SCC: not tried on lattices or large graphs
            colour_ok = kind != "grid" and V <= (1 << 20)
            for si, (piv, trim) in enumerate(SCHEDULES):
                if piv == 0 and not colour_ok:
                    continue
                G.set_param("scc_pivot", piv)
                G.set_param("scc_trim_rounds", trim)
                t = 1e300
                for _ in range(2):
                    G.scc_dag()
                    t = min(t, G.last_ms("scc_dag"))
                feats.append(f)
                kinds.append(1)
                vals.append(si)
                ms.append(t)
                gids.append(gi * 10)
            G.set_param("scc_pivot", 1)
            G.set_param("scc_trim_rounds", 3)
        names.append(name)
        log(f"graph {gi} {name} V={V} E={len(fa)} w={wmax or 'hash'} records={len(ms)} {time.time() - t0:.0f}s")
        # saved after every graph, so an interrupted run keeps what it measured
        np.savez_compressed(out_path, feat=np.array(feats, np.float32), kind=np.array(kinds, np.int8),
                            value=np.array(vals, np.float64), ms=np.array(ms, np.float64),
                            gid=np.array(gids, np.int64), names=np.array(names))
    return len(ms)


def merge_records(paths, out_path):
    """Concatenate record files (graph ids made distinct per file)."""
    cols = {k: [] for k in ("feat", "kind", "value", "ms", "gid", "names")}
    off = 0
    for p in paths:
        with np.load(p, allow_pickle=False) as z:
            for k in ("feat", "kind", "value", "ms"):
                cols[k].append(z[k])
            cols["gid"].append(z["gid"] + off)
            cols["names"].append(z["names"])
            off += 10 * (int(z["gid"].max()) // 10 + 1)
    np.savez_compressed(out_path, **{k: np.concatenate(v) for k, v in cols.items()})
    return sum(len(x) for x in cols["ms"])


def robust_best(values, times, tol=0.05, log=True):
    """Knob label of one timing ladder: the median (in log2 when log) of the
    values within tol of the fastest, so a flat optimum gives its centre
    rather than whichever point timing noise favoured."""
    t = np.asarray(times, np.float64)
    v = np.asarray(values, np.float64)
    good = v[t <= t.min() * (1 + tol)]
    good = np.sort(np.log2(good) if log else good)
    return float(good[(len(good) - 1) // 2])


def dataset(records):
    """Per (graph, k): features -> (log2 best delta, best SCC schedule index)."""
    feat, kind, value, ms, gid = (records[k] for k in ("feat", "kind", "value", "ms", "gid"))
    best_trim = {}
    for g in np.unique(gid[kind == 1]):
        sel = (gid == g) & (kind == 1)
        best_trim[g // 10] = robust_best(value[sel], ms[sel], log=False)
    X, Y = [], []
    for g in np.unique(gid[kind == 0]):
        sel = (gid == g) & (kind == 0)
        if (g // 10) not in best_trim:
            continue
        X.append(feat[sel][0])
        Y.append([robust_best(value[sel], ms[sel]), best_trim[g // 10]])
    return np.array(X, np.float32), np.array(Y, np.float32)


def fit_norm(X, Y):
    xm, xs = X.mean(0), X.std(0)
    xs[xs < 1e-6] = 1.0
    ym, ys = Y.mean(0), Y.std(0)
    ys[ys < 1e-6] = 1.0
    return np.concatenate([xm, xs, ym, ys]).astype(np.float32)


def train(X, Y, epochs=3000, lr=3e-3, seed=1):
    norm = fit_norm(X, Y)
    xn = (X - norm[:16]) / norm[16:32]
    yn = (Y - norm[32:34]) / norm[34:36]
    m = girhip.Mlp(seed)
    loss = None
    for i in range(0, epochs, 500):
        loss = m.train(xn, yn, min(500, epochs - i), lr)
        log(f"epoch {i + 500}: loss {loss:.4f}")
    return m, norm, loss


def save_model(path, params, norm, meta=""):
    np.savez(path, params=np.asarray(params, np.float32), norm=np.asarray(norm, np.float32), meta=np.array(meta))


def load_model(path=MODEL_PATH):
    with np.load(path, allow_pickle=False) as z:
        params, norm = z["params"], z["norm"]
    m = girhip.Mlp(1)
    m.params = params
    return m, norm


def evaluate(model_path, which=("rmat23", "grid4800"), n_src=8, reps=2):
    """Held-out evaluation: learned vs default knobs (delta 16, 3 trim
    rounds) on graphs outside the training families' sizes; every learned
    result is checked identical to the default one."""
    import torch
    model, norm = load_model(model_path)
    dev = torch.device("cuda:0")
    out = []
    for name in which:
        if name.startswith("rmat"):
            V, fb, fa, bb, ba = girhip.rmat_graph(int(name[4:]), 16, 0.57, 0.19, 0.19, seed=5)
            wmax = 64
        else:
            n = int(name[4:])
            V, fb, fa, bb, ba = girhip.grid_graph(n, n, 0.9, 0.8, seed=3)
            wmax = 1000
        w = girhip.weights(len(fa), 1, wmax, 12)
        src = girhip.sources(fb, n_src, 9)
        G = girhip.Graph(*[torch.from_numpy(x).to(dev) for x in (fb, fa, bb, ba)])
        G.set_weights(torch.from_numpy(w.view(np.int32)).to(dev))
        dist = torch.empty((n_src, V), dtype=torch.int64, device=dev)

        def run():
            s_ms, c_ms = [], []
            for _ in range(reps):
                G.sssp(src, 0, out=dist)
                s_ms.append(G.last_ms("sssp"))
                G.scc_dag()
                c_ms.append(G.last_ms("scc_dag"))
            return min(s_ms), min(c_ms)

        G.set_param("sssp_delta", 16)
        G.set_param("scc_trim_rounds", 3)
        G.set_param("scc_pivot", 1)
        sd, cd = run()
        ref_d, ref_s = dist.clone(), G.get_scc()[0].copy()
        res = G.learned_tune(model, norm, src, judge=0)  # the raw prediction
        sl, cl = run()
        same = bool(torch.equal(ref_d, dist)) and bool((G.get_scc()[0] == ref_s).all())
        G.set_param("sssp_delta", 16)
        G.set_param("scc_trim_rounds", 3)
        G.set_param("scc_pivot", 1)
        rj = G.learned_tune(model, norm, src, judge=n_src)  # judged by the default on the same sources
        sj, _ = run()
        same = same and bool(torch.equal(ref_d, dist))
        rec = {"graph": name, "V": V, "E": len(fa), "sources": n_src, "learned_delta": res["delta"],
               "learned_trim": res["trim"], "learned_pivot": res["pivot"], "model_used": res["used"], "sssp_ms_default": round(sd, 3),
               "sssp_ms_learned": round(sl, 3), "sssp_gain": round(sd / sl, 3), "scc_dag_ms_default": round(cd, 3),
               "scc_dag_ms_learned": round(cl, 3), "scc_dag_gain": round(cd / cl, 3),
               "judged_delta": rj["delta"], "judged_kept_prediction": rj["used"], "sssp_ms_judged": round(sj, 3),
               "sssp_gain_judged": round(sd / sj, 3), "results_identical": same}
        log(rec)
        out.append(rec)
        G.close()
        del dist, ref_d
        torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("out")
    c.add_argument("--graphs", type=int, default=400)
    c.add_argument("--seed", type=int, default=1)
    c.add_argument("--budget-s", type=float, default=None)
    t = sub.add_parser("train")
    t.add_argument("records", nargs="+")
    t.add_argument("--out", default=MODEL_PATH)
    t.add_argument("--epochs", type=int, default=3000)
    t.add_argument("--lr", type=float, default=3e-3)
    mg = sub.add_parser("merge")
    mg.add_argument("out")
    mg.add_argument("records", nargs="+")
    e = sub.add_parser("evaluate")
    e.add_argument("--model", default=MODEL_PATH)
    e.add_argument("--graphs", default="rmat23,grid4800")
    e.add_argument("--out", default=None)
    args = ap.parse_args()
    if args.cmd == "evaluate":
        import json
        res = evaluate(args.model, tuple(args.graphs.split(",")))
        if args.out:
            with open(args.out, "w") as f:
                json.dump(res, f, indent=1)
        return
    if args.cmd == "merge":
        log(f"{merge_records(args.records, args.out)} records -> {args.out}")
        return
    if args.cmd == "collect":
        n = collect(args.graphs, args.seed, args.out, args.budget_s)
        log(f"{n} records -> {args.out}")
    else:
        Xs, Ys, nrec = [], [], 0
        for r in args.records:
            with np.load(r, allow_pickle=False) as z:
                rec = {k: z[k] for k in ("feat", "kind", "value", "ms", "gid")}
            nrec += len(rec["ms"])
            X, Y = dataset(rec)
            Xs.append(X)
            Ys.append(Y)
        X, Y = np.concatenate(Xs), np.concatenate(Ys)
        m, norm, loss = train(X, Y, args.epochs, args.lr)
        save_model(args.out, m.params, norm, meta=f"records={nrec} rows={len(X)} loss={loss:.4f}")
        log(f"trained on {len(X)} rows from {nrec} records, final loss {loss:.4f} -> {args.out}")


if __name__ == "__main__":
    main()
