"""The learned substamp screen: contaminated training pairs, method-labelled
samples, training on the GPU and the model file.

The clip-refit loop of the reference (src/cdkscUtil.cpp:585-684) finds bad
substamps one pass at a time.  The screen (XB_FORM_SCREEN, k_screen.hip) is a
32-64-1 network that, after the loop's first pass, predicts which substamps the
loop will replace later and replaces them at once; the loop still runs and
judges.  Its training data come from the method itself: XB_FORM_RECORD keeps
the 32 features of every substamp each recorded pass kept, and whether the chi2
test under the plain loop's converged fit rejects it (the test the screened
form's re-check applies to every substamp the screen replaces).

Crowded-field pairs follow the learned-screen workload: crowded fields with 20 %
of the selected substamps contaminated by cosmic-ray hits (1-3 px, 5e3-5e4
counts), stars saturated and clipped at 6e4, and variable sources (science
flux x U[0.5, 2]).

    python screen.py --pairs 340 --clean 160 --out screen_model.npy   (needs the GPU)
"""
from __future__ import annotations

import argparse
import json
import pathlib
import sys
import time

import numpy as np

sys.path.insert(0, str(pathlib.Path(__file__).resolve().parent))
import xbach  # noqa: E402

MODEL_PATH = pathlib.Path(__file__).resolve().parent / "screen_model.npy"
NF, NH = xbach.MLP_NFEAT, 64
MU, SD, W1 = 0, 32, 64
B1 = W1 + NH * NF
W2 = B1 + NH
B2 = W2 + NH
TAU = B2 + 1

CONTAM_KINDS = ("cosmic_science", "cosmic_template", "saturated", "variable")


def crowded_cfg(side: int, nstars: int, seed: int, fwhm_t: float, fwhm_s: float):
    """Crowded field: the 512-pair flux law and sky, the given size and PSFs."""
    return xbach.synth_cfg("pair512", w=side, h=side, nstars=nstars, fwhmT=fwhm_t,
                           fwhmS0=fwhm_s, seed=seed)


def substamp_centres(trace: dict) -> list[tuple[int, int]]:
    """Every candidate substamp centre sss kept, both images, without repeats."""
    out = set()
    for im in (0, 1):
        cnt = trace[f"ssCount{im}"]
        xs, ys = trace[f"ssX{im}"], trace[f"ssY{im}"]
        for g in range(len(cnt)):
            for c in range(int(cnt[g])):
                out.add((int(xs[g, c]), int(ys[g, c])))
    return sorted(out)


def contaminate(tmpl, sci, centres, frac: float, seed: int):
    """Copies of the pair with `frac` of the substamp centres contaminated;
    returns (tmpl, sci, [(x, y, kind)])."""
    rng = np.random.default_rng(seed)
    t = tmpl.copy()
    s = sci.copy()
    h, w = t.shape
    pick = rng.permutation(len(centres))[: int(round(frac * len(centres)))]
    done = []
    for k in pick:
        x, y = centres[k]
        kind = CONTAM_KINDS[int(rng.integers(len(CONTAM_KINDS)))]
        if kind.startswith("cosmic"):
            img = s if kind == "cosmic_science" else t
            cx = int(np.clip(x + rng.integers(-8, 9), 1, w - 2))
            cy = int(np.clip(y + rng.integers(-8, 9), 1, h - 2))
            npx = int(rng.integers(1, 4))
            for _ in range(npx):
                img[cy, cx] += np.float32(rng.uniform(5e3, 5e4))
                cx = int(np.clip(cx + rng.integers(-1, 2), 1, w - 2))
                cy = int(np.clip(cy + rng.integers(-1, 2), 1, h - 2))
        else:
            y0, y1, x0, x1 = max(y - 8, 0), min(y + 9, h), max(x - 8, 0), min(x + 9, w)
            win = s[y0:y1, x0:x1].astype(np.float64)
            border = np.concatenate([win[0], win[-1], win[:, 0], win[:, -1]])
            sky = float(np.median(border))
            star = win - sky
            if kind == "variable":
                win = sky + star * rng.uniform(0.5, 2.0)
            else:  # saturated: brighten past 6e4 and clip there
                pk = max(float(star.max()), 1.0)
                win = np.minimum(sky + star * (rng.uniform(1.5, 4.0) * 6e4 / pk), 6e4)
            s[y0:y1, x0:x1] = win.astype(np.float32)
        done.append((x, y, kind))
    return t, s, done


def pair_cfgs(n: int, seed0: int, side=(1536, 3072), density=(2e-3, 3.5e-3)):
    """n crowded-field configurations: side ~ U[side], stars ~ U[density] per
    px, template FWHM ~ U[1.5, 2.5], science FWHM ~ U[template + 0.5, 4.0]."""
    rng = np.random.default_rng(seed0)
    out = []
    for i in range(n):
        sd = int(rng.integers(side[0], side[1] + 1))
        dens = float(rng.uniform(*density))
        fwhm_t = float(rng.uniform(1.5, 2.5))
        fwhm_s = float(rng.uniform(fwhm_t + 0.5, 4.0))
        out.append(crowded_cfg(sd, int(dens * sd * sd), seed0 + 13 * i, fwhm_t, fwhm_s))
    return out


def contaminated_pairs(ctx, params, cfgs, frac: float = 0.2, chunk: int = 16):
    """Yields (tmpl, sci) per configuration, contaminated at `frac` of the
    substamps sss selects on the clean pair (synthesis on host threads);
    frac = 0 yields the clean pairs."""
    for c0 in range(0, len(cfgs), chunk):
        part = cfgs[c0:c0 + chunk]
        clean = xbach.synth_pairs(part, threads=16)
        for cfg, (t, s) in zip(part, clean):
            if frac <= 0.0:
                yield t, s
                continue
            ctx.set_form(xbach.FORM_PLAIN)
            try:
                ctx.init(params, t, s)
                ctx.sss()
            except xbach.XbachError:
                continue
            centres = substamp_centres(ctx.trace().as_dict())
            t2, s2, _ = contaminate(t, s, centres, frac, int(cfg.seed) + 7)
            yield t2, s2


def training_sets(ctx, n_crowded: int, n_clean: int, seed0: int = 900000):
    """(params, pair source) of the training mixture: contaminated crowded
    fields on a 20x20 stamp grid, and clean fields of the batching law and of
    the 2048-pair law (ramping science PSF) on the default 10x10 grid — so the
    network also learns where the plain loop needs no help."""
    p20 = xbach.Params.default(stampsx=20, stampsy=20)
    p10 = xbach.Params.default()
    half = n_clean // 2
    pair2048 = [xbach.synth_cfg("pair2048", seed=seed0 + 500000 + i) for i in range(n_clean - half)]
    return [(p20, contaminated_pairs(ctx, p20, pair_cfgs(n_crowded, seed0))),
            (p10, contaminated_pairs(ctx, p10, xbach.batch_cfgs(half, seed=seed0 // 1000 + 1),
                                     frac=0.0)),
            (p10, contaminated_pairs(ctx, p10, pair2048, frac=0.0))]


def label_samples(f, lab, passes, target="final"):
    """Samples and targets of one recorded run.  The recording holds, pass by
    pass, every kept stamp's label (2 replaced by that pass; of the ones the
    pass kept: 1 rejected by the chi2 test under the converged fit, 3 replaced
    later but passing that test, 0 used by the final fit; -1 no substamp).
    target "next": a substamp a pass kept is positive when the next pass
    replaces it (label 2 there), so screening it does that pass's work one
    solve early; the last recorded pass has no next pass unless it is the
    loop's last (which replaces nothing).  target "final": positive when the
    converged fit's chi2 test rejects it."""
    nrec = min(passes, 16)
    n = len(lab) // max(nrec, 1)
    L = lab.reshape(nrec, n)
    F = f.reshape(nrec, n, -1)
    kept = (L == 0) | (L == 1) | (L == 3)
    if target == "final":
        return F[kept], (L[kept] == 1).astype(np.float32)
    y = np.zeros_like(L, dtype=np.float32)
    y[:-1] = (L[1:] == 2)
    if nrec < passes:  # the pass after the last recorded one is unknown
        kept[-1] = False
    return F[kept], y[kept]


def collect(ctx, params, pairs, target="final") -> tuple[np.ndarray, np.ndarray, list[int]]:
    """Method-labelled samples: run the plain method with XB_FORM_RECORD and
    keep, for every recorded pass, the substamps that pass kept, labelled 1
    when the converged fit does not use them and its chi2 test rejects them
    (label 3, replaced but passing that test, and 0, used by the final fit, are
    the negatives)."""
    Xs, ys, its = [], [], []
    try:
        for t, s in pairs:
            ctx.set_form(xbach.FORM_RECORD)  # (the pair source may share ctx)
            try:
                ctx.run(params, t, s)
            except xbach.XbachError:
                continue
            f, lab, _, _ = ctx.screen_samples()
            if len(its) % 20 == 19:
                print(f"  {len(its) + 1} pairs recorded", flush=True)
            it = ctx.trace().kscIterations
            X1, y1 = label_samples(f, lab, it, target)
            Xs.append(X1)
            ys.append(y1)
            its.append(it)
    finally:
        ctx.set_form(xbach.FORM_PLAIN)
    return np.concatenate(Xs), np.concatenate(ys), its


def class_weights(y: np.ndarray) -> np.ndarray:
    """Balance the two classes (replaced substamps are the rarer one)."""
    pos = max(float(y.sum()), 1.0)
    neg = max(float(len(y) - y.sum()), 1.0)
    return np.where(y > 0.5, 0.5 * len(y) / pos, 0.5 * len(y) / neg).astype(np.float32)


def init_model(X: np.ndarray, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    m = np.zeros(xbach.MLP_NPARAM, np.float32)
    m[MU:SD] = X.mean(axis=0)
    m[SD:W1] = np.maximum(X.std(axis=0), 1e-3)
    m[W1:B1] = rng.normal(0.0, np.sqrt(2.0 / NF), NH * NF)
    m[W2:B2] = rng.normal(0.0, np.sqrt(1.0 / NH), NH)
    m[TAU] = 0.5
    return m


def scores(prob: np.ndarray, y: np.ndarray, tau: float) -> dict:
    pred = prob > tau
    tp = int(np.sum(pred & (y > 0.5)))
    fp = int(np.sum(pred & (y < 0.5)))
    fn = int(np.sum(~pred & (y > 0.5)))
    return {"tau": tau, "precision": tp / max(tp + fp, 1), "recall": tp / max(tp + fn, 1),
            "screened_frac": float(pred.mean()), "tp": tp, "fp": fp, "fn": fn}


def pick_threshold(prob: np.ndarray, y: np.ndarray, min_precision: float = 0.9) -> float:
    """Lowest threshold whose precision on held-out samples reaches
    min_precision (a wrongly screened substamp costs a good fit star)."""
    for tau in np.linspace(0.5, 0.99, 50):
        sc = scores(prob, y, float(tau))
        if sc["tp"] > 0 and sc["precision"] >= min_precision:
            return float(tau)
    return 0.99


def train(ctx, X, y, seed=1, epochs=60, batch=1024, lr=3e-3):
    model = init_model(X, seed)
    wt = class_weights(y)
    return ctx.mlp_train(model, X, y, wt, epochs=epochs, batch=batch, lr=lr, seed=seed)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--pairs", type=int, default=340, help="contaminated crowded pairs")
    ap.add_argument("--clean", type=int, default=160, help="clean pairs")
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--out", default=str(MODEL_PATH))
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--min-precision", type=float, default=0.9)
    ap.add_argument("--target", default="final", choices=["next", "final"])
    args = ap.parse_args()
    ctx = xbach.Context(0)
    t0 = time.time()
    Xs, ys, its = [], [], []
    for params, src in training_sets(ctx, args.pairs, args.clean):
        X1, y1, it1 = collect(ctx, params, src, args.target)
        Xs.append(X1)
        ys.append(y1)
        its += it1
        print(f"set: {len(it1)} pairs, {len(y1)} samples, replaced {y1.mean():.3f}, "
              f"passes mean {np.mean(it1):.2f}", flush=True)
    X, y = np.concatenate(Xs), np.concatenate(ys)
    print(f"pairs: {len(its)}, samples: {len(y)}, replaced {int(y.sum())} ({y.mean():.3f}); "
          f"clip-refit passes mean {np.mean(its):.2f} ({time.time() - t0:.1f} s)", flush=True)
    rng = np.random.default_rng(args.seed)
    perm = rng.permutation(len(y))
    nval = len(y) // 5
    va, tr = perm[:nval], perm[nval:]
    t1 = time.time()
    model, loss = train(ctx, X[tr], y[tr], seed=args.seed, epochs=args.epochs)
    print(f"trained {args.epochs} epochs in {time.time() - t1:.2f} s; loss "
          f"{loss[0]:.4f} -> {loss[-1]:.4f}", flush=True)
    pv = ctx.mlp_forward(model, X[va])
    model[TAU] = pick_threshold(pv, y[va], args.min_precision)
    for tau in (0.5, 0.6, 0.7, 0.8, 0.9):
        print(json.dumps(scores(pv, y[va], tau)), flush=True)
    report = {"samples": int(len(y)), "train": int(len(tr)), "n_validation": int(len(va)),
              "positive_frac": float(y.mean()), "loss_first": float(loss[0]),
              "loss_last": float(loss[-1]), "validation": scores(pv, y[va], float(model[TAU])),
              "clip_refit_passes_mean": float(np.mean(its))}
    np.save(args.out, model)
    print(json.dumps(report), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
