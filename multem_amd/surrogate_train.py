"""Trains the frozen-phonon surrogate on pairs made by multem_amd.surrogate_data.

    python -m multem_amd.surrogate_train --data a.npz b.npz --n-eval 24 --steps 6000 --out weights.npz

The last --n-eval pairs (or the last --eval-per-file pairs of each file) are
held out.  Prints one JSON line with the
held-out relative RMS error of the network and of the 2-configuration image
it starts from, both against the 64-configuration image:
rel_rms = sqrt(mean((pred - ref)^2)) / sqrt(mean(ref^2)) per image.
"""
import argparse
import json
import sys
import time

import numpy as np

from .surrogate import SurrogateNet, init_params, normalise


def load_pairs(paths, eval_per_file=0):
    """Training pairs of all files, then held-out pairs: the last
    eval_per_file pairs of each file (none if 0).  Also returns, for the
    held-out pairs, the generation seed of their file and whether the pair
    was tilted (-1 where a file does not record it)."""
    S, V, T, Se, Ve, Te, src, tilt = [], [], [], [], [], [], [], []
    for p in paths:
        d = np.load(p, allow_pickle=False)
        k = len(d["adf_small"]) - eval_per_file
        S.append(d["adf_small"][:k]); V.append(d["vmap"][:k]); T.append(d["adf_big"][:k])
        Se.append(d["adf_small"][k:]); Ve.append(d["vmap"][k:]); Te.append(d["adf_big"][k:])
        n_e = len(d["adf_small"]) - k
        src.append(np.full(n_e, int(d["seed"]) if "seed" in d.files else -1, dtype=np.int64))
        tilt.append((d["theta"][k:] > 0).astype(np.int8) if "theta" in d.files else np.full(n_e, -1, np.int8))
    held = dict(source_seed=np.concatenate(src), tilted=np.concatenate(tilt))
    return (np.concatenate(S + Se), np.concatenate(V + Ve), np.concatenate(T + Te),
            sum(len(x) for x in Se), held)


def rel_rms(pred, ref):
    pred, ref = np.asarray(pred, np.float64), np.asarray(ref, np.float64)
    return np.sqrt(np.mean((pred - ref) ** 2, axis=(-2, -1)) / np.mean(ref ** 2, axis=(-2, -1)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--data", nargs="+", required=True)
    ap.add_argument("--n-eval", type=int, default=24, help="held-out pairs: the last n-eval of all files")
    ap.add_argument("--eval-per-file", type=int, default=0,
                    help="instead, hold out the last k pairs of each file")
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--blocks", type=int, default=6)
    ap.add_argument("--channels", type=int, default=32)
    ap.add_argument("--out", required=True, help="weights (.npz)")
    ap.add_argument("--eval-out", default=None, help="held-out pairs (.npz)")
    a = ap.parse_args()

    S, V, T, n_held, held = load_pairs(a.data, a.eval_per_file)
    n = len(S)
    ntr = n - (n_held if a.eval_per_file else a.n_eval)
    x, s = normalise(S, V)
    t = (T / s).astype(np.float32)
    net = SurrogateNet(a.blocks, a.channels)
    net.set_params(init_params(a.seed, a.blocks, a.channels))
    t0 = time.perf_counter()
    losses = net.train(x[:ntr], t[:ntr], batch=a.batch, steps=a.steps, lr=a.lr, seed=a.seed + 1)
    t_train = time.perf_counter() - t0
    theta = net.get_params()
    pred = net.forward(x[ntr:]).astype(np.float64) * s[ntr:]
    e_net = rel_rms(pred, T[ntr:])
    e_base = rel_rms(S[ntr:], T[ntr:])
    np.savez(a.out, theta=theta, n_blocks=a.blocks, channels=a.channels, losses=losses)
    if a.eval_out:
        extra = held if a.eval_per_file else {}
        np.savez_compressed(a.eval_out, adf_small=S[ntr:], vmap=V[ntr:], adf_big=T[ntr:], **extra)
    print(json.dumps({"pairs_train": int(ntr), "pairs_eval": int(n - ntr), "steps": a.steps, "batch": a.batch,
                      "train_s": t_train, "loss_first": float(losses[:20].mean()),
                      "loss_last": float(losses[-50:].mean()),
                      "eval_rel_rms_net_mean": float(e_net.mean()), "eval_rel_rms_net_max": float(e_net.max()),
                      "eval_rel_rms_2conf_mean": float(e_base.mean()),
                      "eval_rel_rms_2conf_max": float(e_base.max())}))
    sys.stdout.flush()


if __name__ == "__main__":
    main()
