"""Trains the warm-start U-Net on GPU-solved synthetic skies and saves it.

    python -m sparseimaging_amd.train_warm_start --n 4000 --epochs 8 \\
        --out sparseimaging_amd/weights/unet_256.npz

Skies: a ring (random radius, width, centre) plus 1-4 Gaussians at 256^2,
M = 20000 visibilities on 12 random elliptical arcs, 1 % noise, lambda_l1 =
1 and lambda_tsv log-uniform in [1, 1e4]; targets are the GPU MFISTA
solutions at a 1e-6 relative tolerance (12-tap preset)."""
import argparse
import os
import sys
import time


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=4000, help="training skies")
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--M", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=5000)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--solve-batch", type=int, default=1, help="problems per solver launch")
    ap.add_argument("--solve-engine", choices=("persistent", "wide"), default="persistent",
                    help="batch engine of the training solves (--solve-batch > 1)")
    ap.add_argument("--lr-final", type=float, default=None, help="final learning rate (default lr / 10)")
    ap.add_argument("--mse", action="store_true", help="plain MSE instead of the relative squared error")
    ap.add_argument("--in-ch", type=int, default=3, choices=(2, 3),
                    help="input channels: dirty image and log-lambda plane (2), plus the dirty beam (3)")
    ap.add_argument("--init", default=None, help="start from these saved weights (fine-tuning)")
    ap.add_argument("--out", default=None)
    args = ap.parse_args()

    from . import learn

    def log(msg):
        print(msg, flush=True)

    if args.init:
        net = learn.UNet.load(args.init)
        args.in_ch = net.in_ch
    else:
        net = learn.UNet(base=args.base, levels=args.levels, in_ch=args.in_ch, seed=args.seed)
    t0 = time.time()
    from . import ENGINE_PERSISTENT, ENGINE_STREAM

    X, T, lam = learn.make_training_set(args.n, N=args.N, M=args.M, seed=args.seed, batch=args.solve_batch,
                                        log=log, channels=args.in_ch,
                                        engine=ENGINE_STREAM if args.solve_engine == "wide" else ENGINE_PERSISTENT)
    t_data = time.time() - t0
    t1 = time.time()
    hist = learn.train(net, X, T, epochs=args.epochs, batch=args.batch, lr=args.lr,
                       lr_final=args.lr_final if args.lr_final else args.lr / 10, log=log, relative=not args.mse)
    t_train = time.time() - t1
    out = args.out or learn.WEIGHTS
    os.makedirs(os.path.dirname(os.path.abspath(out)), exist_ok=True)
    prev = getattr(net, "meta", None) or {}
    n_total = args.n + (prev.get("n", 0) if args.init else 0)
    ep_total = args.epochs + (prev.get("epochs", 0) if args.init else 0)
    net.save(out, meta={"n": n_total, "N": args.N, "M": args.M, "epochs": ep_total, "seed": args.seed,
                        "fine_tuned_from": os.path.basename(args.init) if args.init else None,
                        "loss_kind": "mse" if args.mse else "relative",
                        "loss": hist, "data_s": round(t_data, 1), "train_s": round(t_train, 1)})
    log(f"saved {out}: data {t_data:.1f} s, training {t_train:.1f} s, loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
