"""Learned warm start for FIRE: data, training and evaluation drivers.

The network and its kernels live in the library (svm_mlp_*, svm_warm_start);
this module drives them:

  * ``fire_batch`` builds a batch of spheres from seeded uniform random cell
    points (GPU Delaunay), with the FIRE parameters of the reference's
    minimisation driver (minimization.cpp:111-114: dt 0.001, alpha 0.99,
    dt_max 50 dt, dt_inc 1.1, dt_dec 0.95, alpha_dec 0.9, n_min 4,
    alpha_min 0) and the rms-force stop test.
  * ``training_set`` minimises held-out spheres with plain FIRE and records,
    per cell, the descriptors of the start and the displacement of the cell
    centre to the minimum, in the cell's tangent frame.
  * ``train`` fits the network on the GPU (Adam, svm_mlp_train).
  * ``run_fire`` minimises a batch, optionally after the warm start, and
    reports wall time and per-sphere iterations.

  * ``run_fire_corrected`` applies the network again every few thousand
    FIRE iterations to the spheres not yet converged; ``snapshot_training_set``
    trains it for that use (states along FIRE trajectories -> displacement
    to the end of the minimisation); ``ablation`` measures every variant
    against plain FIRE.

    python -m curvedspacevertexmodel_amd.warmstart train [--out FILE]
    python -m curvedspacevertexmodel_amd.warmstart train-corrector [--out FILE]
"""
import argparse
import os
import time

import numpy as np

import curvedspacevertexmodel_amd as svm

HERE = os.path.dirname(os.path.abspath(__file__))
WEIGHTS = os.path.join(HERE, "warmstart_weights.npz")
CORRECTOR_WEIGHTS = os.path.join(HERE, "corrector_weights.npz")

FIRE_DT = 0.001
FIRE = dict(dt=FIRE_DT, alpha_start=0.99, dt_max=50 * FIRE_DT, dt_inc=1.1, dt_dec=0.95, alpha_dec=0.9, n_min=4,
            alpha_min=0.0)
TRAIN_SEED0 = 900001    # training spheres use seeds disjoint from the evaluation batch (1137 + i)


def batch_params(n_sys, seed=13377, nmin=500, nmax=5000, p0_range=(3.6, 3.9)):
    """Per-sphere (N, p0): N log-uniform in [nmin, nmax], p0 uniform, one Philox generator."""
    g = np.random.Generator(np.random.Philox(seed))
    out = []
    for _ in range(n_sys):
        u = g.random(2)
        n = int(round(np.exp(np.log(nmin) + u[0] * (np.log(nmax) - np.log(nmin)))))
        out.append((n, p0_range[0] + (p0_range[1] - p0_range[0]) * u[1]))
    return out


def fire_batch(params, seed0=1137, cutoff=1e-10, max_iterations=20000):
    systems = [dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=seed0 + i)
               for i, (n, p0) in enumerate(params)]
    ctx = svm.Context(systems)
    for i, (n, p0) in enumerate(params):
        pts = svm.generate_cell_points(n, seed0 + i)
        ctx.build_from_cell_points(i, pts[svm.spatial_order(pts)])
        ctx.fire_set_parameters(i, max_iterations=max_iterations, force_cutoff=cutoff, stop_rms=1, **FIRE)
    return ctx


def run_fire(ctx, n_sys, net=None, scale=1.0, check_every=100):
    """Minimise every sphere; returns wall seconds (warm start included) and
    per-sphere (iterations, converged, rms force)."""
    ctx.synchronize()
    t0 = time.perf_counter()
    if net is not None:
        ctx.warm_start(net, scale)   # undone per sphere where a cell would fold
    ctx.fire_minimize(check_every)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    st = [ctx.fire_state(i) for i in range(n_sys)]
    its = np.array([s["iterations"] for s in st])
    conv = np.array([bool(s["converged"]) for s in st])
    fmax = np.array([s["force_max"] for s in st])
    return wall, its, conv, fmax


def training_set(n_sys=48, seed0=TRAIN_SEED0, cutoff=1e-8, max_iterations=20000):
    """Descriptors X [cells][16] at random starts and target displacements
    T [cells][2] of the cell centres to the FIRE minimum, in the start frames."""
    params = batch_params(n_sys, seed=seed0)
    ctx = fire_batch(params, seed0=seed0, cutoff=cutoff, max_iterations=max_iterations)
    X, T, frames, C0 = [], [], [], []
    for i in range(n_sys):
        x, f = ctx.cell_descriptors(i)
        X.append(x)
        frames.append(f)
        C0.append(ctx.cell_positions(i))
    run_fire(ctx, n_sys)
    ctx.compute_geometry()
    for i in range(n_sys):
        d = ctx.cell_positions(i) - C0[i]
        f = frames[i].astype(np.float64)
        T.append(np.stack([np.sum(d * f[:, :3], 1), np.sum(d * f[:, 3:], 1)], 1).astype(np.float32))
    return np.concatenate(X), np.concatenate(T)


def fire_states(ctx, n_sys):
    """Per-sphere (iterations, converged, rms force) arrays."""
    st = [ctx.fire_state(i) for i in range(n_sys)]
    return (np.array([s["iterations"] for s in st]), np.array([bool(s["converged"]) for s in st]),
            np.array([s["force_max"] for s in st]))


def run_fire_corrected(ctx, n_sys, net, scale=1.0, every=2000, cap=10000, check_every=100):
    """FIRE with the network as a corrector: the warm start before FIRE and
    again every `every` iterations on the spheres that have not converged
    (converged spheres are never moved), up to `cap` iterations per sphere.
    Returns wall seconds and per-sphere (iterations, converged, rms force)."""
    ctx.synchronize()
    t0 = time.perf_counter()
    ctx.warm_start(net, scale)
    done = 0
    while done < cap:
        done = min(cap, done + every)
        for i in range(n_sys):
            ctx.fire_set_max_iterations(i, done)
        ctx.fire_minimize(check_every)
        its, conv, _ = fire_states(ctx, n_sys)
        if done >= cap or conv.all():
            break
        ctx.warm_start(net, scale)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    return (wall,) + fire_states(ctx, n_sys)


def snapshot_training_set(n_sys=48, seed0=TRAIN_SEED0, every=1000, cutoff=1e-8, max_iterations=20000):
    """Descriptors X [rows][16] of the cells of spheres at the start and every
    `every` FIRE iterations after it (while the sphere has not converged), and
    targets T [rows][2]: the displacement of the cell centre from that state
    to the end of the minimisation, in the state's tangent frames."""
    params = batch_params(n_sys, seed=seed0)
    ctx = fire_batch(params, seed0=seed0, cutoff=cutoff, max_iterations=every)
    snaps = []   # (sphere, X, frames, centres)
    done = 0
    while True:
        _, conv, _ = fire_states(ctx, n_sys) if done else (None, np.zeros(n_sys, bool), None)
        ctx.compute_forces()
        for i in np.flatnonzero(~conv):
            x, f = ctx.cell_descriptors(int(i))
            snaps.append((int(i), x, f, ctx.cell_positions(int(i))))
        if done >= max_iterations or conv.all():
            break
        done += every
        for i in range(n_sys):
            ctx.fire_set_max_iterations(i, done)
        ctx.fire_minimize(100)
    ctx.compute_geometry()
    final = [ctx.cell_positions(i) for i in range(n_sys)]
    X, T = [], []
    for i, x, f, c0 in snaps:
        d = final[i] - c0
        f = f.astype(np.float64)
        X.append(x)
        T.append(np.stack([np.sum(d * f[:, :3], 1), np.sum(d * f[:, 3:], 1)], 1).astype(np.float32))
    X, T = np.concatenate(X), np.concatenate(T)
    # a sphere whose minimisation ran away (NaN positions, the model's own
    # behaviour from some random starts) contributes no rows
    keep = np.isfinite(X).all(1) & np.isfinite(T).all(1)
    return X[keep], T[keep]


def ablation(n_sys=1024, cap=10000, lanes=3, net=None, scale=1.0, corrector=None, every=2000, subset_runs=True):
    """FIRE on the 1,024-sphere batch (rms stop 1e-10, `cap` iterations)
    plain and with each learned variant: the one-shot warm start (`net`) and,
    if given, the corrector (`corrector` applied every `every` iterations).
    Per variant: wall time with the cap, converged fraction, and the time to
    tolerance of its converging population (the same FIRE run again on only
    the spheres it converged, so no sphere waits at the cap); against plain,
    iterations to tolerance on the spheres both converge and their final
    energies (same basin: relative difference <= 1e-9)."""
    params = batch_params(n_sys)
    variants = [("plain", None)]
    if net is not None:
        variants.append(("warm_start", lambda c: run_fire(c, n_sys, net, scale)))
    if corrector is not None:
        variants.append(("corrector", lambda c: run_fire_corrected(c, n_sys, corrector, scale, every, cap)))
    out, ref = {}, None
    for name, runner in variants:
        ctx = fire_batch(params, seed0=1137, cutoff=1e-10, max_iterations=cap)
        ctx.set_lanes(lanes)
        wall, its, conv, fmax = run_fire(ctx, n_sys) if runner is None else runner(ctx)
        ctx.compute_geometry()
        energy = np.array(ctx.energies()[0])
        del ctx
        rec = {"wall_s_with_cap": wall, "converged_fraction": float(conv.mean()),
               "iterations_median_converged": float(np.median(its[conv])) if conv.any() else None}
        solid = np.array([p0 < 3.75 for _, p0 in params])
        rec["converged_fraction_p0_below_3.75"] = float(conv[solid].mean())
        if subset_runs and conv.any():
            sub = [params[i] for i in np.flatnonzero(conv)]
            seeds = [1137 + int(i) for i in np.flatnonzero(conv)]
            systems = [dict(n_cells=n, a0=1.0, p0=p0, kr=1.0, t1_threshold=0.1, seed=sd)
                       for (n, p0), sd in zip(sub, seeds)]
            c2 = svm.Context(systems)
            for k, ((n, p0), sd) in enumerate(zip(sub, seeds)):
                pts = svm.generate_cell_points(n, sd)
                c2.build_from_cell_points(k, pts[svm.spatial_order(pts)])
                c2.fire_set_parameters(k, max_iterations=cap, force_cutoff=1e-10, stop_rms=1, **FIRE)
            c2.set_lanes(lanes)
            w2, _, conv2, _ = run_fire(c2, len(sub)) if runner is None else runner_subset(name, c2, len(sub),
                                                                                       net, corrector, scale,
                                                                                       every, cap)
            rec["time_to_tolerance_converging_s"] = w2
            rec["converging_spheres"] = len(sub)
            rec["subset_all_converged"] = bool(conv2.all())
            del c2
        if ref is None:
            ref = (its, conv, energy)
        else:
            both = conv & ref[1]
            rec["both_converged"] = int(both.sum())
            if both.any():
                rec["iterations_ratio_median_both"] = float(np.median(ref[0][both] / its[both]))
                rec["iterations_total_both"] = [int(ref[0][both].sum()), int(its[both].sum())]
                rel = np.abs(energy[both] - ref[2][both]) / np.abs(ref[2][both])
                rec["final_energy_rel_diff_median_both"] = float(np.median(rel))
                rec["final_energy_within_1e-9_both"] = float(np.mean(rel <= 1e-9))
            rec["lost_vs_plain"] = int((ref[1] & ~conv).sum())
            rec["gained_vs_plain"] = int((conv & ~ref[1]).sum())
        out[name] = rec
        print(f"ablation {name}: {rec}", flush=True)
    return out


def runner_subset(name, ctx, n, net, corrector, scale, every, cap):
    if name == "warm_start":
        return run_fire(ctx, n, net, scale)
    return run_fire_corrected(ctx, n, corrector, scale, every, cap)


# Offsets of the weight blocks in the parameter vector (the layout documented
# with SVM_MLP_NPARAM in include/svm.h): W1 [64][16], b1 [64], W2 [64][64], b2,
# W3 [64][64], b3, W4 [16][64], b4 [16].
_W1, _W2, _W3, _W4 = 0, 1088, 5248, 9408


def init_parameters(seed):
    """He-style initial weights (normal, variance 2/fan_in; 1/fan_in for the
    linear output layer), biases zero; the padding inputs 12..15 of W1 and the
    padding outputs 2..15 of W4 stay zero."""
    rng = np.random.default_rng(seed)
    th = np.zeros(svm.MLP_NPARAM, np.float32)
    th[_W1:_W1 + 1024].reshape(64, 16)[:, :12] = rng.normal(0, np.sqrt(2.0 / 12), (64, 12))
    th[_W2:_W2 + 4096] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    th[_W3:_W3 + 4096] = rng.normal(0, np.sqrt(2.0 / 64), 4096)
    th[_W4:_W4 + 1024].reshape(16, 64)[:2] = rng.normal(0, np.sqrt(1.0 / 64), (2, 64))
    return th


def train(X, T, epochs=40, batch=4096, lr=1e-3, seed=1):
    mu = X[:, :12].mean(0)
    sg = X[:, :12].std(0) + 1e-6
    scale = float(T.std()) + 1e-12
    net = svm.MLP()
    net.set_parameters(init_parameters(seed), mu, sg, scale)
    losses = net.train(X, T, epochs=epochs, batch=batch, lr=lr, seed=seed)
    return net, losses


def save(net, path=WEIGHTS, **meta):
    th, mu, sg, sc = net.get_parameters()
    np.savez(path, theta=th, mu=mu, sigma=sg, out_scale=np.float32(sc), **meta)


def load(path=WEIGHTS):
    z = np.load(path, allow_pickle=False)
    net = svm.MLP()
    net.set_parameters(z["theta"], z["mu"], z["sigma"], float(z["out_scale"]))
    scale = float(z["warm_scale"]) if "warm_scale" in z else 1.0
    return net, scale


def iteration_profile(n_sys=128, cap=10000, chunk=100, lanes=3):
    """Where plain FIRE spends its iterations: the first `n_sys` spheres of the
    1,024-sphere batch minimised in chunks of `chunk` iterations, with every
    sphere's rms force read after each chunk.  Returns (thresholds, rows): for
    each sphere that converged, the first chunk end with rms below each
    threshold."""
    params = batch_params(1024)[:n_sys]
    ctx = fire_batch(params, cutoff=1e-10, max_iterations=cap)
    ctx.set_lanes(lanes)
    hist = []
    for done in range(chunk, cap + 1, chunk):
        for i in range(n_sys):
            ctx.fire_set_max_iterations(i, done)
        ctx.fire_minimize(chunk)
        _, conv, rms = fire_states(ctx, n_sys)
        hist.append(rms.copy())
        if conv.all():
            break
    H = np.array(hist)
    _, conv, _ = fire_states(ctx, n_sys)
    thr = [1e-1, 1e-2, 1e-3, 1e-4, 1e-5, 1e-6, 1e-8, 1e-10]
    rows = [[chunk * (int(np.argmax(H[:, s] < t)) + 1) for t in thr] for s in np.where(conv)[0]]
    return thr, np.array(rows, dtype=np.int64).reshape(-1, len(thr))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train", "train-corrector", "profile"])
    ap.add_argument("--out", default=WEIGHTS)
    ap.add_argument("--spheres", type=int, default=48)
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--validate", type=int, default=32, help="held-out spheres used to pick the step scale")
    args = ap.parse_args()
    if args.cmd == "profile":
        thr, rows = iteration_profile()
        print(f"plain FIRE, 128 spheres of the 1,024-sphere batch, {len(rows)} converged to rms < 1e-10")
        for k, t in enumerate(thr):
            print(f"rms < {t:g}: median iteration {np.median(rows[:, k]):.0f}, "
                  f"median fraction of the sphere's iterations {np.median(rows[:, k] / rows[:, -1]):.2f}")
        return
    if args.cmd == "train-corrector":
        t0 = time.perf_counter()
        X, T = snapshot_training_set(args.spheres)
        print(f"snapshot training set: {len(X)} rows from {args.spheres} spheres, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        net, losses = train(X, T, epochs=args.epochs)
        print("loss per epoch:", " ".join(f"{x:.4g}" for x in losses), flush=True)
        out = args.out if args.out != WEIGHTS else CORRECTOR_WEIGHTS
        save(net, out, warm_scale=np.float32(1.0), train_losses=losses, train_cells=np.int64(len(X)))
        print(f"saved {out}")
        return
    t0 = time.perf_counter()
    X, T = training_set(args.spheres)
    print(f"training set: {len(X)} cells from {args.spheres} spheres, {time.perf_counter() - t0:.1f} s", flush=True)
    net, losses = train(X, T, epochs=args.epochs)
    print("loss per epoch:", " ".join(f"{x:.4g}" for x in losses), flush=True)
    # pick the warm-start step scale on validation spheres (seeds disjoint from both sets)
    vparams = batch_params(args.validate, seed=TRAIN_SEED0 + 7919)
    best = None
    for scale in (0.0, 0.5, 1.0):
        ctx = fire_batch(vparams, seed0=TRAIN_SEED0 + 50000)
        wall, its, conv, _ = run_fire(ctx, len(vparams), net if scale > 0 else None, scale)
        print(f"validation scale {scale}: wall {wall:.2f} s, iterations mean {its.mean():.0f} max {its.max()}, "
              f"converged {conv.mean():.2f}", flush=True)
        if scale > 0 and (best is None or wall < best[0]):
            best = (wall, scale)
    save(net, args.out, warm_scale=np.float32(best[1]), train_losses=losses, train_cells=np.int64(len(X)))
    print(f"saved {args.out} (warm-start scale {best[1]})")


if __name__ == "__main__":
    main()
