"""Train the learned multigrid damping network on the GPU and save its weights.

Usage::

    python ad-tlert_amd/train_damping.py [--fields 24] [--steps 400] [--out ad-tlert_amd/data/damping_gnn.npz]

Training data: the ~4k-node graded 2.5D strip mesh (48 electrodes at 2 m) with ``--fields``
seeded random conductivity fields -- half with independent log-uniform cell resistivities
plus Gaussian anomalies (as the benchmark cases), half smooth random fields -- assembled at
random wavenumbers of the reference quadrature.  For every matrix, plain Jacobi PCG runs a
random number of iterations (10-80) from zero on 32 random right-hand sides; the remaining
errors are the training samples.  A quarter of the fields is held out for validation.
Writes the weights with a manifest (data, seeds, losses, timings) and prints it as JSON.
"""

from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--fields", type=int, default=24)
    ap.add_argument("--waves-per-field", type=int, default=2)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--eval-fields", type=int, default=4)
    ap.add_argument("--out", default=str(HERE / "data" / "damping_gnn.npz"))
    args = ap.parse_args()
    os.environ.setdefault("ADTLERT_ENABLE_FLOAT64", "1")

    import numpy as np
    import torch

    adt = importlib.import_module("ad-tlert_amd")
    learned = importlib.import_module("ad-tlert_amd.learned")
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    t0 = time.perf_counter()
    case = adt.cases.strip_dd_4k()
    mesh = adt.Mesh.from_arrays(case.nodes, case.cells)
    survey = adt.Survey.from_arrays(case.electrodes, case.measurements)
    op = adt.ERTForward2p5D.from_mesh_survey(mesh, survey, preconditioner="jacobi")
    dev = op._dev
    host_pattern = (dev.pattern.indptr, dev.pattern.indices)
    rng = np.random.default_rng(args.seed)
    centers = case.nodes[case.cells].mean(axis=1)
    systems = []
    for f in range(args.fields):
        if f % 2 == 0:
            sigma = adt.cases.log_conductivity_2d(centers, rng)
        else:
            field = adt.cases.smooth_fields_2d(centers, rng, 1, length=rng.uniform(3.0, 15.0))[0]
            sigma = 10.0 ** (-(2.0 + 0.6 * field))
        values = op._assemble_device(dev, torch.as_tensor(sigma)[None])
        for w in rng.choice(dev.W, args.waves_per_field, replace=False):
            iters = int(rng.integers(10, 81))
            errors = learned.error_samples(dev.indptr, dev.indices, values[w], dev.N, 32, iters, rng, host_pattern)
            systems.append(learned.prepare_system(*host_pattern, values[w].contiguous(), errors))
    n_val = max(1, len(systems) // 4)
    train_set, val_set = systems[:-n_val], systems[-n_val:]
    log(f"data: {len(train_set)} training / {len(val_set)} validation systems, N={dev.N}, "
        f"{time.perf_counter() - t0:.1f}s")
    net = learned.DampingNetwork(seed=args.seed)

    def validation() -> float:
        with torch.no_grad():
            return float(np.mean([float(learned.two_grid_loss(s, net.omega(s.graph, s.values, s.omega_p)))
                                  for s in val_set]))

    constant = validation()  # the untrained network returns the constant damping
    t1 = time.perf_counter()
    losses = learned.train(net, train_set, steps=args.steps, lr=args.lr, log=log)
    train_s = time.perf_counter() - t1
    trained = validation()
    net.manifest = {
        "mesh": "2.5D graded P1 strip, 48 electrodes at 2 m (strip_dd_4k)", "nodes": int(dev.N),
        "fields": args.fields, "waves_per_field": args.waves_per_field, "seed": args.seed,
        "samples_per_system": 32, "jacobi_iterations": "uniform 10..80", "steps": args.steps, "lr": args.lr,
        "loss": "mean two-grid energy ratio", "validation_loss_constant": constant,
        "validation_loss_trained": trained, "final_training_loss": float(np.mean(losses[-20:])),
        "training_seconds": train_s, "params": learned.PARAM_COUNT,
    }
    # solver comparison on held-out conductivity fields: every wavenumber x source of the strip
    # survey, multigrid-preconditioned FCG with the constant damping vs the trained network
    evaluation = []
    for f in range(args.eval_fields):
        sigma = torch.as_tensor(adt.cases.log_conductivity_2d(centers, rng) if f % 2 == 0 else
                                10.0 ** (-(2.0 + 0.6 * adt.cases.smooth_fields_2d(centers, rng, 1)[0])))
        row = {}
        for name in ("amg", "learned"):
            op._solver.preconditioner = name
            op._solver.damping_network = net
            torch.cuda.synchronize()
            t = time.perf_counter()
            op._solve_fields(sigma[None])
            torch.cuda.synchronize()
            st = op._solver.last
            row[name] = {"mean_iterations": float(st.iterations.mean()), "max_iterations": int(st.iterations.max()),
                         "seconds": time.perf_counter() - t, "fallback_rhs": int(st.fallback_rhs)}
        evaluation.append(row)
        log(f"held-out field {f}: iterations amg {row['amg']['mean_iterations']:.2f} "
            f"(max {row['amg']['max_iterations']}), learned {row['learned']['mean_iterations']:.2f} "
            f"(max {row['learned']['max_iterations']})")
    net.manifest["evaluation"] = evaluation
    net.manifest["mean_iterations_amg"] = float(np.mean([r["amg"]["mean_iterations"] for r in evaluation]))
    net.manifest["mean_iterations_learned"] = float(np.mean([r["learned"]["mean_iterations"] for r in evaluation]))
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    net.save(args.out)
    print(json.dumps(net.manifest), flush=True)
    log(f"validation two-grid loss: constant {constant:.5f} -> trained {trained:.5f}; saved {args.out}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
