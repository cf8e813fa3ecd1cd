"""Benchmark: He-LDA SCF seconds per solve on one MI355X, the H-apply against the HBM roofline, and
the CPU baseline beside it.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--skip-cpu] [--skip-mixed]

A step is one complete He-LDA (VWN) self-consistent solve on the 16-bohr box at h = 0.25 (65^3,
block k = 5, Chebyshev degree 20, periodic Pulay), geometry build included, float64. Prints ONE
JSON line. ``roofline`` is for the dominant kernel of that solve, the fused Chebyshev step (K2):
its average launch duration is measured live with HIP events around every filter call over the
timed solves, and its algorithmic bytes per launch are
``(2 kNs + gG) / 20 + 19 (3 kNs + gG) / 20`` (s = g = 8 bytes; G = the three face-weight arrays,
``invw`` and ``V``). ``h_apply_microbench`` is the plain apply (K1) on a seeded U(-1,1) block of
k = 8 vectors on the same grid, rotating over enough buffers to defeat the 256 MB Infinity Cache.
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Bytes between L2 and memory per launch of the fused step on this workload, from the rocprofv3
# TCC_EA0_RDREQ_{32B,64B,128B} / TCC_EA0_WRREQ{,_64B} request counts of `bench.py --kernel-only --launch 11:22:6`, all
# L2 channels summed (the 1 GiB copy of the same run reads and writes exactly 1 GiB by the same
# formula; summary and commands: profiles/hdiv_fused_step_he65_k5.md). Re-measured whenever the
# stencil kernel changes.
FUSED_STEP_TRAFFIC_BYTES = (464360 * 128 + 2900 * 64) + (175365 * 64 + (183310 - 175365) * 32)
METRIC = "He-LDA SCF s/solve & H-apply GB/s vs HBM peak; mixed vs FP64 at <=1e-8 Ha"
PUBLISHED_CPU_S_PER_SOLVE = 109.5  # reference, He LDA 65^3, CPU float64 2 threads (convergence_ladders.json)


def he_lda_scenario():
    from types import SimpleNamespace

    return SimpleNamespace(charges=(2.0,), positions=((0.0, 0.0, 0.0),), n_electrons=2.0, xc="lda")


def numerics(spacing=0.25, box=16.0):
    from cdft_amd.contract import Numerics

    return Numerics(box_lengths=(box, box, box), spacing=spacing)


def stencil_bytes(shape, k, s, g, first=False):
    n0, n1, n2 = shape
    n = n0 * n1 * n2
    geom = (n0 + 8) * n1 * n2 + n0 * (n1 + 8) * n2 + n0 * n1 * (n2 + 8) + 2 * n
    return (2 if first else 3) * k * n * s + g * geom


class FilterTimer:
    """HIP events around every Chebyshev filter call (on the stream the kernels run on)."""

    def __init__(self):
        import torch

        self.torch = torch
        self.pairs = []
        self.launches = 0
        self.bytes = 0.0

    def __call__(self, stage, solver, block, degree):
        torch = self.torch
        if stage == "start":
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self.pairs.append([e, None])
        else:
            e = torch.cuda.Event(enable_timing=True)
            e.record()
            self.pairs[-1][1] = e
            k = block.shape[0]
            s = block.element_size()
            shape = solver.grid.shape
            self.launches += degree
            self.bytes += stencil_bytes(shape, k, s, s, first=True) + (degree - 1) * stencil_bytes(shape, k, s, s)

    def seconds(self):
        self.torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in self.pairs) * 1e-3


def h_apply_microbench(pkg, device, geometry_arrays, k=8, reps=200):
    import torch

    geom, shape = geometry_arrays
    n = int(np.prod(shape))
    s = 8
    nbuf = max(2, int(math.ceil(4 * 256e6 / (2 * k * n * s))))
    tuning = autotune_launch(geom, k)
    gen = torch.Generator().manual_seed(0)
    xs = [(torch.rand((k, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device) for _ in range(nbuf)]
    ys = [torch.empty_like(xs[0]) for _ in range(nbuf)]
    for i in range(5):
        pkg.ops.hdiv_apply(xs[i % nbuf], geom, ys[i % nbuf])
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for i in range(reps):
        pkg.ops.hdiv_apply(xs[i % nbuf], geom, ys[i % nbuf])
    e1.record()
    torch.cuda.synchronize()
    t = e0.elapsed_time(e1) * 1e-3 / reps
    b = 2 * k * n * s + s * ((shape[0] + 8) * shape[1] * shape[2] + shape[0] * (shape[1] + 8) * shape[2]
                             + shape[0] * shape[1] * (shape[2] + 8) + 2 * n)
    return {"k": k, "grid": list(shape), "dtype": "f64", "us_per_apply": t * 1e6, "bytes_per_apply": b,
            "GB_s": b / t / 1e9, "frac_of_hbm_peak": b / t / 1e9 / HBM_PEAK_GBS, "buffers_rotated": nbuf,
            "launch": tuning}


def autotune_launch(geom, k):
    """Put the tuned vector-kernel launch for ``(geom.shape, k)`` in force (the sweep a solver runs on a
    shape's fifth solve; a ``--launch`` setting instead, when given) and return it with the launch shape
    it gives."""
    from cdft_amd import autotune, ops

    tuning = autotune.tune_stencil(geom, k)
    shape = ops.hdiv_vec_shape(geom.shape, k, geom.dtype)
    return {"setting": None if tuning is None or tuning["best"] is None else list(tuning["best"]),
            "tile": [shape["tile_y"], shape["tile_z"]], "segment": shape["segment"], "blocks": shape["blocks"],
            "threads": shape["threads"], "candidates_us": None if tuning is None else tuning["us"]}


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads=None, reference=-2.8348195294207272):
    """The oracle (the numpy float64 restatement of the reference's algorithm) running one complete He-LDA
    SCF at 65^3 on this host, timed end to end (geometry, every iteration, the final diagonalisation):
    once with ``threads`` host threads (default: the host's core count, at most 16, the CPU share of one
    GPU box), block applies split over the threads by rows and BLAS/FFT pools limited to the same
    count, and once with 2 threads, the setting of the reference's published 109.5 s."""
    from threadpoolctl import threadpool_limits

    from oracle import core as oc
    from oracle import physics as op

    host = os.cpu_count() or 1
    threads = threads or min(host, 16)
    scen = op.Scenario((2.0,), ((0.0, 0.0, 0.0),), 2.0, "lda")
    num = op.Numerics(box_lengths=(16.0,) * 3, spacing=0.25)
    runs = {}
    for count in (threads, 2):
        with threadpool_limits(limits=count):
            oc.set_threads(count)
            try:
                t0 = time.perf_counter()
                r = op.solve_self_consistent(scen, num)
                runs[count] = (time.perf_counter() - t0, r)
            finally:
                oc.set_threads(1)
    seconds, r = runs[threads]
    two, r2 = runs[2]
    return {"value": seconds, "unit": "s/solve", "cores": threads, "kind": "port",
            "sample": (f"one complete oracle He-LDA SCF at 65^3 on this host ({cpu_model()}, {host} logical CPUs), "
                       f"{threads} threads: {r['scf_iterations']} iterations, E = {float(r['energies']['total']):.13f} Ha "
                       f"(reference {reference}); the reference's published CPU time for this solve is "
                       f"{PUBLISHED_CPU_S_PER_SOLVE} s with 2 threads"),
            "two_threads_s": two, "two_threads_iterations": int(r2["scf_iterations"]),
            "energy_error_vs_reference": float(r["energies"]["total"]) - reference,
            "cpu_model": cpu_model(), "host_cpus": host}


def kernel_only(pkg, device, k=5, filters=3):
    """The dominant kernel alone, for counter runs: ``filters`` degree-20 filters (20 fused steps each) of
    a seeded U(-1,1) block on the bare He 65^3 geometry, then a 1 GiB copy of known size."""
    import torch
    from cdft_amd import scf
    from cdft_amd.chefsi import chebyshev_coefficients
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian

    grid, factor, quad = scf.build_geometry(he_lda_scenario(), numerics(), device)
    geom = CuspFactoredHamiltonian(grid, factor, quad).geometry
    gen = torch.Generator().manual_seed(0)
    x = (torch.rand((k, grid.n_points), generator=gen, dtype=torch.float64) * 2 - 1).to(device)
    table = torch.tensor(chebyshev_coefficients(20, -1.0, 40.0, -1.5), dtype=torch.float64, device=device)
    out, w1, w2 = torch.empty_like(x), torch.empty_like(x), torch.empty_like(x)
    launch = autotune_launch(geom, k)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    pkg.ops.cheb_filter(x, geom, table, 20, out=out, work=(w1, w2))
    torch.cuda.synchronize()
    e0.record()
    for _ in range(filters):
        pkg.ops.cheb_filter(x, geom, table, 20, out=out, work=(w1, w2))
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / (20 * filters)
    a = torch.ones(2**30 // 8, dtype=torch.float64, device=device)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    n = grid.n_points
    bytes_step = (19 * stencil_bytes(grid.shape, k, 8, 8) + stencil_bytes(grid.shape, k, 8, 8, first=True)) / 20
    return {"kernel_only": True, "grid": list(grid.shape), "k": k, "us_per_step": us, "algorithmic_bytes_per_step": bytes_step,
            "GB_s": bytes_step / us / 1e3, "copy_bytes": 2**30, "points": n, "launch": launch,
            "timed_dispatches": 20 * filters}


def solves_only(pkg, device, n, spacing=0.25, mixed=False):
    """One warm-up He-LDA solve on the 16 bohr box at ``spacing`` (65^3 at 0.25, 81^3 at 0.2; float64,
    or with ``mixed`` the float32 filter), a 1 GiB device copy (a marker in a kernel trace), then ``n``
    timed solves: a trace's dispatches after the copy are the timed solves' alone."""
    import torch

    dtype = torch.float32 if mixed else torch.float64
    run_solve(pkg, device, dtype, spacing=spacing)
    a = torch.ones(2**30 // 8, dtype=torch.float64, device=device)
    b = torch.empty_like(a)
    b.copy_(a)
    del a, b
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(n):
        r = run_solve(pkg, device, dtype, spacing=spacing)
    torch.cuda.synchronize()
    shape = tuple(r["step"].grid.shape)
    return {"solves_only": n, "spacing": spacing, "grid": list(shape), "mixed": mixed,
            "s_per_solve": (time.perf_counter() - t0) / n, "energy_total": r["energies"]["total"],
            "scf_iterations": r["scf_iterations"], "filter_precisions": {
                p: r["filter_precisions"].count(p) for p in ("f32", "f64")},
            "launch": pkg.ops.hdiv_vec_shape(shape, 5, torch.float64)}


def apply_only(pkg, device, k=8, reps=60):
    """The k = 8 apply microbench alone (cache-defeating rotation of input and output blocks) on the bare
    He 65^3 geometry, for counter runs, then a 1 GiB copy of known size."""
    import torch
    from cdft_amd import scf
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian

    grid, factor, quad = scf.build_geometry(he_lda_scenario(), numerics(), device)
    geom = CuspFactoredHamiltonian(grid, factor, quad).geometry
    result = h_apply_microbench(pkg, device, (geom, grid.shape), k=k, reps=reps)
    a = torch.ones(2**30 // 8, dtype=torch.float64, device=device)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    return {"apply_only": True, **result, "copy_bytes": 2**30}


def compare_precisions(device, scenario, numerics_, anchor=None, reference=None):
    """One float64 and one mixed (float32 filter, float64 refinement) solve of a configuration, each
    after an untimed warm-up solve of the same precision: seconds, energies and the mixed error."""
    import torch
    from cdft_amd import scf

    out = {}
    for label, dtype in (("f64", torch.float64), ("mixed", torch.float32)):
        scf.solve_self_consistent(scenario, numerics_, device, anchor=anchor, filter_dtype=dtype)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = scf.solve_self_consistent(scenario, numerics_, device, anchor=anchor, filter_dtype=dtype)
        torch.cuda.synchronize()
        out[label] = {"s_per_solve": time.perf_counter() - t0, "energy_total": r["energies"]["total"],
                      "scf_iterations": r["scf_iterations"],
                      "f32_filter_passes": r["filter_precisions"].count("f32"),
                      "f64_filter_passes": r["filter_precisions"].count("f64")}
    out["speedup_mixed_vs_f64"] = out["f64"]["s_per_solve"] / out["mixed"]["s_per_solve"]
    out["energy_error_mixed_vs_f64"] = out["mixed"]["energy_total"] - out["f64"]["energy_total"]
    if reference is not None:
        out["energy_reference_cpu"] = reference
        out["energy_error_f64_vs_reference"] = out["f64"]["energy_total"] - reference
    return out


def adjoint_inversion_step(device, spacing=0.2, seed=0):
    """One Kohn-Sham inversion step on He LDA at h = 0.20 (81^3): V_0 is the converged LDA effective
    potential; the target density is that of V_0 + 8 seeded Gaussian bumps (amplitude U(-0.05, 0.05)
    Ha, width 1 bohr, centres uniform in the 2 bohr ball); L = sum omega f^2 (rho - rho_t)^2; dL/dV on
    the grid by the Sternheimer backward, checked by central finite differences along a further seeded
    bump field (step 1e-6, and 1e-5 to 1e-3 as a sweep) and pointwise at seeded grid points (1e-6 and
    1e-3)."""
    import torch
    from cdft_amd import scf
    from cdft_amd.adjoint import KohnShamDensity, KohnShamProblem, density_loss

    r = scf.solve_self_consistent(he_lda_scenario(), numerics(spacing=spacing), device)
    step = r["step"]
    ham = step.hamiltonian({"effective": r["effective_potential"]})
    grid = step.grid
    pts = grid_points(grid)

    def bumps(s, amplitude):
        rng = np.random.default_rng(s)
        a = rng.uniform(-amplitude, amplitude, 8)
        d = rng.normal(size=(8, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        c = d * (2.0 * rng.uniform(0, 1, 8) ** (1.0 / 3.0))[:, None]
        out = np.zeros(pts.shape[0])
        for j in range(8):
            out += a[j] * np.exp(-((pts - c[j]) ** 2).sum(-1) / 2.0)
        return torch.as_tensor(out).to(device)

    weights = step.q.mass_weights * step.f2
    problem = KohnShamProblem(grid, ham.geometry, ham.measure.q, [2.0])
    rho_t, _ = problem.forward(bumps(seed, 0.05))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    v = torch.zeros(grid.n_points, dtype=torch.float64, device=device, requires_grad=True)
    loss = density_loss(KohnShamDensity.apply(v, problem), rho_t, weights)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    loss.backward()
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    d = bumps(seed + 1, 1.0)
    an = float((v.grad * d).sum())
    cg = problem.history[-1]["sternheimer"][0]

    def loss_at(field):
        return float(density_loss(problem.forward(field)[0], rho_t, weights))

    # directional central differences at the configuration's step 1e-6 and a sweep of larger steps
    directional = {}
    for h in (1e-6, 1e-5, 1e-4, 1e-3):
        fd = (loss_at(h * d) - loss_at(-h * d)) / (2 * h)
        directional[f"{h:g}"] = {"fd": fd, "relative_error": abs(fd - an) / max(abs(an), 1e-300)}
    # pointwise: V perturbed at one grid point, seeded points within 2.5 bohr of the nucleus
    grad = v.grad.detach().cpu().numpy()
    inner = np.nonzero(np.linalg.norm(pts, axis=1) < 2.5)[0]
    picks = np.random.default_rng(seed + 7).choice(inner, size=64, replace=False)
    gmax = float(np.abs(grad).max())
    pointwise = {}
    for hp, count in ((1e-6, 16), (1e-3, 64)):
        worst = 0.0
        for p in picks[:count]:
            e = torch.zeros(grid.n_points, dtype=torch.float64, device=device)
            e[int(p)] = 1.0
            fdp = (loss_at(hp * e) - loss_at(-hp * e)) / (2 * hp)
            worst = max(worst, abs(fdp - float(grad[p])))
        pointwise[f"{hp:g}"] = {"points": int(count), "max_abs_error": worst, "max_error_over_max_gradient": worst / gmax}
    return {"workload": "He LDA 81^3 (h=0.20, 16 bohr box): dL/dV of L = integral (n - n_t)^2 by the Sternheimer backward",
            "loss": float(loss.detach()), "forward_s": t1 - t0, "backward_s": t2 - t1,
            "cg_iterations": cg["iterations"], "cg_relative_residual": cg["relative_residual"],
            "cg_converged": bool(cg["converged"]), "cg_iteration_cap": problem.cg.max_iterations,
            "cg_tolerance": problem.cg.tol, "analytic_directional": an,
            "fd_directional": directional["1e-06"]["fd"], "fd_relative_error": directional["1e-06"]["relative_error"],
            "fd_step": 1e-6, "fd_directional_sweep": directional,
            "fd_pointwise": {"selection": "seeded (seed + 7) among grid points within 2.5 bohr", "max_abs_gradient": gmax,
                             "by_step": pointwise}}


def batched_scan(device, count=64, seed=0, spacing=0.25):
    """The H2+ bond-length scan: ``count`` seeded bond lengths in [0.8, 6.0] bohr, box (R + 14, 14, 14)
    at h = 0.25 (n0 from 60 to 81, 57 x 57 across), one electron, float64. The batch runs as one
    lockstep CheFSI (one descriptor launch per filter step for all grids, each iteration one HIP
    graph); beside it the same grids solved one after another with the single-grid solver."""
    import torch
    from cdft_amd.batch import BatchedCheFSI, bond_scan_geometries
    from cdft_amd.contract import Numerics
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry, solve_noninteracting

    scan = bond_scan_geometries(count=count, seed=seed, spacing=spacing)
    t0 = time.perf_counter()
    geometries, hams, nums = [], [], []
    for r, sc, box in scan:
        num = Numerics(box_lengths=box, spacing=spacing)
        g = build_geometry(sc, num, device)
        geometries.append(g)
        hams.append(CuspFactoredHamiltonian(*g))
        nums.append(num)
    torch.cuda.synchronize()
    build = time.perf_counter() - t0
    solver = BatchedCheFSI(hams, nums[0].eigen, n_states=1)
    solver.solve()  # warm-up (graph capture)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    solver = BatchedCheFSI(hams, nums[0].eigen, n_states=1)
    batched = solver.solve()
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t1
    t2 = time.perf_counter()
    solo = [solve_noninteracting(sc, num, device, geometry=g) for (r, sc, box), num, g in zip(scan, nums, geometries)]
    torch.cuda.synchronize()
    t_solo = time.perf_counter() - t2
    diffs = [abs(b["eigenvalues"][0] - s["eigenvalues"][0]) for b, s in zip(batched, solo)]
    # the same batch with float32 filters while every active grid's residual is above 1e-5
    BatchedCheFSI(hams, nums[0].eigen, n_states=1, filter_dtype=torch.float32).solve()  # warm-up (capture)
    torch.cuda.synchronize()
    t3 = time.perf_counter()
    msolver = BatchedCheFSI(hams, nums[0].eigen, n_states=1, filter_dtype=torch.float32)
    mixed = msolver.solve()
    torch.cuda.synchronize()
    t_mixed = time.perf_counter() - t3
    mixed_diff = max(abs(m["eigenvalues"][0] - b["eigenvalues"][0]) for m, b in zip(mixed, batched))
    shapes = [h.grid.shape for h in hams]
    return {"workload": f"{count} bare H2+ geometries, R = 0.8 + 5.2 U(0,1) (seed {seed}), box (R+14, 14, 14) at h={spacing}, "
                        "float64, CheFSI k=5 degree 20 from the reference's seeded start",
            "n0_range": [min(s[0] for s in shapes), max(s[0] for s in shapes)], "n1": shapes[0][1],
            "geometry_build_s": build, "batched_solve_s": t_batch, "sequential_solve_s": t_solo,
            "speedup_batched_vs_sequential": t_solo / t_batch, "solves_per_s_batched": count / t_batch,
            "lockstep_iterations": solver.iterations, "blocks_per_filter_launch": solver.blocks_per_launch,
            "batched_phases_s": solver.timings,
            "autotuned_launch": solver.tuning,
            "sequential_iterations_total": int(sum(s["n_iterations"] for s in solo)),
            "graph": solver.graph_record,
            "max_abs_eigenvalue_diff_vs_sequential": max(diffs),
            "iterations_equal": all(b["n_iterations"] == s["n_iterations"] for b, s in zip(batched, solo)),
            "eigenvalue_range": [float(min(b["eigenvalues"][0] for b in batched)),
                                 float(max(b["eigenvalues"][0] for b in batched))],
            "mixed": {"batched_solve_s": t_mixed, "speedup_vs_float64_batch": t_batch / t_mixed,
                      "lockstep_iterations": msolver.iterations,
                      "f32_passes": msolver.filter_precisions.count("f32"),
                      "f64_passes": msolver.filter_precisions.count("f64"),
                      "all_converged": all(m["converged"] for m in mixed),
                      "max_abs_eigenvalue_diff_vs_float64_batch": mixed_diff}}


def gate_suite(device):
    """The re-solve gates (G3.1, G3.2, G3.5, G3.4, G5.2) on the two bare molecules and the He atom
    (LDA) at their derived production grids: per scenario the production solve and each gate's wall
    time, verdicts and measured values, for three passes of the whole suite: the first (with the
    per-shape launch tuning and the graph captures), then one with the G3.4 and G5.2 re-solves one
    after another and one with them at once on side streams. Beside it, G3.1's three H2 rungs solved
    one after another instead of as one batch."""
    import torch

    from cdft_amd import fingerprint, gates, scf

    out = {"scenarios": ["h2plus_R2", "h2_R1.4", "he_atom_lda"], "passes": []}
    for label, concurrent in (("first", True), ("sequential", False), ("concurrent", True)):
        record = {}
        t_pass = time.perf_counter()
        for sid in out["scenarios"]:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run = scf.solve_scenario(fingerprint.REGISTRY[sid], device=device)
            t_solve = time.perf_counter() - t0
            results, wall = gates.resolve_gates(run, concurrent=concurrent)
            torch.cuda.synchronize()
            record[sid] = {
                "production_solve_s": t_solve,
                "grid": list(run.result["hamiltonian"].grid.shape if "hamiltonian" in run.result
                             else run.result["step"].grid.shape),
                "gates": {r.gate_id: {"verdict": r.verdict,
                                      "measured": r.measured if math.isfinite(r.measured) else None,
                                      **({"reason": r.reason} if r.reason else {})} for r in results},
                "gate_wall_s": {k: v for k, v in wall.items() if k != "rung_batches"},
                "rung_batches": wall["rung_batches"],
                "scenario_wall_s": time.perf_counter() - t0,
            }
        out["passes"].append({"pass": label, "concurrent_resolves": concurrent, "wall_s": time.perf_counter() - t_pass,
                              "scenarios": record})
    sc = fingerprint.REGISTRY["h2_R1.4"]
    num, _ = fingerprint.derived_numerics(sc)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for h in [num.spacing / (gates.LADDER_RATIO**r) for r in range(gates.LADDER_RUNGS)]:
        scf.solve_scenario(sc, num.with_grid(h, tuple(num.box_lengths)), device, derive_grid=False, n_states=1)
    torch.cuda.synchronize()
    out["h2_G3.1_rungs_one_by_one_s"] = time.perf_counter() - t0
    return out


def grid_points(grid):
    n0, n1, n2 = grid.shape
    o = np.asarray(grid.origin)
    i = np.arange(n0)[:, None, None] * grid.spacing + o[0]
    j = np.arange(n1)[None, :, None] * grid.spacing + o[1]
    k = np.arange(n2)[None, None, :] * grid.spacing + o[2]
    return np.stack(np.broadcast_arrays(i, j, k), axis=-1).reshape(-1, 3)


def run_solve(pkg, device, filter_dtype, timer=None, spacing=0.25):
    import torch
    from cdft_amd import scf

    r = scf.solve_self_consistent(he_lda_scenario(), numerics(spacing=spacing), device, filter_dtype=filter_dtype,
                                  filter_hook=timer)
    torch.cuda.synchronize()
    return r


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-mixed", action="store_true")
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-gates", action="store_true")
    ap.add_argument("--gates-only", action="store_true",
                    help="run only the re-solve gate suite (G3.1, G3.2, G3.5, G3.4, G5.2) on H2+, H2 and He LDA "
                         "and print its record")
    ap.add_argument("--kernel-only", action="store_true",
                    help="run only the fused Chebyshev step on the He 65^3 geometry (k=5, three degree-20 filters) "
                         "and a 1 GiB device copy that calibrates the HBM byte counters; for rocprofv3 --pmc runs")
    ap.add_argument("--batch-only", action="store_true",
                    help="run only the 64-geometry batched H2+ scan and print its record; for rocprofv3 runs")
    ap.add_argument("--apply-only", action="store_true",
                    help="run only the k=8 apply microbench on the He 65^3 geometry and a 1 GiB copy; for rocprofv3 runs")
    ap.add_argument("--solves-only", type=int, default=0, metavar="N",
                    help="one warm-up He 65^3 solve, a 1 GiB device copy (the marker after which a kernel trace "
                         "holds only the timed solves), then N timed solves; for rocprofv3 runs")
    ap.add_argument("--spacing", type=float, default=0.25,
                    help="grid spacing of the --solves-only solves (0.25: 65^3, 0.2: 81^3)")
    ap.add_argument("--mixed", action="store_true", help="--solves-only with the float32 filter")
    ap.add_argument("--launch", default=None, metavar="TY:TZ:SEG",
                    help="vector-kernel launch (tile rows, tile width, planes per segment) for the He 65^3 grid's "
                         "blocks, in place of the timing sweep (--kernel-only, --apply-only, --solves-only), so "
                         "that several profiler runs measure one launch")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the Kohn-Sham path here runs on a "
              "single MI355X; see the sharding paragraph of the design notes)", file=sys.stderr)
        return 2

    import torch

    import __graft_entry__ as ge

    pkg = ge.load_package()
    pkg.load_library()
    device = torch.device("cuda:0")
    if args.launch:
        from cdft_amd import autotune

        autotune.ENABLED = False
        setting = [int(v) for v in args.launch.split(":")]
        for k in (1, 5, 8):
            pkg.ops.hdiv_vec_set((65, 65, 65), k, torch.float64, *setting)
    if args.solves_only:
        print(json.dumps(solves_only(pkg, device, args.solves_only, args.spacing, args.mixed)), flush=True)
        return 0
    if args.kernel_only:
        print(json.dumps(kernel_only(pkg, device)), flush=True)
        return 0
    if args.apply_only:
        print(json.dumps(apply_only(pkg, device)), flush=True)
        return 0
    if args.batch_only:
        print(json.dumps(batched_scan(device)), flush=True)
        return 0
    if args.gates_only:
        print(json.dumps(gate_suite(device)), flush=True)
        return 0
    for _ in range(args.warmup):
        run_solve(pkg, device, torch.float64)
    timer = FilterTimer()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    results = [run_solve(pkg, device, torch.float64, timer) for _ in range(args.steps)]
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    s_per_solve = elapsed / args.steps
    r = results[-1]
    filter_seconds = timer.seconds()
    avg_launch = filter_seconds / max(timer.launches, 1)
    bytes_per_launch = timer.bytes / max(timer.launches, 1)
    achieved = bytes_per_launch / avg_launch / 1e9
    step = r["step"]
    shape = step.grid.shape
    vec = pkg.ops.hdiv_vec_shape(shape, 5, torch.float64)  # the fused step's launch during the timed solves
    micro = h_apply_microbench(pkg, device, (step.bare.geometry, shape))
    line = {
        "metric": METRIC,
        "value": s_per_solve,
        "unit": "s/solve",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": s_per_solve * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": s_per_solve / PUBLISHED_CPU_S_PER_SOLVE,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "He LDA (VWN) SCF, Z=2 at origin, 16 bohr box, h=0.25, 65^3 grid, k=5, Chebyshev "
                               "degree 20, periodic Pulay, float64, geometry build included",
                   "grid": list(shape), "block": 5},
        "energy_total": r["energies"]["total"],
        "hip_graphs": r["graphs"],
        "speculative_filters": r["speculative_filters"],
        "energy_reference_cpu": -2.8348195294207272,
        "scf_iterations": r["scf_iterations"],
        "block_applies": int(sum(r["applications"])),
        "roofline": {"kernel": f"hdiv_vec_kernel<double,2,{vec['tile_z']}> (fused Chebyshev step, k=5, "
                               f"{vec['tile_y']}x{vec['tile_z']} tiles, {vec['segment']}-plane segments)", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": FUSED_STEP_TRAFFIC_BYTES, "avg_launch_us": avg_launch * 1e6, "algorithmic_bytes_per_launch": bytes_per_launch,
                     "launches_timed": timer.launches},
        "h_apply_microbench": micro,
    }
    if not args.skip_mixed:
        run_solve(pkg, device, torch.float32)  # untimed warm-up of the float32 graphs, as for float64
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rm = run_solve(pkg, device, torch.float32)
        torch.cuda.synchronize()
        tm = time.perf_counter() - t0
        line["mixed"] = {"s_per_solve": tm, "speedup_vs_f64": s_per_solve / tm,
                         "energy_total": rm["energies"]["total"],
                         "energy_error_vs_f64": rm["energies"]["total"] - r["energies"]["total"],
                         "scf_iterations": rm["scf_iterations"],
                         "f32_filter_passes": rm["filter_precisions"].count("f32"),
                         "f64_filter_passes": rm["filter_precisions"].count("f64")}
        # The precision target's configuration: He LDA on the 16 bohr box at h = 0.20 (81^3).
        line["he_lda_h020_81"] = compare_precisions(device, he_lda_scenario(), numerics(spacing=0.2),
                                                    reference=-2.834830140859734)
        # H2 LDA, protons at (+-0.7, 0, 0), 20 x 16 x 16 box at h = 0.20 (101 x 81 x 81), lattice
        # anchored on the first proton.
        from types import SimpleNamespace
        from cdft_amd.contract import Numerics
        h2 = SimpleNamespace(charges=(1.0, 1.0), positions=((-0.7, 0.0, 0.0), (0.7, 0.0, 0.0)), n_electrons=2.0, xc="lda")
        line["h2_lda_101x81x81"] = compare_precisions(device, h2, Numerics(box_lengths=(20.0, 16.0, 16.0), spacing=0.2),
                                                      anchor=(-0.7, 0.0, 0.0))
    if not args.skip_adjoint:
        line["adjoint_inversion_step"] = adjoint_inversion_step(device)
    if not args.skip_batch:
        line["h2plus_batched_scan"] = batched_scan(device)
    if not args.skip_gates:
        line["gate_suite"] = gate_suite(device)
    if not args.skip_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
