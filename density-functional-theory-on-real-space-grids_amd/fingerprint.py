"""Full-precision fingerprints of the registered Coulomb scenarios, in the reference's JSON layout, and
their comparison (reference ``scripts/fingerprint.py``).

    python __graft_entry__.py fingerprint out.json [--scenarios h_atom he_atom_lda ...]
    python __graft_entry__.py fingerprint --diff reference.json out.json [--tol 1e-10] [--exact]

Each scenario is solved on the device at production settings (``Numerics()``, the grid derived from
the nuclei as the reference derives it, no LOBPCG cross-check) and recorded as the reference records
it: every energy term and eigenvalue as ``repr`` strings, the occupations, SHA-256 digests of the
density and the orbitals, the grid, the iteration and state counts. A file written here and the
reference's own CPU fingerprint (``reports/fingerprint_cpu_reference.json``) compare with ``--diff``,
which fails a scenario when an energy or eigenvalue differs by more than ``--tol`` (1e-10 Ha: two
float64 solves of one discretisation differ by summation order only), an occupation by more than
``--occupation-tol``, or the status or grid differ; digests and iteration counts are reported, not
judged. ``--exact`` fails on any difference at all.

The analytic model systems of the reference's registry (``harmonic_w1``, ``box_L10``) have no nuclei:
they are solved with the plain operator ``-1/2 lap + V`` (``model.py``) and recorded the same way.
"""

from __future__ import annotations

import argparse
import hashlib
import json
import math
import pathlib
import platform
import sys
import time

import numpy as np
import torch

from .contract import Numerics, Scenario

# ------------------------------------------------------------------------------------------------
# The scenarios (reference ``physics_config.py:137-420``) and their derived grids (``grid.py``)
# ------------------------------------------------------------------------------------------------


def _atom(sid, z, n_electrons, xc="none"):
    return Scenario(sid, (float(z),), ((0.0, 0.0, 0.0),), float(n_electrons), xc)


def _diatomic(sid, bond, n_electrons, xc="none"):
    half = 0.5 * bond
    return Scenario(sid, (1.0, 1.0), ((0.0, 0.0, -half), (0.0, 0.0, half)), float(n_electrons), xc)


#: The Coulomb scenarios of the reference's registry, bonds along z as the reference places them.
REGISTRY: dict[str, Scenario] = {s.scenario_id: s for s in (
    _atom("h_atom", 1, 1),
    _atom("he_plus", 2, 1),
    _diatomic("h2plus_R2", 2.0, 1),
    _atom("he_atom", 2, 2),
    _diatomic("h2_R1.4", 1.4, 2),
    _atom("h_atom_lda", 1, 1, "lda"),
    _atom("he_plus_lda", 2, 1, "lda"),
    _atom("he_atom_lda", 2, 2, "lda"),
    _atom("he_atom_pbe", 2, 2, "pbe"),
    _diatomic("h2_R1.4_lda", 1.4, 2, "lda"),
    _diatomic("h2_R1.4_pbe", 1.4, 2, "pbe"),
    _diatomic("h2plus_R2_lda", 2.0, 1, "lda"),
    _atom("h_atom_N0.5_lda", 1, 0.5, "lda"),
)}

#: The model systems of the reference's registry (no nuclei: the plain operator of ``model.py``).
MODEL_IDS = ("harmonic_w1", "box_L10")
#: Registered in the reference and not solvable here (none at present).
UNSUPPORTED: dict[str, str] = {}

# The reference's sizing constants (``grid.py:61-87``).
SINGLE_CENTRE_HALF_BOX_PER_Z = 16.0
SINGLE_CENTRE_POINTS = 33
MOLECULE_VACUUM = 12.0
TARGET_SPACING = 0.25
INTERACTING_MIN_HALF_BOX = 7.0
INTERACTING_TAIL = 1.0e-9
SLATER_SCREENING = 0.30
LATTICE_RTOL = 1.0e-9


def _span(positions) -> float:
    return max(max(abs(a[d] - b[d]) for d in range(3)) for a in positions for b in positions)


def lattice_spacing(positions, target: float) -> float:
    """The spacing nearest ``target`` that puts every nucleus on the lattice: the smallest non-zero
    coordinate difference split into ``round(d / target)`` cells (``grid.py:190-228``); ``target``
    when the nuclei share every coordinate or admit no common lattice."""
    diffs = sorted({abs(a[d] - b[d]) for i, a in enumerate(positions) for b in positions[i + 1:] for d in range(3)
                    if abs(a[d] - b[d]) > 0.0})
    if not diffs:
        return float(target)
    h = diffs[0] / max(1, int(round(diffs[0] / target)))
    ratios = [d / h for d in diffs]
    if all(abs(r - round(r)) <= LATTICE_RTOL * max(1.0, r) for r in ratios):
        return h
    return float(target)


def whole_cells(edge: float, h: float) -> float:
    """``edge`` rounded up to whole cells of ``h`` (``grid.py:231-238``)."""
    return max(int(math.ceil(edge / h - LATTICE_RTOL)), 1) * h


def derived_numerics(scenario: Scenario, base: Numerics | None = None) -> tuple[Numerics, tuple]:
    """The grid the reference derives from the nuclei (``grid.py:153-183,261-290,338-420``), anchored
    on the first nucleus (``grid.py:448-457``).

    Bare: one centre of charge Z takes the edge ``32 / Z`` with 33 points per edge; several centres
    the nuclear span plus 12 bohr in whole cells of the lattice spacing nearest 0.25. Interacting:
    the span plus twice ``max(7, ln(1e9) / (2 kappa))``, ``kappa = max(min_a (Z_a - 0.3 max(0,
    N/atoms - 1)), 0.25)``, in whole cells of the lattice spacing nearest 0.25."""
    base = Numerics() if base is None else base
    positions = [tuple(float(x) for x in p) for p in scenario.positions]
    charges = [int(z) for z in scenario.charges]
    if scenario.interacting:
        per_atom = scenario.n_electrons / len(charges)
        kappa = max(min(z - SLATER_SCREENING * max(0.0, per_atom - 1.0) for z in scenario.charges), 0.25)
        half = max(INTERACTING_MIN_HALF_BOX, math.log(1.0 / INTERACTING_TAIL) / (2.0 * kappa))
        h = lattice_spacing(positions, TARGET_SPACING)
        edge = whole_cells(_span(positions) + 2.0 * half, h)
    elif len(charges) == 1:
        edge = 2.0 * SINGLE_CENTRE_HALF_BOX_PER_Z / max(charges[0], 1)
        h = edge / (SINGLE_CENTRE_POINTS - 1)
    else:
        h = lattice_spacing(positions, TARGET_SPACING)
        edge = whole_cells(_span(positions) + MOLECULE_VACUUM, h)
    return base.with_grid(h, edge), positions[0]


# ------------------------------------------------------------------------------------------------
# Recording
# ------------------------------------------------------------------------------------------------

#: The energy terms of the reference's breakdown, in its key set (the dispersion, entropy and
#: non-local pseudopotential terms are zero on this all-electron, integer-occupation path).
ENERGY_KEYS = ("dispersion", "entropy", "external", "harris_foulkes", "hartree", "ion_ion", "kinetic",
               "nonlocal_ps", "total", "xc")


def _digest(t: torch.Tensor) -> str:
    array = np.ascontiguousarray(t.detach().cpu().numpy().astype(np.float64))
    return hashlib.sha256(array.tobytes()).hexdigest()


def grid_record(grid) -> dict:
    """The reference's ``UniformGrid.describe()`` for a box grid with a zero halo."""
    n = int(grid.n_points)
    return {"boundary": "zero", "fd_order": int(grid.fd_order), "gradient_order": int(grid.fd_order),
            "mask_fraction": 1.0, "n_box_points": n, "n_points": n, "origin_bohr": [float(x) for x in grid.origin],
            "pad": int(grid.fd_order) // 2, "shape": [int(s) for s in grid.shape], "spacing_bohr": float(grid.spacing)}


def record(scenario: Scenario, device: torch.device, numerics: Numerics | None = None) -> dict:
    """Solve one scenario and return its fingerprint entry."""
    from . import ops, scf

    num, anchor = derived_numerics(scenario, numerics)
    started = time.perf_counter()
    n_states = max(1, int(-(-float(scenario.n_electrons) // 2.0)))
    if scenario.interacting:
        r = scf.solve_self_consistent(scenario, num, device, anchor=anchor, cross_check=False)
        e = r["energies"]
        energies = {"external": e["external"], "harris_foulkes": e["harris_foulkes"], "hartree": e["hartree"],
                    "ion_ion": e["ion_ion"], "kinetic": e["kinetic"], "total": e["total"], "xc": e["xc"]}
        grid = r["step"].grid
        density = r["density"].reshape(1, -1)
        vectors = r["eigenvectors"][:n_states]
        eigenvalues = np.asarray(r["eigenvalues"])[:n_states]
        converged, iterations = r["converged"], r["scf_iterations"]
    else:
        r = scf.solve_noninteracting(scenario, num, device, anchor=anchor, cross_check=False)
        h = r["hamiltonian"]
        grid = h.grid
        ion = r["total"] - r["band"]
        energies = {"external": r["total"] - r["kinetic"] - ion, "harris_foulkes": r["total"], "hartree": 0.0,
                    "ion_ion": ion, "kinetic": r["kinetic"], "total": r["total"], "xc": 0.0}
        vectors = r["eigenvectors"][:n_states]
        eigenvalues = np.asarray(r["eigenvalues"])[:n_states]
        occ_dev = scf.aufbau(eigenvalues, scenario.n_electrons)
        rho = grid.empty()
        n = grid.empty()
        ops.density(vectors.contiguous(), occ_dev, h.factor.weight, rho, n)
        density = n.reshape(1, -1)
        converged, iterations = r["converged"], r["n_iterations"]
    occupations = scf.aufbau(eigenvalues, scenario.n_electrons)
    energies.update({"dispersion": 0.0, "entropy": 0.0, "nonlocal_ps": 0.0})
    peak = torch.cuda.max_memory_allocated(device) if device.type == "cuda" else None
    return {
        "status": "valid" if converged else "unconverged",
        "energies": {k: repr(float(energies[k])) for k in ENERGY_KEYS},
        "eigenvalues": [repr(float(x)) for x in eigenvalues],
        "occupations": [repr(float(x)) for x in np.asarray(occupations)[:n_states]],
        "density_sha256": _digest(density),
        "orbitals_sha256": _digest(vectors.reshape(1, n_states, -1)),
        "grid": grid_record(grid),
        "n_iterations": int(iterations),
        "n_states": n_states,
        "wall_time_s": round(time.perf_counter() - started, 3),
        "gpu_peak_bytes": int(peak) if peak is not None else None,
    }


def model_record(scenario_id: str, device: torch.device) -> dict:
    """Solve one model system (``model.solve_model``) and return its fingerprint entry."""
    from . import model

    started = time.perf_counter()
    r = model.solve_model(scenario_id, device)
    system, grid = r["system"], r["grid"]
    vectors = r["eigenvectors"]
    n_states = int(vectors.shape[0])
    occ = np.asarray(r["occupations"], dtype=np.float64)[:n_states]
    density = (torch.as_tensor(occ, dtype=torch.float64, device=vectors.device)[:, None] * vectors * vectors).sum(0)
    energies = dict.fromkeys(ENERGY_KEYS, 0.0)
    energies.update({"kinetic": r["kinetic"], "external": r["external"], "total": r["total"],
                     "harris_foulkes": r["band"]})
    rec = grid_record(grid)
    rec["boundary"] = "odd_reflection" if system.boundary == 1 else "zero"
    peak = torch.cuda.max_memory_allocated(device) if device.type == "cuda" else None
    return {
        "status": "valid" if r["converged"] else "unconverged",
        "energies": {k: repr(float(energies[k])) for k in ENERGY_KEYS},
        "eigenvalues": [repr(float(x)) for x in r["eigenvalues"]],
        "occupations": [repr(float(x)) for x in occ],
        "density_sha256": _digest(density.reshape(1, -1)),
        "orbitals_sha256": _digest(vectors.reshape(1, n_states, -1)),
        "grid": rec,
        "n_iterations": int(r["n_iterations"]),
        "n_states": n_states,
        "wall_time_s": round(time.perf_counter() - started, 3),
        "gpu_peak_bytes": int(peak) if peak is not None else None,
    }


def _write(target: pathlib.Path, data: dict) -> None:
    target.parent.mkdir(parents=True, exist_ok=True)
    partial = target.with_suffix(target.suffix + ".partial")
    partial.write_text(json.dumps(data, indent=2, sort_keys=True), encoding="utf-8")
    partial.replace(target)


def fingerprint(device: str | torch.device = "cuda", scenario_ids=None, target: pathlib.Path | None = None) -> dict:
    """Fingerprint the registered scenarios (all, or ``scenario_ids``); with ``target`` the file is
    rewritten after every scenario."""
    dev = torch.device(device)
    if dev.type != "cuda" or not torch.cuda.is_available():
        raise RuntimeError("the fingerprint solves on the GPU (the HIP kernels); no GPU is visible")
    names = list(MODEL_IDS) + list(REGISTRY) + list(UNSUPPORTED) if scenario_ids is None else list(scenario_ids)
    unknown = [s for s in names if s not in REGISTRY and s not in UNSUPPORTED and s not in MODEL_IDS]
    if unknown:
        raise SystemExit(f"unknown scenario id(s): {', '.join(unknown)}")
    out: dict = {"_meta": {"device_requested": str(device), "device": str(dev),
                           "device_name": torch.cuda.get_device_name(dev), "torch": torch.__version__,
                           "torch_cuda": str(getattr(torch.version, "hip", None)), "threads": int(torch.get_num_threads()),
                           "hostname": platform.node(), "deterministic_requested": False,
                           "deterministic_algorithms": "off"}}
    for sid in names:
        print(f"[fingerprint] {sid} on {dev} ...", flush=True)
        if sid in UNSUPPORTED:
            out[sid] = {"status": "unsupported", "error": UNSUPPORTED[sid]}
        elif sid in MODEL_IDS:
            out[sid] = model_record(sid, dev)
        else:
            out[sid] = record(REGISTRY[sid], dev)
        print(f"[fingerprint] {sid}: {out[sid]['status']}", flush=True)
        if target is not None:
            _write(target, out)
    return out


# ------------------------------------------------------------------------------------------------
# Comparison
# ------------------------------------------------------------------------------------------------

#: Compared numerically: energies (a mapping), eigenvalues and occupations (lists).
NUMERIC = ("energies", "eigenvalues", "occupations")
#: Compared for equality and reported, never failing unless ``exact``.
DIGESTS = ("density_sha256", "orbitals_sha256")
#: Facts of a run rather than results: ignored by ``exact``.
RUN_FACTS = ("wall_time_s", "gpu_peak_bytes", "n_states")


def _number(v) -> float:
    return math.nan if v is None or v == "None" else float(v)


def largest_difference(left, right) -> float:
    """The largest ``|a - b|`` over two lists or mappings of numbers (as stored, ``repr`` strings or
    numbers); infinite when their shapes or keys differ, NaN on both sides counting as equal."""
    if isinstance(left, dict) and isinstance(right, dict):
        if left.keys() != right.keys():
            return math.inf
        pairs = [(left[k], right[k]) for k in left]
    elif isinstance(left, list) and isinstance(right, list):
        if len(left) != len(right):
            return math.inf
        pairs = list(zip(left, right))
    else:
        return 0.0 if left == right else math.inf
    worst = 0.0
    for a, b in pairs:
        x, y = _number(a), _number(b)
        if math.isnan(x) and math.isnan(y):
            continue
        d = abs(x - y)
        worst = max(worst, math.inf if math.isnan(d) else d)
    return worst


def diff(a: dict, b: dict, tol: float = 1.0e-10, occupation_tol: float = 0.0, exact: bool = False,
         out=print) -> int:
    """Report the differences of two fingerprints; return the number of failing scenarios."""
    for side, fp in (("left", a), ("right", b)):
        if fp.get("_meta"):
            out(f"{side:<6} {json.dumps(fp['_meta'], sort_keys=True)}")
    ids = sorted(k for k in set(a) | set(b) if not k.startswith("_"))
    failing = 0
    worst = dict.fromkeys(NUMERIC, 0.0)
    digests_equal = dict.fromkeys(DIGESTS, 0)
    digests_seen = dict.fromkeys(DIGESTS, 0)
    for sid in ids:
        left, right = a.get(sid), b.get(sid)
        if left is None or right is None:
            failing += 1
            out(f"{sid:<16} FAIL   present on one side only")
            continue
        if exact:
            l_cmp = {k: v for k, v in left.items() if k not in RUN_FACTS}
            r_cmp = {k: v for k, v in right.items() if k not in RUN_FACTS}
            if None not in (left.get("n_states"), right.get("n_states")):
                l_cmp["n_states"], r_cmp["n_states"] = left["n_states"], right["n_states"]
            if l_cmp == r_cmp:
                for key in DIGESTS:
                    if left.get(key) is not None:
                        digests_seen[key] += 1
                        digests_equal[key] += 1
                out(f"{sid:<16} identical")
                continue
        problems = []
        deltas = {}
        for name in NUMERIC:
            if name not in left and name not in right:
                continue
            d = largest_difference(left.get(name), right.get(name))
            deltas[name] = d
            worst[name] = max(worst[name], d)
            limit = occupation_tol if name == "occupations" else tol
            if exact:
                if left.get(name) != right.get(name):
                    problems.append(f"{name} differ")
            elif d > limit:
                problems.append(f"{name} |delta| {d:.3e} > {limit:g}")
        for key in ("status", "grid"):
            if left.get(key) != right.get(key):
                problems.append(f"{key}: {left.get(key)} -> {right.get(key)}")
        if left.get("error") or right.get("error"):
            problems.append(f"error: {left.get('error')!r} / {right.get('error')!r}")
        notes = []
        for key in DIGESTS:
            if left.get(key) is None and right.get(key) is None:
                continue
            digests_seen[key] += 1
            same = left.get(key) == right.get(key)
            digests_equal[key] += same
            notes.append(f"{key.split('_')[0]}={'equal' if same else 'unequal'}")
            if exact and not same:
                problems.append(f"{key} differs")
        if None not in (left.get("n_states"), right.get("n_states")) and left["n_states"] != right["n_states"]:
            problems.append(f"n_states {left['n_states']} -> {right['n_states']}")
        if exact and left.get("n_iterations") != right.get("n_iterations"):
            problems.append("n_iterations differ")
        failing += bool(problems)
        numbers = "  ".join(f"{k}={v:.3e}" for k, v in deltas.items())
        out(f"{sid:<16} {'FAIL' if problems else 'ok':<6} {numbers}  {' '.join(notes)}  "
            f"iter {left.get('n_iterations')}/{right.get('n_iterations')}")
        for p in problems:
            out(f"   {p}")
    out("-" * 78)
    out("max |delta| over all scenarios: " + "  ".join(f"{k}={v:.3e}" for k, v in worst.items())
        + ("  (exact mode)" if exact else f"  (tol {tol:g}, occupation tol {occupation_tol:g})"))
    out("hashes equal: " + "  ".join(f"{k}={digests_equal[k]}/{digests_seen[k]}" for k in DIGESTS)
        + ("" if exact else "  (informational: summation order changes them)"))
    out(f"{failing} failing scenario(s) of {len(ids)}")
    return failing


def main(argv=None) -> int:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("paths", nargs="+", help="output path, or two paths with --diff")
    parser.add_argument("--diff", action="store_true", help="compare two stored fingerprints")
    parser.add_argument("--device", default="cuda", help="device to solve on (default cuda)")
    parser.add_argument("--scenarios", nargs="+", default=None, help="restrict to these scenario ids")
    parser.add_argument("--tol", type=float, default=1.0e-10, help="energy/eigenvalue tolerance for --diff")
    parser.add_argument("--occupation-tol", type=float, default=0.0, help="occupation tolerance for --diff")
    parser.add_argument("--exact", action="store_true", help="--diff: any difference at all fails")
    args = parser.parse_args(argv)
    if args.diff:
        if len(args.paths) != 2:
            parser.error("--diff takes exactly two paths")
        a, b = (json.loads(pathlib.Path(p).read_text(encoding="utf-8")) for p in args.paths)
        return 1 if diff(a, b, tol=args.tol, occupation_tol=args.occupation_tol, exact=args.exact) else 0
    target = pathlib.Path(args.paths[0])
    data = fingerprint(args.device, args.scenarios, target)
    _write(target, data)
    print(f"fingerprint of {len(data) - 1} scenarios written to {target}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
