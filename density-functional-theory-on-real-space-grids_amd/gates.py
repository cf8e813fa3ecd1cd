"""The re-solve gates of the reference's gate suite, on the device (reference
``src/cdft/gates/tier3.py:30-480``, ``gates/tier5.py:236-363``, verdict rule ``gates/base.py:40,134-177``).

Each gate re-solves the scenario of a completed run and thresholds a difference:

- G3.1 grid-spacing convergence: three rungs at ``h, h/1.25, h/1.5625`` on the run's box, the verdict
  ``|E(h/1.25) - E(h/1.5625)| / atoms`` of the ground eigenvalue, the fitted order, the Richardson
  limit and the monotonicity reported beside it (``tier3.py:76-180``);
- G3.2 domain-size convergence: the run's box and the box grown by 4 bohr at the run's spacing,
  ``|E_large - E_small| / atoms`` (``tier3.py:183-287``);
- G3.4 initial-guess independence: the SCF from every other starting density, the spread of the
  total energies (``tier3.py:409-460``);
- G3.5 stencil-order independence: the ground eigenvalue at order 12 against the run's order at the
  run's spacing (``tier3.py:290-352``);
- G5.2 determinism: two strict re-solves of the run's configuration, geometry rebuilt each time,
  ``max |eps_A - eps_B|`` (``tier5.py:236-363``).

G3.1, G3.2 and G3.5 solve one state on the run's resolved grid with the varied quantity changed and
the coordinate origin on the lattice (the reference's ``derive_grid=False``) and skip an interacting
scenario, as the reference does. Their rungs are independent fixed-potential eigenproblems of
different sizes: ``resolve_gates`` solves every distinct rung of the gates it runs as ONE lockstep
batch (``batch.BatchedCheFSI``: one stencil launch per filter step for all rungs, each rung's
eigenvalue and iteration count bitwise those of its own solve), and G3.1's production rung, G3.2's
small box and G3.5's run's-order solve, one problem, are solved once; G3.5's order-12 solve runs on
its own (a batch runs one stencil order). The re-solves of G3.4 and of the G5.2 pair are
independent too: they run at once, one host thread and one stream each (``at_once``), every scratch
buffer, cached solver and captured graph being keyed by the stream.
"""

from __future__ import annotations

import dataclasses
import math
import threading
import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import scf

#: One meV in hartree, the per-atom threshold of the convergence gates (``tier3.py:30``).
MILLI_EV = 3.674932e-5
#: A passing measurement above this fraction of its threshold is MARGINAL (``base.py:40``).
MARGINAL_FRACTION = 0.9

PASS, MARGINAL, FAIL, SKIPPED = "pass", "marginal", "fail", "skipped"


@dataclass(frozen=True)
class GateSpec:
    gate_id: str
    name: str
    threshold: float
    kind: str
    citation: str
    units: str
    solves: int = 0


@dataclass
class GateResult:
    """One gate's outcome: verdict, measured value, threshold and the detail record
    (reference ``contract.py:889-906``)."""

    gate_id: str
    name: str
    verdict: str
    measured: float
    threshold: float
    kind: str
    citation: str = ""
    units: str = ""
    reason: str = ""
    detail: dict = field(default_factory=dict)

    @property
    def margin(self) -> float:
        """The fraction of the threshold the measurement uses."""
        return math.inf if self.threshold == 0.0 else abs(self.measured) / abs(self.threshold)


SPECS = {
    "G3.1": GateSpec("G3.1", "Grid-spacing convergence", MILLI_EV, "derived", "D6; D9", "Ha/atom", 3),
    "G3.2": GateSpec("G3.2", "Domain-size convergence", MILLI_EV, "derived", "D5", "Ha/atom", 2),
    "G3.4": GateSpec("G3.4", "Initial-guess independence", 1.0e-7, "exact", "F13", "Ha", 2),
    "G3.5": GateSpec("G3.5", "Stencil-order independence", MILLI_EV, "derived", "D2", "Ha/atom", 2),
    "G5.2": GateSpec("G5.2", "Determinism", 1.0e-9, "exact", "G5.2; D13; D-61", "Ha", 2),
}

#: Refinement ratio and rung count of the spacing ladder (``tier3.py:95-98``).
LADDER_RATIO = 1.25
LADDER_RUNGS = 3
#: Box growth in bohr of the domain-size gate (``tier3.py:201-202``).
BOX_GROWTH = 4.0
#: Comparison order of the stencil-order gate (``tier3.py:304-305``).
HIGH_ORDER = 12

INTERACTING_LADDER_REASON = (
    "self-consistent scenario: the ladder gates re-solve the fixed-potential problem only; the "
    "interacting path's ladders are measured once and recorded, not re-run per suite pass")


def make_result(spec: GateSpec, measured: float, detail: dict | None = None) -> GateResult:
    """The verdict rule (``base.py:134-160``): a non-finite measurement fails, above the threshold
    fails, above ``MARGINAL_FRACTION`` of it is marginal, else passes."""
    if not math.isfinite(measured) or abs(measured) > spec.threshold:
        verdict = FAIL
    elif spec.threshold > 0.0 and abs(measured) > MARGINAL_FRACTION * spec.threshold:
        verdict = MARGINAL
    else:
        verdict = PASS
    return GateResult(spec.gate_id, spec.name, verdict, float(measured), float(spec.threshold), spec.kind,
                      spec.citation, spec.units, detail=detail or {})


def skipped(spec: GateSpec, reason: str) -> GateResult:
    """A SKIPPED result; the reason is mandatory (``base.py:163-177``)."""
    if not reason:
        raise ValueError("a skipped gate must say why")
    return GateResult(spec.gate_id, spec.name, SKIPPED, math.nan, float(spec.threshold), spec.kind, spec.citation,
                      spec.units, reason=reason)


# ------------------------------------------------------------------------------------------------
# Rungs: fixed-potential solves on a given spacing and box, batched
# ------------------------------------------------------------------------------------------------


def _n_atoms(run: scf.Run) -> int:
    return max(len(run.scenario.charges), 1)


def ladder_spacings(run: scf.Run) -> list[float]:
    """The refining spacings of G3.1, coarse to fine (``tier3.py:118-124``)."""
    base = float(run.numerics.spacing)
    return [base / (LADDER_RATIO**rung) for rung in range(LADDER_RUNGS)]


def grown_box(run: scf.Run) -> tuple:
    """The box of G3.2's second solve: every edge grown by ``BOX_GROWTH`` (``tier3.py:227``)."""
    return tuple(float(edge) + BOX_GROWTH for edge in run.numerics.box_lengths)


class Rungs:
    """Ground eigenvalues of the run's scenario on (spacing, box) rungs, one state each, the origin on
    the lattice. ``solve`` runs every rung not yet known as one lockstep batch and keeps the results,
    so gates sharing a rung solve it once."""

    def __init__(self, run: scf.Run):
        self.run = run
        self.known: dict = {}
        self.batches: list[dict] = []

    @staticmethod
    def key(spacing: float, box) -> tuple:
        return (float(spacing), tuple(float(b) for b in box))

    def solve(self, rungs) -> None:
        from .batch import BatchedCheFSI
        from .hamiltonian import CuspFactoredHamiltonian

        todo = []
        for spacing, box in rungs:
            k = self.key(spacing, box)
            if k not in self.known and k not in todo:
                todo.append(k)
        if not todo:
            return
        started = time.perf_counter()
        run = self.run
        hams = []
        for spacing, box in todo:
            num = dataclasses.replace(run.numerics, spacing=spacing, box_lengths=box)
            hams.append(CuspFactoredHamiltonian(*scf.build_geometry(run.scenario, num, run.device, anchor=None)))
        solver = BatchedCheFSI(hams, run.numerics.eigen, n_states=1, ritz_stop=True, seed=run.numerics.seed)
        results = solver.solve()
        for k, h, r in zip(todo, hams, results):
            self.known[k] = {"eigenvalue": float(r["eigenvalues"][0]), "n_iterations": int(r["n_iterations"]),
                             "converged": bool(r["converged"]), "shape": tuple(int(s) for s in h.grid.shape),
                             "spacing": float(h.grid.spacing)}
        self.batches.append({"rungs": len(todo), "shapes": [self.known[k]["shape"] for k in todo],
                             "lockstep_iterations": int(solver.iterations),
                             "wall_s": time.perf_counter() - started})

    def __getitem__(self, rung) -> dict:
        k = self.key(*rung)
        if k not in self.known:
            self.solve([rung])
        return self.known[k]


# ------------------------------------------------------------------------------------------------
# Independent re-solves at once: one host thread and one stream each
# ------------------------------------------------------------------------------------------------

#: Side streams per device, kept for the process so that the solvers, plans and graphs cached per
#: stream are reused by the next gate evaluation.
_side_streams: dict = {}
_side_lock = threading.Lock()


def side_streams(device: torch.device, n: int) -> list:
    with _side_lock:
        pool = _side_streams.setdefault(str(device), [])
        while len(pool) < n:
            pool.append(torch.cuda.Stream(device))
        return pool[:n]


def at_once(device: torch.device, jobs) -> list:
    """Run each job (a callable without arguments) on a host thread of its own with a stream of its
    own on ``device`` current, and return their results in order, each with its wall time. Every
    scratch buffer, cached solver and captured graph of the solvers is keyed by the caller's stream,
    so the jobs share none. The side streams wait for the caller's stream before the jobs start, and
    the caller's stream waits for them after; the first error of a job is raised."""
    jobs = list(jobs)
    streams = side_streams(device, len(jobs))
    origin = torch.cuda.current_stream(device)
    for s in streams:
        s.wait_stream(origin)
    out: list = [None] * len(jobs)
    errors: list = [None] * len(jobs)

    def work(i):
        try:
            with torch.cuda.stream(streams[i]):
                started = time.perf_counter()
                value = jobs[i]()
                streams[i].synchronize()
                out[i] = (value, time.perf_counter() - started)
        except BaseException as exc:  # re-raised on the caller's thread below
            errors[i] = exc

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for s in streams:
        origin.wait_stream(s)
    for exc in errors:
        if exc is not None:
            raise exc
    return out


def _sequentially(jobs) -> list:
    out = []
    for job in jobs:
        started = time.perf_counter()
        value = job()
        out.append((value, time.perf_counter() - started))
    return out


# ------------------------------------------------------------------------------------------------
# The gates
# ------------------------------------------------------------------------------------------------


def spacing_ladder(run: scf.Run, rungs: Rungs | None = None) -> GateResult:
    """G3.1 (``tier3.py:76-180``)."""
    spec = SPECS["G3.1"]
    if run.scenario.interacting:
        return skipped(spec, INTERACTING_LADDER_REASON)
    rungs = Rungs(run) if rungs is None else rungs
    box = tuple(run.numerics.box_lengths)
    spacings = ladder_spacings(run)
    rungs.solve([(h, box) for h in spacings])
    solved = [rungs[(h, box)] for h in spacings]
    if not all(r["converged"] for r in solved):
        return skipped(spec, "a rung of the spacing ladder did not converge")
    energies = [r["eigenvalue"] for r in solved]
    n_atoms = _n_atoms(run)
    steps = [energies[i + 1] - energies[i] for i in range(len(energies) - 1)]
    monotone = all(s <= 0 for s in steps) or all(s >= 0 for s in steps)
    shrinking = all(abs(steps[i]) > abs(steps[i + 1]) for i in range(len(steps) - 1))
    order = math.nan
    if len(steps) >= 2 and steps[0] != 0.0 and steps[1] != 0.0:
        order = math.log(abs(steps[0] / steps[1])) / math.log(LADDER_RATIO)
    measured = abs(steps[-1]) / n_atoms
    extrapolated = None
    if monotone and math.isfinite(order) and order > 0.0:
        extrapolated = energies[-1] + (energies[-1] - energies[-2]) / (LADDER_RATIO**order - 1.0)
    return make_result(spec, measured, {
        "spacings_coarse_to_fine": spacings,
        "energies": energies,
        "successive_differences": steps,
        "monotone": monotone,
        "differences_shrinking": shrinking,
        "measured_order": order,
        "extrapolated_zero_spacing": extrapolated,
        "verdict_basis": "measured difference between the two finest rungs",
        "remaining_error_estimate": None if extrapolated is None else abs(energies[-1] - extrapolated) / n_atoms,
        "production_spacing": spacings[0],
        "per_atom": measured,
        "grid_shapes": [r["shape"] for r in solved],
        "n_iterations": [r["n_iterations"] for r in solved],
    })


def box_growth(run: scf.Run, rungs: Rungs | None = None) -> GateResult:
    """G3.2 (``tier3.py:183-287``). The reference reads the boundary decay of ``psi = f phi`` from a
    measurement its solves never record, so its decay is always absent and the verdict is the energy
    shift; the same holds here."""
    spec = SPECS["G3.2"]
    if run.scenario.interacting:
        return skipped(spec, INTERACTING_LADDER_REASON)
    rungs = Rungs(run) if rungs is None else rungs
    h = float(run.numerics.spacing)
    box = tuple(float(b) for b in run.numerics.box_lengths)
    grown = grown_box(run)
    rungs.solve([(h, box), (h, grown)])
    small, large = rungs[(h, box)], rungs[(h, grown)]
    if not (small["converged"] and large["converged"]):
        return skipped(spec, "a solve of the box ladder did not converge")
    shift = abs(large["eigenvalue"] - small["eigenvalue"]) / _n_atoms(run)
    return make_result(spec, shift, {
        "box_lengths": list(box),
        "grown_box_lengths": list(grown),
        "energy_small_box": small["eigenvalue"],
        "energy_large_box": large["eigenvalue"],
        "energy_shift_per_atom": shift,
        "boundary_decay": None,
        "decay_dominates_verdict": False,
        "grid_shapes": [small["shape"], large["shape"]],
    })


def stencil_order(run: scf.Run, rungs: Rungs | None = None) -> GateResult:
    """G3.5 (``tier3.py:290-352``): the ground eigenvalue at the run's order against order 12, both on
    the run's spacing and box with the origin on the lattice. The run's-order solve is the production
    rung of the other ladders (taken from ``rungs``); the order-12 one is a fixed-potential solve of
    its own (a batch runs one stencil order)."""
    spec = SPECS["G3.5"]
    if run.scenario.interacting:
        return skipped(spec, INTERACTING_LADDER_REASON)
    low_order = int(run.numerics.fd_order)
    if low_order >= HIGH_ORDER:
        return skipped(spec, f"the run's order {low_order} is not below the comparison order {HIGH_ORDER}")
    rungs = Rungs(run) if rungs is None else rungs
    h = float(run.numerics.spacing)
    box = tuple(float(b) for b in run.numerics.box_lengths)
    low = rungs[(h, box)]
    high_numerics = dataclasses.replace(run.numerics, spacing=h, box_lengths=box, fd_order=HIGH_ORDER,
                                        eigen=dataclasses.replace(run.numerics.eigen, cross_check=False))
    started = time.perf_counter()
    high = scf.solve_scenario(run.scenario, high_numerics, run.device, derive_grid=False, n_states=1)
    wall = time.perf_counter() - started
    if not (low["converged"] and high.converged):
        return skipped(spec, "a solve of the order comparison did not converge")
    e_low, e_high = low["eigenvalue"], float(high.eigenvalues[0])
    difference = abs(e_high - e_low) / _n_atoms(run)
    return make_result(spec, difference, {
        "order_low": low_order,
        "order_high": HIGH_ORDER,
        "energy_low_order": e_low,
        "energy_high_order": e_high,
        "difference_per_atom": difference,
        "spacing": h,
        "n_iterations": [low["n_iterations"], high.n_iterations],
        "high_order_wall_s": round(wall, 4),
    })


def initial_guesses(run: scf.Run, concurrent: bool = True) -> GateResult:
    """G3.4 (``tier3.py:409-460``): the SCF from every other starting density; with ``concurrent``
    those solves run at once (``at_once``)."""
    spec = SPECS["G3.4"]
    if not run.scenario.interacting:
        return skipped(spec, "a fixed-potential solve has no starting density")
    numerics = dataclasses.replace(run.numerics, eigen=dataclasses.replace(run.numerics.eigen, cross_check=False))
    guesses = [g for g in scf.INITIAL_GUESSES if g != run.initial_guess]

    def job(guess):
        def solve():
            r = scf.solve_self_consistent(run.scenario, numerics, run.device, anchor=run.anchor, cross_check=False,
                                          initial_guess=guess)
            return bool(r["converged"]), float(r["energies"]["total"]), int(r["scf_iterations"])
        return solve

    jobs = [job(g) for g in guesses]
    solved = at_once(run.device, jobs) if concurrent else _sequentially(jobs)
    energies = {run.initial_guess: run.total}
    for guess, ((converged, total, _), _) in zip(guesses, solved):
        if not converged:
            return skipped(spec, f"the SCF from the {guess!r} guess did not converge; the comparison is not available")
        energies[guess] = total
    spread = max(energies.values()) - min(energies.values())
    return make_result(spec, spread, {
        "energies": energies,
        "iterations_from_used_guess": run.n_iterations,
        "iterations": {g: r[0][2] for g, r in zip(guesses, solved)},
        "wall_s": {g: round(r[1], 4) for g, r in zip(guesses, solved)},
        "concurrent": bool(concurrent),
    })


def determinism_pair(run: scf.Run, concurrent: bool = True) -> GateResult:
    """G5.2 (``tier5.py:236-363``): solve the run's configuration twice, geometry rebuilt each time,
    the cross-check off, and compare every eigenvalue; the run's own eigenvalues against the first
    are recorded, never thresholded. With ``concurrent`` the two solves run at once (``at_once``),
    on separate streams with separate scratch memory, solvers and graphs."""
    spec = SPECS["G5.2"]
    numerics = dataclasses.replace(run.numerics, eigen=dataclasses.replace(run.numerics.eigen, cross_check=False))
    primary = np.asarray(run.eigenvalues, dtype=np.float64)
    n_states = None if run.scenario.interacting else int(primary.shape[-1])

    def solve():
        if run.scenario.interacting:
            r = scf.solve_self_consistent(run.scenario, numerics, run.device, anchor=run.anchor, cross_check=False)
        else:
            r = scf.solve_noninteracting(run.scenario, numerics, run.device, anchor=run.anchor, cross_check=False,
                                         n_states=n_states)
        return np.asarray(r["eigenvalues"], dtype=np.float64).copy()

    pair = at_once(run.device, [solve, solve]) if concurrent else _sequentially([solve, solve])
    (first, wall_a), (second, wall_b) = pair
    n = min(primary.shape[-1], first.shape[-1], second.shape[-1])
    delta = np.abs(first[:n] - second[:n])
    primary_delta = np.abs(primary[:n] - first[:n])
    return make_result(spec, float(delta.max()), {
        "first_run": first[:n].tolist(),
        "repeat_run": second[:n].tolist(),
        "primary_run": primary[:n].tolist(),
        "per_state_difference": delta.tolist(),
        "primary_vs_strict": float(primary_delta.max()),
        "seed": run.numerics.seed,
        "device": str(run.device),
        "gpu_leg_exercised": run.device.type == "cuda",
        "solves": 2,
        "bitwise_identical": bool((delta == 0.0).all()),
        "strict_pair_wall_s": [round(wall_a, 4), round(wall_b, 4)],
        "primary_wall_s": round(run.wall_s, 4),
        "concurrent": bool(concurrent),
    })


#: The re-solve gates in the order ``resolve_gates`` runs them.
GATES = ("G3.1", "G3.2", "G3.5", "G3.4", "G5.2")


def resolve_gates(run: scf.Run, gate_ids=GATES, concurrent: bool = True) -> tuple[list[GateResult], dict]:
    """Run the re-solve gates ``gate_ids`` on ``run``; returns the results and the wall time of each
    gate (seconds; the shared rung batch under ``"rungs"``). ``concurrent`` runs the independent SCF
    or eigenproblem re-solves of G3.4 and G5.2 at once."""
    unknown = [g for g in gate_ids if g not in SPECS]
    if unknown:
        raise ValueError(f"unknown re-solve gate(s): {', '.join(unknown)}")
    rungs = Rungs(run)
    wall: dict = {}
    if not run.scenario.interacting:
        wanted = []
        box = tuple(float(b) for b in run.numerics.box_lengths)
        if "G3.1" in gate_ids:
            wanted += [(h, box) for h in ladder_spacings(run)]
        if "G3.2" in gate_ids:
            wanted += [(float(run.numerics.spacing), box), (float(run.numerics.spacing), grown_box(run))]
        if "G3.5" in gate_ids:
            wanted += [(float(run.numerics.spacing), box)]
        started = time.perf_counter()
        rungs.solve(wanted)
        wall["rungs"] = time.perf_counter() - started
    out = []
    for gate_id in gate_ids:
        started = time.perf_counter()
        if gate_id == "G3.1":
            result = spacing_ladder(run, rungs)
        elif gate_id == "G3.2":
            result = box_growth(run, rungs)
        elif gate_id == "G3.5":
            result = stencil_order(run, rungs)
        elif gate_id == "G3.4":
            result = initial_guesses(run, concurrent)
        else:
            result = determinism_pair(run, concurrent)
        wall[gate_id] = time.perf_counter() - started
        out.append(result)
    wall["rung_batches"] = rungs.batches
    return out, wall
