"""The reference's fingerprint format (``scripts/fingerprint.py``): the grids this path derives for
the registered Coulomb scenarios against the grids the reference recorded (every field equal, the
origins bit for bit), and the ``--diff`` rules on fingerprints built from the reference's own file.
The GPU fingerprint of every supported scenario against the reference's file is in
``test_fingerprint_gpu.py``."""

from __future__ import annotations

import copy
import json

import torch

from conftest import ROOT

REFERENCE = json.loads((ROOT / "tests" / "golden" / "fingerprint_cpu_reference.json").read_text())


def test_derived_grids_are_the_reference_grids(pkg):
    from cdft_amd import fingerprint as fp
    from cdft_amd.geometry import UniformGrid

    assert set(fp.REGISTRY) | set(fp.MODEL_IDS) == {k for k in REFERENCE if not k.startswith("_")} and not fp.UNSUPPORTED
    for sid, scenario in fp.REGISTRY.items():
        num, anchor = fp.derived_numerics(scenario)
        grid = UniformGrid.from_box(num.box_lengths, num.spacing, torch.device("cpu"), anchor=anchor,
                                    fd_order=num.fd_order)
        assert fp.grid_record(grid) == REFERENCE[sid]["grid"], sid


def test_lattice_rule_closed_cases(pkg):
    from cdft_amd import fingerprint as fp

    assert fp.lattice_spacing([(0.0, 0.0, -0.7), (0.0, 0.0, 0.7)], 0.25) == 1.4 / 6
    assert fp.lattice_spacing([(0.0, 0.0, 0.0)], 0.25) == 0.25
    # 1.0 and 0.3 share no lattice near 0.25: the target is kept
    assert fp.lattice_spacing([(0.0, 0.0, 0.0), (1.0, 0.3, 0.0)], 0.25) == 0.25
    assert fp.whole_cells(13.4, 1.4 / 6) == 58 * (1.4 / 6)
    assert fp.whole_cells(14.0, 0.25) == 14.0  # an exact multiple is not rounded up


def _lines():
    out = []
    return out, out.append


def test_diff_rules(pkg):
    from cdft_amd import fingerprint as fp

    a = {k: v for k, v in REFERENCE.items() if not k.startswith("_")}
    lines, say = _lines()
    assert fp.diff(a, copy.deepcopy(a), out=say) == 0
    assert lines[-1] == f"0 failing scenario(s) of {len(a)}"
    # an energy moved by 5e-11 passes at the default 1e-10, by 2e-10 fails
    b = copy.deepcopy(a)
    b["he_atom_lda"]["energies"]["kinetic"] = repr(float(a["he_atom_lda"]["energies"]["kinetic"]) + 5e-11)
    assert fp.diff(a, b, out=_lines()[1]) == 0
    b["he_atom_lda"]["energies"]["kinetic"] = repr(float(a["he_atom_lda"]["energies"]["kinetic"]) + 2e-10)
    assert fp.diff(a, b, out=_lines()[1]) == 1
    assert fp.diff(a, b, tol=1e-9, out=_lines()[1]) == 0
    # eigenvalues, occupations (exact by default), status, grid, a missing scenario, a missing key
    for change in (lambda e: e["h_atom"]["eigenvalues"].__setitem__(0, repr(-0.4999)),
                   lambda e: e["h_atom_N0.5_lda"]["occupations"].__setitem__(0, "0.5000000001"),
                   lambda e: e["h2_R1.4"].__setitem__("status", "unconverged"),
                   lambda e: e["h2_R1.4"]["grid"].__setitem__("shape", [59, 59, 60]),
                   lambda e: e.pop("he_plus"),
                   lambda e: e["he_plus"]["energies"].pop("xc")):
        b = copy.deepcopy(a)
        change(b)
        assert fp.diff(a, b, out=_lines()[1]) == 1
    # digests and iteration counts are informational, except under --exact
    b = copy.deepcopy(a)
    b["h_atom"]["density_sha256"] = "0" * 64
    b["h_atom"]["n_iterations"] += 1
    b["h_atom"]["wall_time_s"] = 0.0
    assert fp.diff(a, b, out=_lines()[1]) == 0
    assert fp.diff(a, b, exact=True, out=_lines()[1]) == 1
    b = copy.deepcopy(a)
    b["h_atom"]["wall_time_s"] = 0.0  # a run fact: ignored even by --exact
    assert fp.diff(a, b, exact=True, out=_lines()[1]) == 0
