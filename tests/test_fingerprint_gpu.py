"""The fingerprint of every registered Coulomb scenario on the GPU against the reference's CPU
fingerprint (``reports/fingerprint_cpu_reference.json``, committed as data under ``tests/golden``),
through the reference's own comparison rules: every energy term and eigenvalue within 1e-10 Ha (the
``--diff`` default), occupations equal, status and grid equal; and the SCF / eigensolver iteration
counts equal, which ``--diff`` only reports."""

from __future__ import annotations

import json

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu
REFERENCE = json.loads((ROOT / "tests" / "golden" / "fingerprint_cpu_reference.json").read_text())


def test_fingerprint_of_every_coulomb_scenario_matches_the_reference(pkg, device):
    from cdft_amd import fingerprint as fp

    ours = fp.fingerprint(device, list(fp.REGISTRY))
    lines: list[str] = []
    failing = fp.diff({k: REFERENCE[k] for k in fp.REGISTRY}, {k: ours[k] for k in fp.REGISTRY}, out=lines.append)
    print("\n".join(lines), flush=True)
    assert failing == 0
    iterations = {k: (ours[k]["n_iterations"], REFERENCE[k]["n_iterations"]) for k in fp.REGISTRY}
    assert all(a == b for a, b in iterations.values()), iterations
