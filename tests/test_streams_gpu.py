"""Work on two streams of one device at once: scratch memory is per (device, stream).

The reference runs independent solves from several host threads, one CUDA stream each
(``scripts/scan_throughput.py:264-283``). Every scratch buffer here (reduction partials, the
residual-norm partials, Poisson plans, graph-capture side streams, cached subspace solvers) is keyed
by the caller's stream, so such solves never read each other's partial sums. The checks: the
reductions issued on two streams take two distinct workspaces and give bitwise the serial results,
and two SCF solves run at once from two threads on two streams give bitwise the energies of the same
solves run one after the other.
"""

from __future__ import annotations

import threading

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_reductions_on_two_streams_use_separate_workspaces(pkg, device):
    from cdft_amd import ops
    from cdft_amd.contract import Numerics, helium_lda
    from cdft_amd.hamiltonian import CuspFactoredHamiltonian
    from cdft_amd.scf import build_geometry

    ham = CuspFactoredHamiltonian(*build_geometry(helium_lda(), Numerics(box_lengths=(16.0,) * 3, spacing=0.5),
                                                  device))
    geom, q = ham.geometry, ham.measure.q
    n = ham.grid.n_points
    gen = torch.Generator().manual_seed(3)
    blocks = [(torch.rand((5, n), generator=gen, dtype=torch.float64) * 2 - 1).to(device) for _ in range(2)]
    evals = torch.tensor([-0.9, -0.1, 0.2, 0.4, 0.7], dtype=torch.float64, device=device)

    def work(x):
        gram = ops.wgram(x, None, q)
        norms = torch.empty(5, dtype=torch.float64, device=device)
        ops.hdiv_resnorm(x, geom, evals, q, norms)
        return gram, norms

    serial = [work(x) for x in blocks]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(device), torch.cuda.Stream(device)]
    results = [None, None]
    keys = [None, None]
    for _ in range(20):  # interleaved launches on the two streams, no synchronisation between them
        for i, (s, x) in enumerate(zip(streams, blocks)):
            s.wait_stream(torch.cuda.current_stream(device))
            with torch.cuda.stream(s):
                keys[i] = ops.workspace_key(device)
                results[i] = work(x)
    torch.cuda.synchronize()
    assert keys[0] != keys[1]
    ws = [ops._workspaces[("partials", *k)] for k in keys]
    assert ws[0].data_ptr() != ws[1].data_ptr()
    assert ops._stencil_partials[keys[0]].data_ptr() != ops._stencil_partials[keys[1]].data_ptr()
    for (g, r), (gs, rs) in zip(results, serial):
        assert torch.equal(g, gs) and torch.equal(r, rs)


def test_two_scf_solves_on_two_streams_match_serial(pkg, device):
    from cdft_amd import scf
    from cdft_amd.contract import Numerics, helium_lda, hydrogen_molecule_lda

    cases = [(helium_lda(), Numerics(box_lengths=(16.0,) * 3, spacing=0.5), None),
             (hydrogen_molecule_lda(1.4, axis=0), Numerics(box_lengths=(9.8, 8.4, 8.4), spacing=0.35),
              (-0.7, 0.0, 0.0))]
    serial = [scf.solve_self_consistent(sc, num, device, anchor=a) for sc, num, a in cases]
    torch.cuda.synchronize()
    out = [None, None]
    errors = []

    def run(i):
        try:
            stream = torch.cuda.Stream(device)
            with torch.cuda.stream(stream):
                sc, num, a = cases[i]
                out[i] = scf.solve_self_consistent(sc, num, device, anchor=a)
            stream.synchronize()
        except Exception as exc:  # noqa: BLE001 - reported below
            errors.append(repr(exc))

    threads = [threading.Thread(target=run, args=(i,)) for i in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
    assert not any(t.is_alive() for t in threads), "a threaded solve did not finish"
    assert not errors, errors
    for a, b in zip(out, serial):
        assert a["energy_history"] == b["energy_history"]
        assert np.array_equal(a["eigenvalues"], b["eigenvalues"])
        assert a["scf_iterations"] == b["scf_iterations"]
