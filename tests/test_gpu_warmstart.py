"""What the trained warm-start network gains over the method without it, with the benchmark's
protocol at a small scale: a network trained on the batch stepper's states of small channels from
the sweep's generator starts held-out channels closer to their steady state, so the same steady
rules (the same absolute residual threshold: the freestream run's first residual times 10^-orders)
stop them in fewer iterations. Every run still ends under the plain solver's rules; the network only
moves where they start."""

from __future__ import annotations

import numpy as np
import pytest

from autan_amd import sweep, warmstart

pytestmark = pytest.mark.gpu

CELLS = (2e4, 3e4)


def test_the_trained_network_saves_iterations():
    quiet = lambda *_: None  # noqa: E731
    net, info = warmstart.train_network(sweep.specs(48, seed=21, cells=CELLS), iterations=8000, epochs=100, log=quiet)
    assert info["loss"] < 0.2 * info["first_loss"]
    cmp = warmstart.compare(net, sweep.specs(4, seed=22, cells=CELLS), orders=2.0, max_iterations=20000, log=quiet)
    it_free = sum(f["result"].iterations for f in cmp["free"])
    it_warm = sum(w["result"].iterations for w in cmp["warm_abs"])
    assert all(w["result"].status != 2 for w in cmp["warm_abs"])  # no warm start diverges
    assert np.all(np.isfinite([w["result"].residuals for w in cmp["warm_abs"]]))
    print(f"iterations to the threshold: freestream {it_free}, warm {it_warm} ({it_free / it_warm:.2f}x)")
    assert it_warm < 0.8 * it_free, (it_warm, it_free)


def test_the_warm_start_travels_as_an_initial_state_file(tmp_path):
    """The network's state of a case written as an autan.initial-state/1 file, read back with the
    runtime's checks and paired with the cells by centroid, is the warm start's state bit for bit."""
    from autan_amd import initial_state
    from autan_amd.mlp import MLP

    net = MLP(7, 256, 4, 5, seed=1)
    net.set_normalization(np.zeros(7), np.ones(7), np.array([1.0, 0.5, 0.0, 0.0, 1 / 1.4]), np.full(5, 0.05))
    spec = sweep.specs(1, seed=5, cells=CELLS)[0]
    digest, case = warmstart.write_initial_state(tmp_path / "state.json", net, spec)
    st = initial_state.read(tmp_path / "state.json", sha256=digest, mesh_sha256=warmstart.mesh_digest(case.mesh))
    m = case.mesh
    u = initial_state.start_state(st, m.centroids(), np.column_stack([m.x, m.y, m.z]), warmstart.GAMMA)
    states, _ = warmstart.warm_states(net, [spec])
    net.close()
    assert u.shape == states[0].shape
    assert np.array_equal(u.view(np.uint64), np.asarray(states[0], dtype=np.float64).view(np.uint64))
