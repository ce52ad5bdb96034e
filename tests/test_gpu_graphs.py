"""Captured step pairs (meso_hip_set_graphs) against eager steps, bit for bit
(GPU, through the C ABI).

Each check replays the graphs several times, with eager work of the program
between the replays (macro and state downloads, residuals, single eager steps
that switch the buffer parity, a parameter change that forces a new capture),
and compares every replay with the eager context.
"""
import numpy as np
import pytest

from meso_hip import capi, configs

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def same(a, b, what):
    ma, mb = a.macros(), b.macros()
    for key in ("rho", "u", "T", "q"):
        assert np.array_equal(ma[key], mb[key]), f"{what}: {key}"
    fa, fb = a.flux_moments(), b.flux_moments()
    for x, y in zip(fa, fb):
        assert np.array_equal(x, y), f"{what}: flux moments"


@pytest.mark.parametrize("name,storage", [("shak_kn1_quad10", capi.STORAGE_FP32),
                                          ("shak_kn1_quad10", capi.STORAGE_FP64),
                                          ("shak_quad6_nc6_lim", capi.STORAGE_FP64),
                                          ("bgk_quad8", capi.STORAGE_FP32)])
def test_graph_replays_equal_eager_steps(case_factory, name, storage):
    cs = capi.Case(case_factory(name))
    _say(f"[gpu] graphs {name} NC={cs.n_cells} NV={cs.n_vel} storage={storage}")
    eager, graph = capi.Device(cs, storage), capi.Device(cs, storage)
    graph.set_graphs(True)
    for d in (eager, graph):
        d.init_equilibrium()
    for r in range(6):
        n = 1 if r == 3 else 2  # one odd step: the next pair starts from the other parity
        if r == 4:  # a parameter change: both contexts take it, the graphs are captured again
            p = capi.StepParams.from_buffer_copy(cs.params)
            p.dt *= 0.9
            eager.set_params(p)
            graph.set_params(p)
        re, rg = eager.step(n), graph.step(n)
        assert re.steps_done == rg.steps_done == n and re.bad_cells == rg.bad_cells == 0
        same(eager, graph, f"{name} round {r}")
        assert np.array_equal(eager.residual(), graph.residual(), equal_nan=True)  # 2D: w and qz are 0/0
    assert np.array_equal(eager.get_state(), graph.get_state())
    eager.close()
    graph.close()


def test_graph_batch_equals_eager_batch(tmp_path):
    cases = [capi.Case(configs.build_instance(i, str(tmp_path), 0.3))
             for i in configs.batch_instances(4, base_seed=3000)]
    _say(f"[gpu] graph batch of {len(cases)}:", ", ".join(f"{c.n_cells}x{c.n_vel}" for c in cases))
    eager, graph = [], []
    for cs in cases:
        e, g = capi.Device(cs, capi.STORAGE_FP32), capi.Device(cs, capi.STORAGE_FP32)
        g.set_graphs(True)
        e.init_equilibrium()
        g.init_equilibrium()
        eager.append(e)
        graph.append(g)
    for n in (4, 3, 4):
        capi.step_batch(eager, n)
        reps = capi.step_batch(graph, n)
        assert all(r.steps_done == n for r in reps)
        for k, (e, g) in enumerate(zip(eager, graph)):
            same(e, g, f"instance {k} after {n}")
    for d in eager + graph:
        d.close()
