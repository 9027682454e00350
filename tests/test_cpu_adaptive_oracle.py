"""CPU pins of the oracle's adaptive time stepper (TransientAdaptive, transient_adaptive.rs:81-350)
on transient heat: fixed mode takes exactly t_end / dt_init steps; every output time is hit
exactly; the error against a fine fixed-step solution falls as rtol falls; a smaller rtol takes
more steps."""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from problems import heat_c3


@pytest.fixture(scope="module")
def heat5():
    m = vfvm.Mesh.lattice(5, 0, 0.3)
    k = np.exp(0.5 * np.sin(3 * m.get("pos")[:, 0]))
    p = heat_c3(m, k)
    g = O.dual(m.get("pos"), m.get("tets"), m.num_pts)
    scl, sb, flux, bn, ops = p.oracle_inputs(g)
    sysf = lambda x, xp, dt: O._csr_system(g, ops, scl, sb, flux, bn, p.unknown, xp, dt)
    x0 = O.apply_initial_dirichlet(g, ops, sb, bn, np.zeros(p.N))
    return sysf, x0


def test_fixed_steps_and_output_times(heat5):
    sysf, x0 = heat5
    x, st, steps = O.transient_adaptive(sysf, x0, 0.05, 1e-3, fixed=True)
    assert st["accepted"] == 50 and st["rejected_error"] == 0 and st["outputs"] == 10
    assert np.allclose([s[1] for s in steps], 1e-3, rtol=1e-9)
    outs = [0.05 * i / 10 for i in range(1, 11)]
    x, st, steps = O.transient_adaptive(sysf, x0, 0.05, 3e-4, rtol=1e-3, out_times=outs)
    ts = [s[0] for s in steps]
    for t in outs:
        assert min(abs(t - u) for u in ts) == 0.0
    assert ts[-1] == 0.05


def test_error_falls_with_rtol(heat5):
    sysf, x0 = heat5
    ref, _, _ = O.transient_adaptive(sysf, x0, 0.05, 1e-4, fixed=True)
    errs, counts = [], []
    for rtol in (1e-2, 1e-3, 1e-4):
        x, st, steps = O.transient_adaptive(sysf, x0, 0.05, 1e-4, rtol=rtol)
        errs.append(np.abs(x - ref).max())
        counts.append(st["accepted"])
    assert errs[0] > errs[1] > errs[2]
    assert counts[0] < counts[1] < counts[2]
