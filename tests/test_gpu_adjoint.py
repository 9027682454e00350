"""Discrete adjoint on the GPU (vfvm_adjoint_forward / vfvm_adjoint_backward through the
C ABI) against the oracle's adjoint (sparse LU, pinned by finite differences in
tests/test_cpu_adjoint_oracle.py) and against central finite differences of the GPU's
own forward re-solves.

Bars: gradients within 1e-8 (max-abs, relative to the largest entry) of the oracle's
with FP64 solves at 1e-13; bit-identical gradients for every checkpoint interval (the
segment replays reproduce the forward states bit for bit); central differences of the
GPU forward agree to 1e-5 relative; FP32-inner refined solves within 1e-8 of FP64.
"""
import numpy as np
import pytest

import vfvm
import vfvm_oracle as O
from vfvm.fields import lognormal_conductivity
from problems import X0, X1, Y0, face_mean_weights, heat_c3, mass_c2

pytestmark = pytest.mark.gpu

LIN = dict(tol=1e-13, restart=60, max_outer=400, precond=vfvm.PRECOND_ILU0)


@pytest.fixture(scope="module")
def ctx():
    c = vfvm.Context(0)
    yield c
    c.close()


def _heat(n=8, seed=3):
    mesh = vfvm.Mesh.lattice(n, seed=seed)
    pos = mesh.get("pos")
    k = lognormal_conductivity(pos, ell=0.3, modes=64, seed=1)
    q = 0.5 + 0.25 * np.sin(3.0 * pos[:, 1])
    wt = 1.0 + 0.2 * pos[:, 2]
    return heat_c3(mesh, k, w=wt, q=q)


def _rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def _oracle_grads(p, w, T0, dt, num_ts):
    g = O.dual(p.pos, p.mesh.get("tets"), p.mesh.num_pts)
    scl, sb, flux, bnd_nid, ops = p.oracle_inputs(g)
    states = O.linear_map(g, ops, scl, sb, flux, bnd_nid, p.unknown, T0, dt, num_ts)
    acc = O.adjoint_accumulate(g, ops, scl, sb, flux, bnd_nid, p.unknown, states, dt, num_ts, w)
    out = {"state": states[-1]}
    for name, fid in (("k", p.k_id), ("q", p.q_id), ("w", p.w_id)):
        out[name] = O.adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, dt, fid, constant=False)
    out["tin"] = O.adjoint_dirichlet_gradient(g, ops, bnd_nid, acc, p.tin_id)
    out["init"] = O.adjoint_initial_gradient(g, ops, bnd_nid, acc, num_ts)
    return out


def _gpu_grads(ctx, p, w, T0, dt, num_ts, every, lin=LIN, lin_adj=LIN):
    p.to_device(ctx)
    ctx.scldom_set(p.unknown, T0)
    fst = ctx.adjoint_forward(p.ops, dt=dt, num_ts=num_ts, checkpoint_every=every, lin=lin)
    state = ctx.scldom_get(p.unknown)
    wrt = [(vfvm.WRT_SCLDOM, p.k_id), (vfvm.WRT_SCLDOM, p.q_id), (vfvm.WRT_SCLDOM, p.w_id),
           (vfvm.WRT_SCLBND, p.tin_id), (vfvm.WRT_INITIAL, 0)]
    grads, st = ctx.adjoint_backward(w, wrt, lin=lin_adj)
    assert np.array_equal(ctx.scldom_get(p.unknown), state)  # final state restored
    return dict(state=state, k=grads[0], q=grads[1], w=grads[2], tin=grads[3], init=grads[4]), fst, st


@pytest.mark.parametrize("every", [2, 1, 10])
def test_transient_adjoint_vs_oracle(ctx, every):
    print(f"[adjoint transient n=8 every={every}]", flush=True)
    p = _heat()
    w = face_mean_weights(p.mesh, X1)
    T0 = 0.3 * np.cos(2.0 * p.pos[:, 0])
    dt, num_ts = 0.02, 6
    got, fst, st = _gpu_grads(ctx, p, w, T0, dt, num_ts, every)
    ref = _oracle_grads(p, w, T0, dt, num_ts)
    assert _rel(got["state"], ref["state"]) < 1e-10
    for key in ("k", "q", "w", "tin", "init"):
        r = _rel(got[key], ref[key])
        print(f"  {key}: rel err {r:.2e}")
        assert r < 1e-8, key
    # segments replayed backwards from the last but one; each copies its last state from the next
    # checkpoint, but the first one replayed runs to its end to compare it with that checkpoint
    expected_replays = {2: 3, 1: 1, 10: 0}[every]
    assert st.recomputed_steps == expected_replays
    assert st.replay_mismatches == 0
    assert st.steps == num_ts and st.checkpoints == num_ts // every + 1


def test_checkpoint_interval_bitwise(ctx):
    print("[adjoint checkpoint intervals]", flush=True)
    p = _heat()
    w = face_mean_weights(p.mesh, X1)
    T0 = np.zeros(p.N)
    a, _, _ = _gpu_grads(ctx, p, w, T0, 0.02, 7, 3)
    b, _, _ = _gpu_grads(ctx, p, w, T0, 0.02, 7, 100)
    for key in a:
        assert np.array_equal(a[key], b[key]), key


def test_steady_adjoint_vs_oracle(ctx):
    print("[adjoint steady heat]", flush=True)
    p = _heat()
    p.ops = [o for o in p.ops if o[0] not in (vfvm.OP_SCL_DOM_TIME, vfvm.OP_SCL_BND_TIME)]
    w = face_mean_weights(p.mesh, X1)
    T0 = np.zeros(p.N)
    got, _, st = _gpu_grads(ctx, p, w, T0, 0.0, 0, 1)
    ref = _oracle_grads(p, w, T0, 0.0, 0)
    for key in ("k", "q", "tin"):
        assert _rel(got[key], ref[key]) < 1e-8, key
    assert np.all(got["w"] == 0.0) and np.all(got["init"] == 0.0)
    assert st.recomputed_steps == 0


def test_advection_adjoint_vs_oracle(ctx):
    print("[adjoint steady advection]", flush=True)
    mesh = vfvm.Mesh.lattice(8, seed=5)
    D = 1e-2 * np.exp(0.5 * np.sin(4.0 * mesh.get("pos")[:, 0] + mesh.get("pos")[:, 2]))
    p = mass_c2(mesh, D=D, c_out=0.0)
    w = np.random.default_rng(2).uniform(size=p.N) / p.N
    p.to_device(ctx)
    ctx.adjoint_forward(p.ops, lin=LIN)
    (gd, gc), _ = ctx.adjoint_backward(w, [(vfvm.WRT_SCLDOM, p.d_id), (vfvm.WRT_SCLBND, p.cin_id)], lin=LIN)
    g = O.dual(p.pos, mesh.get("tets"), mesh.num_pts)
    scl, sb, flux, bnd_nid, ops = p.oracle_inputs(g)
    states = O.linear_map(g, ops, scl, sb, flux, bnd_nid, p.unknown, np.zeros(p.N), 0.0, 0)
    acc = O.adjoint_accumulate(g, ops, scl, sb, flux, bnd_nid, p.unknown, states, 0.0, 0, w)
    rd = O.adjoint_field_gradient(g, ops, scl, flux, bnd_nid, acc, 0.0, p.d_id, constant=False)
    rc = O.adjoint_dirichlet_gradient(g, ops, bnd_nid, acc, p.cin_id)
    assert _rel(gd, rd) < 1e-8
    assert _rel(gc, rc) < 1e-8


def test_gpu_central_differences(ctx):
    """dJ/dk from the adjoint against central differences of the GPU forward map."""
    print("[adjoint vs GPU finite differences]", flush=True)
    p = _heat(n=10)
    w = face_mean_weights(p.mesh, X1)
    T0 = np.zeros(p.N)
    dt, num_ts = 0.02, 5
    got, _, _ = _gpu_grads(ctx, p, w, T0, dt, num_ts, 2)
    kind, k = p.scl[p.k_id]
    NP = p.mesh.num_pts
    verts = [int(np.argmax(np.abs(got["k"]))), NP // 2, NP + int(p.mesh.bnd_nodes(X1)[10])]
    for v in verts:
        d = 1e-5 * k[v]
        J = []
        for s in (+1, -1):
            kk = k.copy()
            kk[v] += s * d
            ctx.scldom_set(p.k_id, kk)
            ctx.scldom_set(p.unknown, T0)
            ctx.adjoint_forward(p.ops, dt=dt, num_ts=num_ts, checkpoint_every=num_ts, lin=LIN)
            J.append(w @ ctx.scldom_get(p.unknown))
        fd = (J[0] - J[1]) / (2 * d)
        print(f"  v={v}: adjoint {got['k'][v]:.9e} fd {fd:.9e}")
        assert abs(got["k"][v] - fd) <= 1e-5 * np.max(np.abs(got["k"]))


def test_mixed_precision_adjoint(ctx):
    print("[adjoint mixed precision]", flush=True)
    p = _heat()
    w = face_mean_weights(p.mesh, X1)
    T0 = np.zeros(p.N)
    lin32 = dict(LIN, tol=1e-12, precision=vfvm.FP32_REFINED)
    lin64 = dict(LIN, tol=1e-12)
    a, _, _ = _gpu_grads(ctx, p, w, T0, 0.02, 6, 2, lin=lin32, lin_adj=lin32)
    b, _, _ = _gpu_grads(ctx, p, w, T0, 0.02, 6, 2, lin=lin64, lin_adj=lin64)
    for key in ("k", "q", "w", "tin", "init"):
        assert _rel(a[key], b[key]) < 1e-8, key


def test_torch_autograd(ctx):
    print("[adjoint torch.autograd]", flush=True)
    import torch
    from vfvm.autograd import field_to_state
    p = _heat()
    p.to_device(ctx)
    w = face_mean_weights(p.mesh, X1)
    T0 = np.zeros(p.N)
    dt, num_ts = 0.02, 4
    k = torch.tensor(p.scl[p.k_id][1], dtype=torch.float64, device="cuda", requires_grad=True)
    wt = torch.tensor(w, dtype=torch.float64, device="cuda")
    T = field_to_state(k, ctx, p.k_id, p.ops, T0, dt=dt, num_ts=num_ts, every=2, lin=LIN, lin_adj=LIN)
    J = (wt * T).sum()
    J.backward()
    (direct,), _ = ctx.adjoint_backward(w, [(vfvm.WRT_SCLDOM, p.k_id)], lin=LIN)
    assert np.array_equal(k.grad.cpu().numpy(), direct)
    ref = _oracle_grads(p, w, T0, dt, num_ts)
    assert _rel(direct, ref["k"]) < 1e-8


def test_adjoint_errors(ctx):
    print("[adjoint errors]", flush=True)
    p = _heat(n=6)
    p.to_device(ctx)
    w = np.zeros(p.N)
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.adjoint_backward(w, [])
    assert e.value.code == vfvm._lib.ERR_INVALID_ARGUMENT
    bad = [o if o[0] != vfvm.OP_SCL_DOM_PRODUCTION else (o[0], o[1], o[2], p.k_id, o[4]) for o in p.ops]
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.adjoint_forward(bad, dt=0.02, num_ts=2)
    assert e.value.code == vfvm._lib.ERR_OPERATOR_UNSUPPORTED
    ctx.adjoint_forward(p.ops, dt=0.02, num_ts=2, lin=LIN)
    neu = [o for o in p.ops if o[0] == vfvm.OP_SCL_BND_NEUMANN][0][3]
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.adjoint_backward(w, [(vfvm.WRT_SCLBND, neu)])
    assert e.value.code == vfvm._lib.ERR_OPERATOR_UNSUPPORTED
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.adjoint_backward(w, [(vfvm.WRT_SCLDOM, p.unknown)])
    assert e.value.code == vfvm._lib.ERR_INVALID_ARGUMENT
    with pytest.raises(vfvm.VFVChemError) as e:
        ctx.adjoint_forward(p.ops, dt=0.02, num_ts=2, checkpoint_every=0)
    assert e.value.code == vfvm._lib.ERR_INVALID_ARGUMENT
