"""The hand-written adjoint of the FP64 step (pdipc_rollout / pdipc_backward)
against the CPU reference derivative -- the oracle run on forward-mode dual
numbers, which applies the chain rule mechanically to the restated step -- and
against central finite differences of the GPU forward, plus the
torch.autograd.Function harness over the C ABI."""

import numpy as np
import pytest
import torch

import oracle
import pdipc
import scenes
from pdipc_autograd import PdIpcRollout

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def _dense_solver(ref, diag, P):
    """The Chebyshev solve as dense scalar matrices: x_K = Px x0 + Pb b per coordinate."""
    H = ref.sys_matrix()
    R = -(H - np.diag(np.diag(H)))
    R[np.abs(R) < 1e-12] = 0
    n = len(diag)
    w, K, r = P["omega"], int(P["cheb_iters"]), P["rho"] ** 2
    J = (1 - w) * np.eye(n) + w * (R / diag[:, None])
    M = J @ J
    C = w * w * (R / diag[None, :]) / diag[:, None] + (2 - w) * w * np.diag(1 / diag)
    Px_prev, Pb_prev = np.eye(n), np.zeros((n, n))
    Px, Pb = M.copy(), C.copy()
    W = 2 / (2 - r)
    for _ in range(1, K):
        Px, Px_prev = W * (M @ Px - Px_prev) + Px_prev, Px
        Pb, Pb_prev = W * (M @ Pb + C - Pb_prev) + Pb_prev, Pb
        W = 4 / (4 - r * W)
    return Px, Pb


def test_solve_vjp_matches_dense_transpose():
    """One Chebyshev solve's VJP (the reverse sweeps on D M D^-1 and the first
    term's adjoint) against the transposed dense linear map."""
    sc = scenes.cloth_over_sphere(6, 1, ground=False, seed=0, patch_dim=6, sphere_patch_dim=4)
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    n = ref.sim
    rng = np.random.default_rng(1)
    diag = np.diag(ref.sys_matrix()) * (1 + rng.uniform(0, 0.5, n))  # as if collision terms were added
    rhs = rng.normal(size=(n, 3))
    x0 = ref.x().copy()
    x0[:n] += rng.normal(0, 1e-3, (n, 3))
    a = rng.normal(size=(n, 3))
    print("case: solve VJP", flush=True)
    xb, bb = sim.debug_solve_vjp(diag, rhs, x0, a)
    Px, Pb = _dense_solver(ref, diag, sc["params"])
    assert _rel(xb, Px.T @ a) < 1e-10
    assert _rel(bb, Pb.T @ a) < 1e-10


@pytest.mark.parametrize("kind", ["cloth", "tets"])
def test_local_vjp_matches_oracle_jacobian(kind):
    """The local step's VJP (springs, bending rescale, SVD strain projection)
    against central differences of the oracle's local step."""
    if kind == "cloth":
        sc = scenes.cloth_over_sphere(6, 1, ground=False, seed=0, patch_dim=6, sphere_patch_dim=4)
    else:
        sc = scenes.config2(cells=2, cloth_n=4)
    ref = oracle.Oracle(sc)
    sim = pdipc.Simulator.from_scene(sc)
    n = ref.sim
    rng = np.random.default_rng(2)
    x = ref.x().copy()
    x[:n] += rng.normal(0, 3e-3, (n, 3))  # stretch, bend and (for tets) leave the clamp interval
    init = np.zeros((n, 3))
    bbar = rng.normal(size=(n, 3))
    print(f"case: local VJP {kind}", flush=True)
    got = sim.debug_local_vjp(x, bbar)
    want = np.zeros((n, 3))
    h = 1e-7
    for i in range(n):
        for k in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, k] += h
            xm[i, k] -= h
            want[i, k] = np.sum(bbar * (ref.debug_local(xp, init) - ref.debug_local(xm, init))) / (2 * h)
    assert _rel(got, want) < 1e-5


def _cloth_case(seed=0, n=12, steps=15):
    sc, v0, target = scenes.config3(seed, n_cloth=n, steps=steps)
    return sc, v0, target


def _directions(sc, v0, seed):
    K = oracle.dual_dirs()
    rng = np.random.default_rng(seed)
    k = sc["cloths"][0]["tri_k"]
    dk = rng.normal(size=(K, len(k))) * k  # relative stiffness directions
    dv = rng.normal(size=(K,) + v0.shape) * 0.05
    return dk, dv


def _dual_reference(sc, v0, target, steps, dk, dv):
    d = oracle.Oracle(sc, dual=True, tri_k_tangents={0: dk})
    d.dual_set("v", v0, dv)
    d.step(steps)
    xo, xt = d.dual_get("x")
    gl = 2.0 * (xo - target)
    return xo, np.array([np.sum(gl * xt[q]) for q in range(len(dk))])


@pytest.mark.parametrize("seed,lanes", [(0, 8), (1, 8), (0, 4), (0, 1)])
def test_cloth_adjoint_matches_dual_oracle(seed, lanes):
    """dL/dk and dL/dv0 of L = |x_T - target|^2 after 15 contact steps of a
    12 x 12 cloth over the obstacle, projected on 8 random directions, against
    the dual-number oracle's directional derivatives (8, 4 and 1 lanes per SELL
    row: 8, 16 and 64 rows per slice in the reverse pass)."""
    steps = 15
    sc, v0, target = _cloth_case(seed, steps=steps)
    dk, dv = _directions(sc, v0, 100 + seed)
    print(f"case: cloth adjoint vs dual oracle, seed {seed}, {lanes} lanes", flush=True)
    xo, want = _dual_reference(sc, v0, target, steps, dk, dv)
    sim = pdipc.Simulator.from_scene(sc, sell_lanes=lanes)
    sim.set_v(v0)
    sim.rollout(steps, 4)
    xg = sim.x()[: sim.sim]
    assert _rel(xg, xo) < 1e-8
    ntri = len(sc["cloths"][0]["f"])
    dx0, dv0, dkg = sim.backward(dx=2.0 * (xg - target), n_tris=ntri)
    got = np.array([dkg @ dk[q] + np.sum(dv0 * dv[q]) for q in range(len(dk))])
    print("adjoint", got, "\ndual   ", want, "\nstats", sim.adjoint_stats(), flush=True)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()
    # the handle is left at the final state and a second pass repeats the first
    # (to the rounding of the FP64 atomics that gather the barrier adjoint)
    assert np.array_equal(sim.x()[: sim.sim], xg)
    _, dv0b, dkb = sim.backward(dx=2.0 * (xg - target), n_tris=ntri)
    assert np.abs(dkb - dkg).max() <= 1e-7 * np.abs(dkg).max()
    assert np.abs(dv0b - dv0).max() <= 1e-7 * np.abs(dv0).max()


def test_cloth_adjoint_central_differences():
    """Directional derivatives against central differences (eps = 1e-6) of the
    GPU forward itself on 8 random directions."""
    steps = 12
    sc, v0, target = _cloth_case(2, steps=steps)
    dk, dv = _directions(sc, v0, 7)
    k0 = sc["cloths"][0]["tri_k"].copy()
    sim = pdipc.Simulator.from_scene(sc)
    x0 = sim.x()[: sim.sim].copy()

    def loss(k, v):
        sim.set_cloth_stiffness(0, k)
        sim.set_x(x0)
        sim.set_v(v)
        sim.rollout(steps, 5)
        x = sim.x()[: sim.sim]
        return float(np.sum((x - target) ** 2)), x

    print("case: cloth adjoint vs central differences", flush=True)
    _, xT = loss(k0, v0)
    _, dv0, dkg = sim.backward(dx=2.0 * (xT - target), n_tris=len(k0))
    errs = []
    eps = 1e-6
    for q in range(len(dk)):
        lp, _ = loss(k0 + eps * dk[q], v0 + eps * dv[q])
        lm, _ = loss(k0 - eps * dk[q], v0 - eps * dv[q])
        fd = (lp - lm) / (2 * eps)
        ad = dkg @ dk[q] + np.sum(dv0 * dv[q])
        errs.append(abs(fd - ad) / max(abs(ad), 1e-30))
        print(f"direction {q}: fd {fd:.9e} adjoint {ad:.9e} rel {errs[-1]:.2e}", flush=True)
    assert np.median(errs) < 1e-4
    assert np.sum(np.array(errs) < 1e-3) >= len(errs) - 1  # one direction may cross a contact switch


def test_autograd_function_gradcheck():
    """torch.autograd.Function over the C ABI: torch's own gradcheck on a
    two-parameter family k = k0 (1 + a dk), v = v0 + b dv."""
    steps = 8
    sc, v0, target = _cloth_case(3, steps=steps)
    dk, dv = _directions(sc, v0, 9)
    k0 = torch.tensor(sc["cloths"][0]["tri_k"], dtype=torch.float64)
    sim = pdipc.Simulator.from_scene(sc)
    x0 = sim.x()[: sim.sim].copy()
    tgt = torch.tensor(target, dtype=torch.float64)
    dkt = torch.tensor(dk[0] / sc["cloths"][0]["tri_k"], dtype=torch.float64)
    dvt = torch.tensor(dv[0], dtype=torch.float64)
    v0t = torch.tensor(v0, dtype=torch.float64)

    def fn(a, b):
        xT = PdIpcRollout.apply(k0 * (1 + a * dkt), v0t + b * dvt, sim, x0, steps, 3)
        return ((xT - tgt) ** 2).sum()

    print("case: autograd.Function gradcheck", flush=True)
    a = torch.zeros((), dtype=torch.float64, requires_grad=True)
    b = torch.zeros((), dtype=torch.float64, requires_grad=True)
    # the backward pass through the Function is the library's adjoint, bit for bit
    loss = fn(a, b)
    loss.backward()
    _, dv0, dk = sim.backward(dx=2.0 * (sim.x()[: sim.sim] - target), n_tris=len(k0))
    assert abs(a.grad.item() - float(dk @ (k0.numpy() * dkt.numpy()))) <= 1e-7 * abs(a.grad.item())
    assert abs(b.grad.item() - float(np.sum(dv0 * dv[0]))) <= 1e-7 * abs(b.grad.item())
    # torch's own central differences (eps = 1e-6) at the FD truncation level
    assert torch.autograd.gradcheck(fn, (a.detach().requires_grad_(), b.detach().requires_grad_()), eps=1e-6,
                                    atol=1e-9, rtol=2e-3, nondet_tol=1e-10)


def test_tet_adjoint_matches_dual_oracle():
    """The SVD strain projection's derivative: a tet block with a cloth over it,
    dL/dv0 on 8 random directions against the dual oracle."""
    steps = 6
    sc = scenes.config2(cells=3, cloth_n=8)
    rng = np.random.default_rng(4)
    ref = oracle.Oracle(sc)
    nsim = ref.sim
    v0 = rng.normal(0, 0.05, (nsim, 3))
    target = ref.x()[:nsim] + rng.normal(0, 0.01, (nsim, 3))
    K = oracle.dual_dirs()
    dv = rng.normal(size=(K, nsim, 3)) * 0.05
    print("case: tet adjoint vs dual oracle", flush=True)
    d = oracle.Oracle(sc, dual=True)
    d.dual_set("v", v0, dv)
    d.step(steps)
    xo, xt = d.dual_get("x")
    want = np.array([np.sum(2 * (xo - target) * xt[q]) for q in range(K)])
    sim = pdipc.Simulator.from_scene(sc)
    sim.set_v(v0)
    sim.rollout(steps, 3)
    xg = sim.x()[:nsim]
    assert _rel(xg, xo) < 1e-8
    _, dv0, _ = sim.backward(dx=2.0 * (xg - target))
    got = np.array([np.sum(dv0 * dv[q]) for q in range(K)])
    print("adjoint", got, "\ndual   ", want, flush=True)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_adjoint_graph_replay_matches_direct_launches():
    """The recorded forward and the reverse local/global loops replayed from HIP
    graphs give the direct launches' trajectory bit for bit and the same
    gradient (to the rounding of the adjoint's FP64 atomics); two backward
    passes in a row on the graph path agree as well."""
    steps = 12
    sc, v0, target = _cloth_case(4, steps=steps)
    ntri = len(sc["cloths"][0]["f"])
    out = {}
    for graphs in (0, 1):
        print(f"case: adjoint with use_graphs={graphs}", flush=True)
        sim = pdipc.Simulator.from_scene(sc, use_graphs=graphs)
        sim.set_v(v0)
        sim.rollout(steps, 5)
        xT = sim.x()[: sim.sim].copy()
        _, dv0, dk = sim.backward(dx=2.0 * (xT - target), n_tris=ntri)
        _, dv0b, dkb = sim.backward(dx=2.0 * (xT - target), n_tris=ntri)
        assert np.abs(dkb - dk).max() <= 1e-7 * np.abs(dk).max()
        out[graphs] = (xT, dv0, dk)
    assert np.array_equal(out[0][0], out[1][0])
    assert np.abs(out[0][1] - out[1][1]).max() <= 1e-7 * np.abs(out[0][1]).max()
    assert np.abs(out[0][2] - out[1][2]).max() <= 1e-7 * np.abs(out[0][2]).max()


def test_recorded_rollout_matches_checkpoints():
    """checkpoint_every = 0 keeps the rollout's own tape: the backward runs no
    forward step, gives the checkpointed backward's gradient (the re-run
    segments record the same trajectory) and can be repeated.  A checkpointed
    rollout records its last segment as it runs: the first backward reverses
    that tape, a second one re-runs the segment (the same trajectory bit for
    bit), and both give the same gradient up to the rounding of the FP64
    atomics that gather the barrier adjoint."""
    steps = 9
    sc, v0, target = _cloth_case(1, steps=steps)
    k0 = sc["cloths"][0]["tri_k"].copy()
    sim = pdipc.Simulator.from_scene(sc)
    x0 = sim.x()[: sim.sim].copy()
    out = {}
    for every in (4, 0):
        print(f"case: rollout with checkpoint_every = {every}", flush=True)
        sim.set_cloth_stiffness(0, k0)
        sim.set_x(x0)
        sim.set_v(v0)
        sim.rollout(steps, every)
        xT = sim.x()[: sim.sim].copy()
        out[every] = (xT, sim.backward(dx=2.0 * (xT - target), n_tris=len(k0)))
        if every:
            for a, b in zip(sim.backward(dx=2.0 * (xT - target), n_tris=len(k0)), out[every][1]):
                assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()
    (xa, (dxa, dva, dka)), (xb, (dxb, dvb, dkb)) = out[4], out[0]
    assert np.array_equal(xa, xb)
    for a, b in ((dxa, dxb), (dva, dvb), (dka, dkb)):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    again = sim.backward(dx=2.0 * (xb - target), n_tris=len(k0))
    for a, b in zip(again, (dxb, dvb, dkb)):
        assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max()


def test_adjoint_after_stiffness_change_matches_dual_oracle():
    """A rollout and backward at one stiffness field, then a stiffness change
    on the same handle and a second rollout and backward: the second gradient
    must follow the new stiffness (the reverse pass's copies of the M^2
    2-path products are refreshed), checked against the dual-number oracle
    run at the new stiffness."""
    steps = 10
    sc, v0, target = _cloth_case(5, steps=steps)
    k0 = sc["cloths"][0]["tri_k"].copy()
    rng = np.random.default_rng(11)
    k1 = k0 * np.exp(rng.uniform(-1.0, 1.0, len(k0)))  # a different field of the same pattern
    ntri = len(k0)
    sim = pdipc.Simulator.from_scene(sc)
    x0 = sim.x()[: sim.sim].copy()
    print("case: adjoint at k0, then at k1 on the same handle", flush=True)
    sim.set_v(v0)
    sim.rollout(steps, 5)
    xa = sim.x()[: sim.sim].copy()
    sim.backward(dx=2.0 * (xa - target), n_tris=ntri)
    sim.set_cloth_stiffness(0, k1)
    sim.set_x(x0)
    sim.set_v(v0)
    sim.rollout(steps, 5)
    xg = sim.x()[: sim.sim].copy()
    _, dv0, dkg = sim.backward(dx=2.0 * (xg - target), n_tris=ntri)
    sc1 = dict(sc)
    sc1["cloths"] = [dict(sc["cloths"][0], tri_k=k1)] + list(sc["cloths"][1:])
    dk, dv = _directions(sc1, v0, 21)
    xo, want = _dual_reference(sc1, v0, target, steps, dk, dv)
    assert _rel(xg, xo) < 1e-8
    got = np.array([dkg @ dk[q] + np.sum(dv0 * dv[q]) for q in range(len(dk))])
    print("adjoint", got, "\ndual   ", want, flush=True)
    assert np.abs(got - want).max() <= 1e-6 * np.abs(want).max()


def test_staged_reverse_sweeps_match_l2_gather():
    """The reverse sweeps staged in LDS (the default) against the L2-gather
    ones: the same trajectory bit for bit and the same gradient (to the
    rounding of the barrier adjoint's FP64 atomics)."""
    steps = 10
    sc, v0, target = _cloth_case(6, steps=steps)
    ntri = len(sc["cloths"][0]["f"])
    out = {}
    for st in (1, 0):
        print(f"case: adjoint with sell_stage={st}", flush=True)
        sim = pdipc.Simulator.from_scene(sc, sell_stage=st)
        sim.set_v(v0)
        sim.rollout(steps, 5)
        xT = sim.x()[: sim.sim].copy()
        out[st] = (xT,) + sim.backward(dx=2.0 * (xT - target), n_tris=ntri)
    assert np.array_equal(out[0][0], out[1][0])
    for a, b in zip(out[1][1:], out[0][1:]):
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max()
