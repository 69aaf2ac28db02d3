"""Pins the numpy surrogate oracle (oracle/surrogate.py): descriptor, MLP, analytic position gradient and
training gradients against torch autograd (FP64, CPU) and central finite differences; bf16 rounding
against torch's bfloat16 conversion."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import surrogate as S  # noqa: E402

RC = 2.5


def _config(N=64, rho=0.6, seed=1):
    L = (N / rho) ** (1 / 3)
    rng = np.random.default_rng(seed)
    n = int(np.ceil(N ** (1 / 3)))
    g = np.stack(np.meshgrid(*[np.arange(n)] * 3, indexing="ij"), -1).reshape(-1, 3)[:N]
    x = (g + 0.5) * (L / n) + rng.uniform(-0.1, 0.1, (N, 3))
    return x % L, L


def _torch_energy(x, L, p):
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    d = xt[:, None, :] - xt[None, :, :]
    d = d - L * torch.round(d / L)
    r2 = (d ** 2).sum(-1) + torch.eye(len(x), dtype=torch.float64) * 1e6
    r = torch.sqrt(r2)
    mask = (r < RC).double()
    mu = torch.tensor(S.MU)
    e = torch.exp(-((r[..., None] - mu) ** 2) / (2 * S.WIDTH ** 2))
    fc = 0.5 * (torch.cos(np.pi * r / RC) + 1) * mask
    G = (e * fc[..., None]).sum(1)
    T = {k: torch.tensor(v, dtype=torch.float64) for k, v in p.items()}
    z = (G - T["shift"]) * T["scale"]
    h1 = torch.nn.functional.silu(z @ T["W1"].T + T["b1"])
    h2 = torch.nn.functional.silu(h1 @ T["W2"].T + T["b2"])
    e_at = h2 @ T["w3"] + T["b3"][0]
    E = e_at.sum()
    E.backward()
    return E.item(), xt.grad.numpy(), G.detach().numpy()


def test_descriptor_energy_and_position_gradient_vs_torch():
    x, L = _config()
    p = S.init_params(3)
    p["scale"] = np.full(32, 0.5)
    p["shift"] = np.full(32, 0.3)
    E, grad, _ = S.energy_grad(x, L, RC, p)
    Et, gt, Gt = _torch_energy(x, L, p)
    assert np.allclose(S.descriptors(x, L, RC), Gt, rtol=1e-12, atol=1e-12)
    assert E == pytest.approx(Et, rel=1e-12)
    assert np.allclose(grad, gt, rtol=1e-9, atol=1e-10)


def test_position_gradient_finite_difference():
    x, L = _config(N=27, rho=0.5, seed=2)
    p = S.init_params(4)
    _, grad, _ = S.energy_grad(x, L, RC, p)
    h = 1e-6
    for i, d in [(0, 0), (5, 1), (26, 2)]:
        xp, xm = x.copy(), x.copy()
        xp[i, d] += h
        xm[i, d] -= h
        num = (S.energy_grad(xp, L, RC, p)[0] - S.energy_grad(xm, L, RC, p)[0]) / (2 * h)
        assert num == pytest.approx(grad[i, d], rel=1e-6, abs=1e-8)


def test_training_gradients_vs_torch():
    x, L = _config(N=64, seed=5)
    G = S.descriptors(x, L, RC)
    u = S.lj_atom_energies(x, L, RC)
    p = S.init_params(7)
    loss, grads = S.loss_grad(p, G, u)
    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k in grads) for k, v in p.items()}
    z = (torch.tensor(G) - T["shift"]) * T["scale"]
    h1 = torch.nn.functional.silu(z @ T["W1"].T + T["b1"])
    h2 = torch.nn.functional.silu(h1 @ T["W2"].T + T["b2"])
    e = h2 @ T["w3"] + T["b3"][0]
    lt = torch.mean((e - torch.tensor(u)) ** 2)
    lt.backward()
    assert loss == pytest.approx(lt.item(), rel=1e-12)
    for k, g in grads.items():
        assert np.allclose(g, T[k].grad.numpy(), rtol=1e-9, atol=1e-12), k


def test_lj_atom_energies_sum_to_potential(orc):
    x, L = _config(N=125, rho=0.7, seed=9)
    u = S.lj_atom_energies(x, L, RC)
    assert u.sum() == pytest.approx(orc.potential_all(x, [L] * 3, RC), rel=1e-12)


def test_bf16_rounding_matches_torch():
    a = np.random.default_rng(0).normal(size=1000) * 10
    assert np.array_equal(S.bf16(a), torch.tensor(a, dtype=torch.float32).bfloat16().double().numpy())


def test_adam_reduces_loss():
    x, L = _config(N=125, rho=0.7, seed=11)
    G = S.descriptors(x, L, RC)
    u = S.lj_atom_energies(x, L, RC)
    p = S.init_params(1)
    p["shift"] = G.mean(0)
    p["scale"] = 1 / (G.std(0) + 1e-3)
    state = {}
    l0, _ = S.loss_grad(p, G, u)
    for _ in range(300):
        _, g = S.loss_grad(p, G, u)
        S.adam_step(p, g, state, lr=1e-2)
    l1, _ = S.loss_grad(p, G, u)
    assert l1 < 0.1 * l0


def test_adam_matches_torch_optimizer():
    """Three Adam steps of the oracle against torch.optim.Adam (same betas, eps, lr) on a fixed
    gradient sequence; the first step is also the closed form -lr * g / (|g| + eps)."""
    rng = np.random.default_rng(3)
    w0 = rng.normal(size=(4, 5))
    grads = [rng.normal(size=(4, 5)) for _ in range(3)]
    p = {"W": w0.copy()}
    state = {}
    t = torch.tensor(w0.copy(), dtype=torch.float64, requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-3, betas=(0.9, 0.999), eps=1e-8)
    for k, g in enumerate(grads):
        S.adam_step(p, {"W": g}, state, lr=1e-3)
        opt.zero_grad()
        t.grad = torch.tensor(g, dtype=torch.float64)
        opt.step()
        assert np.allclose(p["W"], t.detach().numpy(), rtol=0, atol=1e-15)
        if k == 0:
            assert np.allclose(p["W"], w0 - 1e-3 * grads[0] / (np.abs(grads[0]) + 1e-8), atol=1e-15)


def _torch_force_loss(configs, L, p, w_force):
    """The energy + force-matching loss built in torch (FP64), forces by autograd with create_graph, so
    that its parameter gradient contains the mixed derivative d(dE/dx)/dtheta by double backward."""
    T = {k: torch.tensor(v, dtype=torch.float64, requires_grad=k not in ("shift", "scale")) for k, v in p.items()}
    mu = torch.tensor(S.MU)
    n = sum(len(x) for x in configs)
    le = lf = 0.0
    for x in configs:
        xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
        d = xt[:, None, :] - xt[None, :, :]
        d = d - L * torch.round(d / L)
        r2 = (d ** 2).sum(-1) + torch.eye(len(x), dtype=torch.float64) * 1e6
        r = torch.sqrt(r2)
        mask = (r < RC).double()
        e = torch.exp(-((r[..., None] - mu) ** 2) / (2 * S.WIDTH ** 2))
        fc = 0.5 * (torch.cos(np.pi * r / RC) + 1) * mask
        G = (e * fc[..., None]).sum(1)
        z = (G - T["shift"]) * T["scale"]
        h1 = torch.nn.functional.silu(z @ T["W1"].T + T["b1"])
        h2 = torch.nn.functional.silu(h1 @ T["W2"].T + T["b2"])
        e_at = h2 @ T["w3"] + T["b3"][0]
        gradE, = torch.autograd.grad(e_at.sum(), xt, create_graph=True)
        u = torch.tensor(S.lj_atom_energies(x, L, RC))
        R = gradE - torch.tensor(S.lj_forces(x, L, RC))
        le = le + ((e_at - u) ** 2).sum()
        lf = lf + (R ** 2).sum()
    loss = le / n + w_force * lf / (3 * n)
    loss.backward()
    return (le / n).item(), (lf / (3 * n)).item(), {k: T[k].grad.numpy() for k in ("W1", "b1", "W2", "b2", "w3", "b3")}


def test_force_matching_gradient_vs_torch_double_backward():
    """The analytic force-matching gradient (forward-mode tangent + reverse pass) against torch's double
    backward on two configurations; a dropped tangent term or sign fails it."""
    cfgs = [_config(N=64, rho=0.6, seed=s)[0] for s in (12, 13)]
    L = _config(N=64, rho=0.6, seed=12)[1]
    p = S.init_params(8)
    p["shift"] = np.full(32, 0.2)
    p["scale"] = np.linspace(0.5, 1.5, 32)
    p["b1"] = np.random.default_rng(2).normal(0, 0.3, 64)
    p["b2"] = np.random.default_rng(3).normal(0, 0.3, 64)
    me, mf, grads = S.force_loss_grad(p, cfgs, L, RC, w_force=0.7)
    te, tf, tg = _torch_force_loss(cfgs, L, p, 0.7)
    assert me == pytest.approx(te, rel=1e-12) and mf == pytest.approx(tf, rel=1e-10)
    for k, g in grads.items():
        assert np.allclose(g, tg[k], rtol=1e-8, atol=1e-11), (k, np.abs(g - tg[k]).max())
    # without the force term it is the energy-only gradient
    _, _, g0 = S.force_loss_grad(p, cfgs, L, RC, w_force=0.0)
    G = np.concatenate([S.descriptors(x, L, RC) for x in cfgs])
    u = np.concatenate([S.lj_atom_energies(x, L, RC) for x in cfgs])
    _, ge = S.loss_grad(p, G, u)
    for k in ge:
        assert np.allclose(g0[k], ge[k], rtol=1e-12, atol=1e-14)


def test_lj_forces_finite_difference(orc):
    x, L = _config(N=125, rho=0.7, seed=14)          # L >= 2 r_c: minimum image = all images
    f = S.lj_forces(x, L, RC)
    assert np.allclose(f, orc.force_all(x, [L] * 3, RC), rtol=1e-10, atol=1e-10)


def test_tree_pairs_equal_dense_pairs():
    """The k-d tree pair search (used above 2000 atoms) against the dense minimum-image table."""
    for N, rho, seed in ((1000, 0.8442, 21), (2197, 0.5, 22)):
        x, L = _config(N=N, rho=rho, seed=seed)
        a = S.pairs_dense(x, L, RC)
        b = S.pairs_tree(x, L, RC)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
        assert np.allclose(a[2], b[2], rtol=0, atol=1e-12) and np.allclose(a[3], b[3], rtol=0, atol=1e-12)
