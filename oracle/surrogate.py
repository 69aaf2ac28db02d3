"""FP64 numpy oracle of the learned surrogate (TEST INFRASTRUCTURE ONLY; no reference counterpart).

Model (per atom i, neighbours j within r_c under the minimum image, L >= 2 r_c):
  descriptor  G_k(i) = sum_j exp(-(r_ij - mu_k)^2 / (2 w^2)) fc(r_ij),  fc(r) = (cos(pi r / r_c) + 1) / 2,
              mu_k uniform on [0.8, 2.5] (32 centres), w = 0.1
  input       z = (G - shift) * scale
  MLP         h1 = silu(W1 z + b1) (64), h2 = silu(W2 h1 + b2) (64), e_i = w3 . h2 + b3
  energy      E = sum_i e_i;  gradient dE/dx_i = sum_j [sum_k (g_k(i) + g_k(j)) phi'_k(r_ij)] (x_i - x_j) / r_ij
              with g(a) = dE/dG(a) and phi_k(r) = exp(...) fc(r)
  training    per-atom energy targets u_i = 2 eps sum_j sr6 (sr6 - 1) (the LJ energy split evenly over
              pairs, sum_i u_i = V), loss = mean_i (e_i - u_i)^2, gradients by reverse mode
bf16=True rounds weights and layer inputs to bfloat16 (round to nearest even) where the GPU kernels
store them, so that GPU and oracle differ only by FP32 summation order.
Pinned in tests/test_surrogate_oracle.py against torch autograd (FP64, CPU) and finite differences.
"""
from __future__ import annotations

import numpy as np

NFEAT, NH = 32, 64
MU = np.linspace(0.8, 2.5, NFEAT)
WIDTH = 0.1


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = {
        "W1": rng.normal(0, 1 / np.sqrt(NFEAT), (NH, NFEAT)), "b1": np.zeros(NH),
        "W2": rng.normal(0, 1 / np.sqrt(NH), (NH, NH)), "b2": np.zeros(NH),
        "w3": rng.normal(0, 1 / np.sqrt(NH), NH), "b3": np.zeros(1),
        "shift": np.zeros(NFEAT), "scale": np.ones(NFEAT),
    }
    return p


def pairs(x, L, rc):
    """All ordered pairs i != j with r_ij < rc under the minimum image (valid for L >= 2 rc), sorted by
    (i, j).  Above 2000 atoms the candidates come from a periodic k-d tree (scipy) instead of the dense
    N x N distance table; both give the same pairs (tested)."""
    if x.shape[0] > 2000:
        return pairs_tree(x, L, rc)
    return pairs_dense(x, L, rc)


def pairs_tree(x, L, rc):
    from scipy.spatial import cKDTree
    L = np.broadcast_to(np.asarray(L, float), (3,))
    xw = x - L * np.floor(x / L)
    xw[xw >= L] = 0.0                    # a coordinate that rounds up to L wraps to 0
    ij = cKDTree(xw, boxsize=L).query_pairs(rc * (1 + 1e-9), output_type="ndarray")
    i = np.concatenate([ij[:, 0], ij[:, 1]])
    j = np.concatenate([ij[:, 1], ij[:, 0]])
    d = x[i] - x[j]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    keep = r2 < rc * rc
    i, j, d, r2 = i[keep], j[keep], d[keep], r2[keep]
    order = np.lexsort((j, i))
    return i[order], j[order], d[order], np.sqrt(r2[order])


def pairs_dense(x, L, rc):
    L = np.broadcast_to(np.asarray(L, float), (3,))
    d = x[:, None, :] - x[None, :, :]
    d -= L * np.round(d / L)
    r2 = (d ** 2).sum(-1)
    np.fill_diagonal(r2, np.inf)
    i, j = np.nonzero(r2 < rc * rc)
    return i, j, d[i, j], np.sqrt(r2[i, j])


def basis(r, rc):
    """phi_k(r) and phi'_k(r), shape (npairs, 32)."""
    e = np.exp(-((r[:, None] - MU[None, :]) ** 2) / (2 * WIDTH ** 2))
    fc = 0.5 * (np.cos(np.pi * r / rc) + 1.0)
    dfc = -0.5 * np.pi / rc * np.sin(np.pi * r / rc)
    phi = e * fc[:, None]
    dphi = e * (-(r[:, None] - MU[None, :]) / WIDTH ** 2) * fc[:, None] + e * dfc[:, None]
    return phi, dphi


def descriptors(x, L, rc):
    i, j, d, r = pairs(x, L, rc)
    phi, _ = basis(r, rc)
    G = np.zeros((x.shape[0], NFEAT))
    np.add.at(G, i, phi)
    return G


def silu(a):
    return a / (1 + np.exp(-a))


def dsilu(a):
    s = 1 / (1 + np.exp(-a))
    return s * (1 + a * (1 - s))


def mlp(G, p, bf=False):
    """Per-atom energies and the cache of the forward pass."""
    q = (lambda a: bf16(a)) if bf else (lambda a: a)
    z = q((G - p["shift"]) * p["scale"])
    a1 = z @ q(p["W1"]).T + p["b1"]
    h1 = q(silu(a1))
    a2 = h1 @ q(p["W2"]).T + p["b2"]
    h2 = q(silu(a2))
    e = h2 @ q(p["w3"]) + p["b3"][0]
    return e, (z, a1, h1, a2, h2)


def input_grad(G, p, bf=False, cache=None):
    """g = d e_i / d G(i), shape (N, 32).  With bf the back-propagated vectors are rounded to bfloat16
    before each matrix product, as the MFMA kernel feeds them."""
    q = (lambda a: bf16(a)) if bf else (lambda a: a)
    if cache is None:
        _, cache = mlp(G, p, bf)
    z, a1, h1, a2, h2 = cache
    d2 = q(p["w3"])[None, :] * dsilu(a2)
    d1 = (q(d2) @ q(p["W2"])) * dsilu(a1)
    return (q(d1) @ q(p["W1"])) * p["scale"][None, :]


def energy_grad(x, L, rc, p, bf=False):
    """Surrogate energy E and its gradient +dE/dx (N, 3) (the engine's force convention)."""
    i, j, d, r = pairs(x, L, rc)
    phi, dphi = basis(r, rc)
    G = np.zeros((x.shape[0], NFEAT))
    np.add.at(G, i, phi)
    e, cache = mlp(G, p, bf)
    g = input_grad(G, p, bf, cache)
    # every ordered pair (i, j) appears once with i first: its term (x_i - x_j) / r_ij acts on atom i
    coef = ((g[i] + g[j]) * dphi).sum(1) / r
    grad = np.zeros_like(x)
    np.add.at(grad, i, coef[:, None] * d)
    return e.sum(), grad, e


def lj_atom_energies(x, L, rc, eps=1.0, sigma=1.0):
    i, j, d, r = pairs(x, L, rc)
    sr6 = (sigma / r) ** 6
    u = np.zeros(x.shape[0])
    np.add.at(u, i, 2 * eps * sr6 * (sr6 - 1))
    return u


def loss_grad(p, G, u, bf=False):
    """Per-atom energy MSE and its gradient w.r.t. W1, b1, W2, b2, w3, b3 (reverse mode).  With bf the
    forward pass is the bfloat16 one of mlp() and d2 is rounded before the product with W2."""
    q = (lambda a: bf16(a)) if bf else (lambda a: a)
    e, (z, a1, h1, a2, h2) = mlp(G, p, bf)
    n = G.shape[0]
    res = e - u
    loss = np.mean(res ** 2)
    de = 2 * res / n
    gw3 = h2.T @ de
    gb3 = np.array([de.sum()])
    d2 = de[:, None] * q(p["w3"])[None, :] * dsilu(a2)
    gW2 = d2.T @ h1
    gb2 = d2.sum(0)
    d1 = (q(d2) @ q(p["W2"])) * dsilu(a1)
    gW1 = d1.T @ z
    gb1 = d1.sum(0)
    return loss, {"W1": gW1, "b1": gb1, "W2": gW2, "b2": gb2, "w3": gw3, "b3": gb3}


def adam_step(p, grads, state, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015) of the trainable tensors, in place."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k, gk in grads.items():
        m = state.setdefault("m_" + k, np.zeros_like(gk))
        v = state.setdefault("v_" + k, np.zeros_like(gk))
        m[...] = b1 * m + (1 - b1) * gk
        v[...] = b2 * v + (1 - b2) * gk * gk
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p[k] = p[k] - lr * mh / (np.sqrt(vh) + eps)
    return p


def lj_forces(x, L, rc, eps=1.0, sigma=1.0):
    """+grad U of the truncated LJ potential (the engine's force convention), FP64, minimum image."""
    i, j, d, r = pairs(x, L, rc)
    sr6 = (sigma / r) ** 6
    coef = 48 * eps * sr6 * (0.5 - sr6) / r ** 2
    f = np.zeros_like(x)
    np.add.at(f, i, coef[:, None] * d)
    return f


def dsilu2(a):
    """Second derivative of silu: s (1 - s) (2 + a (1 - 2 s))."""
    s = 1 / (1 + np.exp(-a))
    return s * (1 - s) * (2 + a * (1 - 2 * s))


def force_loss_grad(p, configs, L, rc, w_force, bf=False):
    """Energy + force matching on whole configurations (a list of (N, 3) position arrays of one box):

      loss = mean_i (e_i - u_i)^2 + w_force * mean_{i,d} (dE/dx_id - dU/dx_id)^2

    over all atoms i of all configurations, and its gradient w.r.t. W1, b1, W2, b2, w3, b3.  The force
    term needs the mixed derivative d(dE/dx)/dtheta: with R_i = dE/dx_i - dU/dx_i the loss depends on the
    input gradients g through v_i = dL/dg_i = c sum_j phi'(r_ij) (x_i - x_j)/r_ij . (R_i - R_j),
    c = 2 w_force / (3 n), and S = sum_i g_i . v_i is differentiated by forward mode along u = v * scale
    (t1 = W1 u, ta1 = silu'(a1) t1, t2 = W2 ta1) followed by reverse mode.  Returns (energy MSE, force MSE,
    gradient dict).  bf: the bfloat16 forward of mlp() (weights and layer inputs rounded), and the
    tangent pass on the rounded weights with the operands the training kernel feeds to its bfloat16
    products rounded the same way (u, ta1, zb2; the rounded u and ta1 also enter the weight
    gradients), FP64 otherwise."""
    q = (lambda a: bf16(a)) if bf else (lambda a: a)
    W1, W2, w3 = q(p["W1"]), q(p["W2"]), q(p["w3"])
    rows = []
    for x in configs:
        i, j, d, r = pairs(x, L, rc)
        phi, dphi = basis(r, rc)
        G = np.zeros((x.shape[0], NFEAT))
        np.add.at(G, i, phi)
        rows.append((x, i, j, d, r, dphi, G))
    n = sum(x.shape[0] for x, *_ in rows)
    c = 2.0 * w_force / (3.0 * n)
    gr = {k: np.zeros_like(p[k]) for k in ("W1", "b1", "W2", "b2", "w3", "b3")}
    se = sf = 0.0
    for x, i, j, d, r, dphi, G in rows:
        e, (z, a1, h1, a2, h2) = mlp(G, p, bf)
        g = input_grad(G, p, bf, (z, a1, h1, a2, h2))
        coef = ((g[i] + g[j]) * dphi).sum(1) / r
        gradE = np.zeros_like(x)
        np.add.at(gradE, i, coef[:, None] * d)
        u = lj_atom_energies(x, L, rc)
        R = gradE - lj_forces(x, L, rc)
        res = e - u
        se += np.sum(res ** 2)
        sf += np.sum(R ** 2)
        de = 2 * res / n
        # v_i[k] = c sum_j dphi_k(r_ij) (d_ij / r_ij) . (R_i - R_j)
        proj = (d / r[:, None] * (R[i] - R[j])).sum(1)
        v = np.zeros((x.shape[0], NFEAT))
        np.add.at(v, i, c * dphi * proj[:, None])
        uu = q(v * p["scale"][None, :])
        # backward quantities of e (de = 1): delta2 = w3 silu'(a2), gamma1 = delta2 W2, delta1 = silu'(a1) gamma1
        s1, s2 = dsilu(a1), dsilu(a2)
        delta2 = w3[None, :] * s2
        gamma1 = q(delta2) @ W2
        delta1 = gamma1 * s1
        # tangent along u, then its reverse pass
        t1 = uu @ W1.T
        ta1 = q(s1 * t1)
        t2 = ta1 @ W2.T
        zb2 = w3[None, :] * dsilu2(a2) * t2
        ab1 = q(zb2) @ W2
        zb1 = dsilu2(a1) * t1 * gamma1 + s1 * ab1
        D2 = de[:, None] * delta2 + zb2
        D1 = de[:, None] * delta1 + zb1
        gr["w3"] += h2.T @ de + (s2 * t2).sum(0)
        gr["b3"] += np.array([de.sum()])
        gr["W2"] += D2.T @ h1 + delta2.T @ ta1
        gr["b2"] += D2.sum(0)
        gr["W1"] += D1.T @ z + delta1.T @ uu
        gr["b1"] += D1.sum(0)
    return se / n, sf / (3 * n), gr
