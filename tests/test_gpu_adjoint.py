"""Path-replay adjoint (shb_bake_adjoint) and the torch fitting harness.

The bake's light selection uses radiance without the fitted scales, so each
sample's contribution is a monomial in the light / albedo scales: the
replay gradient is the exact derivative of the estimator for the same
random numbers. Checks: central finite differences under common random
numbers (h = 1e-2), the exact light gradient from per-light forward bakes
(one-hot light scales), non-unit scales, and the autograd wrapper."""
import numpy as np
import pytest

import shb
from shb import scenes

pytestmark = pytest.mark.gpu

SPP, B, SEED = 16, 3, 99


@pytest.fixture(scope="module")
def setup():
    s = scenes.scaled_synthetic_scene(6000, 96, 21)
    b = shb.Baker(0)
    b.upload_scene(s)
    b.rasterize(96, 96, 1, want=False)
    nl, nm = len(s.lights), len(s.materials)
    rng = np.random.default_rng(3)
    sh0 = b.bake(SPP, B, seed=SEED)[0]
    target = sh0 * rng.uniform(0.5, 1.5, sh0.shape).astype(np.float32)
    yield b, nl, nm, target
    b.close()


def loss_and_grad(sh, target):
    d = sh.astype(np.float64) - target
    return 0.5 * float((d * d).sum()), d.astype(np.float32)


def bake_loss(b, ls, al, target):
    b.set_params(ls, al)
    sh = b.bake(SPP, B, seed=SEED)[0]
    return loss_and_grad(sh, target)[0]


def check_fd(b, nl, nm, target, ls, al, picks_l, picks_m, h=1e-2):
    # every contribution is a polynomial of degree <= bounces + 2 in the
    # scales, so central differences are exact to O(h^2); h = 1e-2 keeps the
    # loss difference far above the float32 rounding of the SH outputs
    b.set_params(ls, al)
    sh = b.bake(SPP, B, seed=SEED)[0]
    _, g = loss_and_grad(sh, target)
    dl, dm, _ = b.bake_adjoint(SPP, B, g, seed=SEED)
    for (kind, idx, ch) in [("l", i, c) for i, c in picks_l] + [("m", i, c) for i, c in picks_m]:
        lp, lm_, ap, am = ls.copy(), ls.copy(), al.copy(), al.copy()
        if kind == "l":
            base = ls[idx, ch]
            lp[idx, ch] = base * (1 + h)
            lm_[idx, ch] = base * (1 - h)
            fd = (bake_loss(b, lp, al, target) - bake_loss(b, lm_, al, target)) / (2 * h * base)
            adj = dl[idx, ch]
        else:
            base = al[idx, ch]
            ap[idx, ch] = base * (1 + h)
            am[idx, ch] = base * (1 - h)
            fd = (bake_loss(b, ls, ap, target) - bake_loss(b, ls, am, target)) / (2 * h * base)
            adj = dm[idx, ch]
        scale = max(abs(fd), 1e-3 * (np.abs(dl).max() + np.abs(dm).max()))
        assert abs(adj - fd) <= 1e-2 * scale, (kind, idx, ch, adj, fd)
    b.set_params(np.ones_like(ls), np.ones_like(al))
    return dl, dm


def strongest(g, k):
    flat = np.argsort(-np.abs(g).ravel())[:k]
    return [(int(i // 3), int(i % 3)) for i in flat]


def test_adjoint_matches_central_fd(setup):
    b, nl, nm, target = setup
    ls, al = np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32)
    b.set_params(ls, al)
    _, g = loss_and_grad(b.bake(SPP, B, seed=SEED)[0], target)
    dl, dm, _ = b.bake_adjoint(SPP, B, g, seed=SEED)
    assert np.isfinite(dl).all() and np.isfinite(dm).all()
    assert np.abs(dl).max() > 0 and np.abs(dm).max() > 0
    check_fd(b, nl, nm, target, ls, al, strongest(dl, 4), strongest(dm, 4))


def test_adjoint_fd_with_non_unit_scales(setup):
    b, nl, nm, target = setup
    rng = np.random.default_rng(8)
    ls = rng.uniform(0.5, 2.0, (nl, 3)).astype(np.float32)
    al = rng.uniform(0.5, 1.2, (nm, 3)).astype(np.float32)
    b.set_params(ls, al)
    _, g = loss_and_grad(b.bake(SPP, B, seed=SEED)[0], target)
    dl, dm, _ = b.bake_adjoint(SPP, B, g, seed=SEED)
    b.set_params(ls, al)
    check_fd(b, nl, nm, target, ls, al, strongest(dl, 3), strongest(dm, 3))


def test_light_gradient_exact_from_per_light_bakes(setup):
    # linear loss L = sum(w * SH): dL/ds_{l,c} = sum over texels and
    # coefficients of w * (light l's own SH, channel c), and light l's own
    # SH is a forward bake with every other light's scale 0 (same samples)
    b, nl, nm, target = setup
    w = np.random.default_rng(5).normal(size=target.shape).astype(np.float32)
    b.set_params(np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32))
    dl, _, _ = b.bake_adjoint(SPP, B, w, seed=SEED)
    for l in strongest(dl, 6):
        li = l[0]
        onehot = np.zeros((nl, 3), np.float32)
        onehot[li] = 1.0
        b.set_params(onehot, None)
        sh_l = b.bake(SPP, B, seed=SEED)[0].astype(np.float64).reshape(-1, 9, 3)
        exact = (w.astype(np.float64).reshape(-1, 9, 3) * sh_l).sum(axis=(0, 1))
        np.testing.assert_allclose(dl[li], exact, rtol=2e-3,
                                   atol=1e-4 * np.abs(exact).max() + 1e-7)
    b.set_params(np.ones((nl, 3), np.float32), None)


def test_adjoint_deterministic_inputs_and_errors(setup):
    b, nl, nm, target = setup
    b.set_params(np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32))
    zero = np.zeros_like(target)
    dl, dm, _ = b.bake_adjoint(SPP, B, zero, seed=SEED)
    assert not dl.any() and not dm.any()
    ls = np.ones((nl, 3), np.float32)
    ls[0, 0] = 0.0
    b.set_params(ls, None)
    with pytest.raises(shb.ShbError) as e:
        b.bake_adjoint(SPP, B, zero, seed=SEED)
    assert e.value.code == -1
    b.set_params(np.ones((nl, 3), np.float32), None)
    with pytest.raises(shb.ShbError):
        b.set_params(-ls, None)


def test_torch_fit_reduces_loss(setup):
    b, nl, nm, _ = setup
    from shb import fit
    true_l, true_a = fit.lognormal_factors(nl, nm, 0.3, seed=1)
    b.set_params(true_l, true_a)
    target = b.bake(SPP, B, seed=SEED)[0]
    b.set_params(np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32))
    ls, al, losses = fit.fit(b, target, SPP, B, iters=12, lr=0.05, seed=SEED)
    assert losses[-1] < 0.5 * losses[0], losses
    b.set_params(np.ones((nl, 3), np.float32), np.ones((nm, 3), np.float32))
