"""Degridding as a torch.autograd.Function (wsclean_amd.autograd).

The differentiation workload: an FP32 2048^2 model of 500 point sources
(log-uniform flux) convolved with a 2-pixel-FWHM Gaussian; the 2016-baseline
3 km-disc array over 600 x 10 s and 16 channels (19,353,600 samples);
chi-squared against seeded noisy observed visibilities. Checks: the dot-product
(adjoint) test, central finite differences on 64 seeded pixels, the backward
against the oracle's invert of the same residual, and 20 gradient-descent
steps that lower chi-squared every step."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

ARCSEC = np.pi / (180.0 * 3600.0)


def _model(width, n_src, rng, fwhm_px=2.0):
    img = np.zeros((width, width), np.float64)
    xy = rng.integers(width // 8, width - width // 8, size=(n_src, 2))
    flux = np.exp(rng.uniform(np.log(0.01), np.log(10.0), n_src))
    np.add.at(img, (xy[:, 1], xy[:, 0]), flux)
    sigma = fwhm_px / (2.0 * np.sqrt(2.0 * np.log(2.0)))
    r = np.arange(-6, 7)
    k = np.exp(-0.5 * (r / sigma) ** 2)
    k /= k.sum()
    img = np.apply_along_axis(lambda a: np.convolve(a, k, mode="same"), 0, img)
    img = np.apply_along_axis(lambda a: np.convolve(a, k, mode="same"), 1, img)
    return img.astype(np.float32)


@pytest.fixture(scope="module")
def problem():
    from wsclean_amd import synth
    from wsclean_amd.autograd import WStackingOperator, chi_squared
    rng = np.random.default_rng(505)
    obs = synth.config1(scale=6.0)  # the 3 km-disc array over 600 x 10 s, 16 channels
    W, dl = 2048, 20.0 * ARCSEC
    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(obs.uvw_m).to(dev)
    wl = torch.from_numpy(obs.wavelengths).to(dev)
    op = WStackingOperator(W, W, dl, dl, uvw, wl, 32)
    x_true = torch.from_numpy(_model(W, 500, rng)).to(dev)
    y = op.predict(x_true)
    gen = torch.Generator(device=dev)
    gen.manual_seed(7)
    noise = 0.1 * torch.randn(y.shape, dtype=torch.complex64, device=dev, generator=gen)
    y_obs = (y + noise).contiguous()
    print(f"samples {obs.nrow * obs.nchan}, passes {op.n_passes}, "
          f"chi2(x_true) {float(chi_squared(y, y_obs)):.4e}")
    yield dict(op=op, x_true=x_true, y_obs=y_obs, W=W, dl=dl, uvw=obs.uvw_m,
               wl=obs.wavelengths, rng=rng)
    op.close()


def test_workload_size(problem):
    assert problem["op"].nrow * problem["op"].nchan == 2016 * 600 * 16


def test_dot_product(problem):
    from wsclean_amd.autograd import dot_product_test
    lhs, rhs = dot_product_test(problem["op"], seed=3)
    print(f"<Ax,y> = {lhs:.9e}, <x,A^H y> = {rhs:.9e}")
    assert abs(lhs - rhs) <= 2e-6 * max(abs(lhs), abs(rhs))


def test_finite_differences(problem):
    from wsclean_amd.autograd import chi_squared, degrid
    op, y_obs = problem["op"], problem["y_obs"]
    x = (0.5 * problem["x_true"]).clone().requires_grad_(True)
    loss = chi_squared(degrid(x, op), y_obs)
    loss.backward()
    grad = x.grad.detach().cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(64)
    W = problem["W"]
    pix = rng.integers(0, W, size=(64, 2))
    # chi-squared is quadratic in x, so central differences carry no
    # truncation error; a large step keeps FP32 rounding of the forward small
    eps = 2.0
    worst = 0.0
    scale = np.abs(grad).max()
    with torch.no_grad():
        for px, py in pix:
            xp = x.detach().clone()
            xp[py, px] += eps
            lp = float(chi_squared(op.predict(xp), y_obs))
            xp[py, px] -= 2 * eps
            lm = float(chi_squared(op.predict(xp), y_obs))
            fd = (lp - lm) / (2 * eps)
            worst = max(worst, abs(fd - grad[py, px]) / scale)
    print(f"finite differences: worst |fd - grad| / max|grad| = {worst:.3e}")
    assert worst < 1e-3


def test_backward_matches_oracle_invert(problem, oracle):
    """The gradient equals 2 Re(A^H r): the reference invert (unnormalised)
    of the residual, here from the oracle on a 1/16 subset of the rows."""
    from wsclean_amd.autograd import WStackingOperator, chi_squared, degrid
    W, dl = 512, problem["dl"] * 4
    rows = np.arange(0, problem["uvw"].shape[0], 16)
    uvw_m, wl = problem["uvw"][rows], problem["wl"]
    dev = torch.device("cuda:0")
    uvw_l = (uvw_m[:, None, :] / wl[None, :, None]).reshape(-1, 3)
    aw = np.abs(uvw_l[:, 2])
    op = WStackingOperator(W, W, dl, dl, torch.from_numpy(uvw_m).to(dev),
                           torch.from_numpy(wl).to(dev), 16,
                           w_limits=(float(aw.min()), float(aw.max())))
    rng = np.random.default_rng(9)
    y_obs = torch.from_numpy((rng.standard_normal((len(rows), len(wl))) + 1j *
                              rng.standard_normal((len(rows), len(wl)))).astype(np.complex64)
                             ).to(dev)
    x = torch.from_numpy(rng.standard_normal((W, W)).astype(np.float32)).to(dev)
    x.requires_grad_(True)
    pred = degrid(x, op)
    chi_squared(pred, y_obs).backward()
    residual = (pred.detach() - y_obs).cpu().numpy().reshape(-1)
    ref = oracle.RefGridder(W, W, dl, dl, 4, 7, 63, "double")
    ref.prepare_w_layers(16, 1e12, float(aw.min()), float(aw.max()))
    ref.start_inversion_pass(0)
    ref.add_samples(uvw_l, (2.0 * residual).astype(np.complex64))
    ref.finish_inversion_pass()
    ref.finalize_image(1.0)
    got = x.grad.detach().cpu().numpy()
    err = np.linalg.norm(got - ref.image()) / np.linalg.norm(ref.image())
    op.close()
    print(f"backward vs oracle invert: rel err {err:.2e}")
    assert err < 3e-5


def test_gradient_descent_lowers_chi_squared(problem):
    from wsclean_amd.autograd import chi_squared, degrid
    op, y_obs = problem["op"], problem["y_obs"]
    # chi-squared has the Hessian 2 A^H A; its largest eigenvalue (the peak
    # of the gridded sampling density, far above the number of samples for a
    # dense uv centre) comes from a few power iterations, and a step of 0.8/L
    # with L = 2 lambda_max decreases chi-squared monotonically
    gen = torch.Generator(device="cuda:0")
    gen.manual_seed(11)
    v = torch.randn((problem["W"], problem["W"]), dtype=torch.float32, device="cuda:0",
                    generator=gen)
    lam = 0.0
    for _ in range(12):
        v = v / torch.linalg.vector_norm(v)
        av = op.adjoint(op.predict(v))
        lam = float(torch.sum(v.double() * av.double()))
        v = av
    step = 0.8 / (2.0 * lam)
    x = torch.zeros((problem["W"], problem["W"]), dtype=torch.float32, device="cuda:0",
                    requires_grad=True)
    losses = []
    for _ in range(20):
        loss = chi_squared(degrid(x, op), y_obs)
        losses.append(float(loss.detach()))
        (g,) = torch.autograd.grad(loss, x)
        with torch.no_grad():
            x -= step * g
    print(f"lambda_max(A^H A) ~ {lam:.4e} (n = {op.nrow * op.nchan})")
    print("chi2:", " ".join(f"{v:.4e}" for v in losses))
    assert all(b < a for a, b in zip(losses, losses[1:]))
    assert losses[-1] < 0.9 * losses[0]
