"""GPU FP64 direct Fourier transform (hws_direct_invert_device /
hws_direct_predict_device) against the oracle's FP64 DFT, the restatement of
DirectMSGridder<double> (gridding/directmsgridder.cpp:16-114, 175-196) and of
its adjoint predict."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _obs(nrow, nchan, dl, seed, frac=0.45, wmax=60.0):
    rng = np.random.default_rng(seed)
    wl = C_LIGHT / np.linspace(120e6, 160e6, nchan)
    umax = frac / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-wmax, wmax, nrow)])
    return rng, uvw, wl


def _lambda(uvw, wl):
    return (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


@pytest.mark.parametrize("w,h,shift", [(48, 40, (0.0, 0.0)), (64, 64, (0.013, -0.021)),
                                       (34, 50, (-0.004, 0.007))])
def test_direct_invert_matches_oracle_dft(oracle, w, h, shift):
    from wsclean_amd.direct import direct_invert
    dl, dm = 0.004, 0.0035
    rng, uvw, wl = _obs(300, 3, dl, 11)
    vis = (rng.standard_normal((300, 3, 2)) + 1j * rng.standard_normal((300, 3, 2))
           ).astype(np.complex64)
    got = direct_invert(_dev(uvw), _dev(wl), _dev(vis), w, h, dl, dm, shift[0], shift[1],
                        multiplication_factor=0.25, image=1)
    torch.cuda.synchronize()
    ref = oracle.dft_invert(w, h, dl, dm, _lambda(uvw, wl), vis[:, :, 1].reshape(-1),
                            factor=0.25, shift=shift)
    got = got.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"direct invert {w}x{h} shift={shift}: rel err {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("dense", [False, True])
def test_direct_predict_matches_oracle_dft(oracle, dense):
    from wsclean_amd.direct import direct_predict
    w, h, dl, dm = 40, 36, 0.004, 0.004
    rng, uvw, wl = _obs(257, 2, dl, 12)
    model = np.zeros((h, w), np.float32)
    if dense:
        model[:] = rng.standard_normal((h, w))  # 1440 terms: several LDS chunks
    else:
        idx = rng.choice(w * h, 30, replace=False)
        model.reshape(-1)[idx] = rng.uniform(0.1, 3.0, 30)
    shift = (0.006, -0.003)
    got = direct_predict(_dev(model), _dev(uvw), _dev(wl), dl, dm, *shift)
    torch.cuda.synchronize()
    ref = oracle.dft_predict(w, h, dl, dm, _lambda(uvw, wl), model.astype(np.float64),
                             shift=shift)
    got = got.cpu().numpy().reshape(-1)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print(f"direct predict dense={dense}: rel err {err:.2e}")
    assert err < 1e-6  # complex64 output


def test_direct_pair_is_adjoint():
    from wsclean_amd.direct import direct_invert, direct_predict
    w, h, dl = 64, 48, 0.003
    rng, uvw, wl = _obs(500, 4, dl, 13)
    vis = (rng.standard_normal((500, 4)) + 1j * rng.standard_normal((500, 4))).astype(
        np.complex64)
    model = rng.standard_normal((h, w)).astype(np.float32)
    img = direct_invert(_dev(uvw), _dev(wl), _dev(vis), w, h, dl, dl, 0.01, 0.02)
    pred = direct_predict(_dev(model), _dev(uvw), _dev(wl), dl, dl, 0.01, 0.02)
    torch.cuda.synchronize()
    lhs = float(np.sum(img.cpu().numpy() * model.astype(np.float64)))
    rhs = float(np.real(np.sum(vis.astype(np.complex128) *
                               np.conj(pred.cpu().numpy().astype(np.complex128)))))
    assert abs(lhs - rhs) <= 1e-6 * (abs(lhs) + abs(rhs))


def test_direct_empty_inputs():
    from wsclean_amd.direct import direct_invert, direct_predict
    uvw = torch.zeros((0, 3), dtype=torch.float64, device="cuda:0")
    wl = torch.tensor([2.0], dtype=torch.float64, device="cuda:0")
    vis = torch.zeros((0, 1), dtype=torch.complex64, device="cuda:0")
    img = direct_invert(uvw, wl, vis, 16, 16, 0.01, 0.01)
    out = direct_predict(torch.ones((16, 16), dtype=torch.float32, device="cuda:0"), uvw, wl,
                         0.01, 0.01)
    torch.cuda.synchronize()
    assert torch.count_nonzero(img).item() == 0 and out.numel() == 0
    _, uvw2, wl2 = _obs(10, 1, 0.01, 3)
    zero = direct_predict(torch.zeros((16, 16), dtype=torch.float32, device="cuda:0"),
                          _dev(uvw2), _dev(wl2), 0.01, 0.01)
    torch.cuda.synchronize()
    assert torch.count_nonzero(zero).item() == 0

