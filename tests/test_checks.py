"""Argument checks on the tensors whose raw pointers go to the C ABI
(wsclean_amd/_checks.py): every mismatch in dtype, size, layout or device
raises ValueError before the library is called. Runs without a GPU: the
gridder object is built without a device context and its library handle is
a stub that fails the test if any entry point is reached."""
import pytest

torch = pytest.importorskip("torch")

from wsclean_amd import _checks  # noqa: E402
from wsclean_amd.gridder import DeviceGridder  # noqa: E402


class _NoCalls:
    def __getattr__(self, name):
        raise AssertionError(f"library entry point {name} reached")


def _gridder(nrow=10, nchan=3, n_images=2, complex_images=(1,), H=8, W=16):
    g = DeviceGridder.__new__(DeviceGridder)
    g.L, g.h = _NoCalls(), None
    g.nrow, g.nchan, g.n_images = nrow, nchan, n_images
    g.height, g.width, g.complex_images, g.device = H, W, complex_images, 0
    g.planned = nrow > 0
    return g


def _good(g):
    vis = torch.zeros((g.nrow, g.nchan, g.n_images), dtype=torch.complex64)
    img = torch.zeros((g.n_images, g.height, g.width), dtype=torch.float32)
    return vis, img, img.clone()


def test_plan_rejects_bad_uvw_and_wavelengths():
    g = _gridder()
    wl = torch.ones(3, dtype=torch.float64)
    for bad in (torch.zeros((10, 3), dtype=torch.float32), torch.zeros((10, 2), dtype=torch.float64),
                torch.zeros(30, dtype=torch.float64),
                torch.zeros((3, 10), dtype=torch.float64).t()):
        with pytest.raises(ValueError):
            g.plan(bad, wl)
    uvw = torch.zeros((10, 3), dtype=torch.float64)
    for bad in (torch.ones(3, dtype=torch.float32), torch.ones((3, 1), dtype=torch.float64),
                torch.ones(6, dtype=torch.float64)[::2]):
        with pytest.raises(ValueError):
            g.plan(uvw, bad)
    # right dtypes and shapes, but host tensors: rejected by the device check
    with pytest.raises(ValueError, match="expected cuda:0"):
        g.plan(uvw, wl)
    # empty sample sets are valid (the library plans nothing)
    with pytest.raises(ValueError, match="expected cuda:0"):
        g.plan(torch.zeros((0, 3), dtype=torch.float64), wl)


@pytest.mark.parametrize("which", ["vis_dtype", "vis_size", "vis_strided", "real_dtype",
                                   "real_shape", "imag_missing", "imag_shape", "real_strided"])
def test_invert_and_predict_reject_mismatches(which):
    g = _gridder()
    vis, re, im = _good(g)
    if which == "vis_dtype":
        vis = torch.zeros(vis.shape, dtype=torch.complex128)
    elif which == "vis_size":
        vis = torch.zeros((g.nrow, g.nchan, g.n_images - 1), dtype=torch.complex64)
    elif which == "vis_strided":
        vis = torch.zeros((g.nrow, g.nchan, 2 * g.n_images), dtype=torch.complex64)[..., ::2]
    elif which == "real_dtype":
        re = re.double()
    elif which == "real_shape":
        re = torch.zeros((g.n_images, g.width, g.height), dtype=torch.float32)
    elif which == "imag_missing":
        im = None
    elif which == "imag_shape":
        im = torch.zeros((1, g.height, g.width), dtype=torch.float32)
    elif which == "real_strided":
        re = torch.zeros((g.n_images, g.height, 2 * g.width), dtype=torch.float32)[..., ::2]
    with pytest.raises(ValueError) as e:
        g.invert(vis, 1.0, re, im)
    assert "expected cuda" not in str(e.value), e.value
    with pytest.raises(ValueError) as e:
        g.predict(re, im, vis)
    assert "expected cuda" not in str(e.value), e.value
    with pytest.raises(ValueError):
        g.capture_invert(vis, 1.0, re, im)
    with pytest.raises(ValueError):
        g.capture_predict(re, im, vis)


def test_well_formed_host_tensors_fail_only_on_the_device_check():
    g = _gridder()
    vis, re, im = _good(g)
    with pytest.raises(ValueError, match="expected cuda:0"):
        g.invert(vis, 1.0, re, im)
    with pytest.raises(ValueError, match="expected cuda:0"):
        g.predict(re, im, vis)
    # imag may be omitted when no image is complex
    g2 = _gridder(complex_images=())
    vis, re, _ = _good(g2)
    with pytest.raises(ValueError, match="expected cuda:0"):
        g2.invert(vis, 1.0, re, None)


def test_invert_before_plan_is_rejected():
    g = _gridder(nrow=0, nchan=0)
    vis, re, im = _good(_gridder())
    with pytest.raises(ValueError, match="plan"):
        g.invert(vis, 1.0, re, im)


def test_direct_transform_checks():
    from wsclean_amd import direct
    uvw = torch.zeros((4, 3), dtype=torch.float64)
    wl = torch.ones(2, dtype=torch.float64)
    with pytest.raises(ValueError):
        direct.direct_invert(uvw.float(), wl, torch.zeros((4, 2), dtype=torch.complex64),
                             8, 8, 0.01, 0.01)
    with pytest.raises(ValueError):
        direct.direct_invert(uvw, wl, torch.zeros((4, 3), dtype=torch.complex64), 8, 8, 0.01,
                             0.01)
    with pytest.raises(ValueError):
        direct.direct_predict(torch.zeros((8, 8), dtype=torch.float64), uvw, wl, 0.01, 0.01)
    with pytest.raises(ValueError, match="expected cuda"):
        direct.direct_predict(torch.zeros((8, 8), dtype=torch.float32), uvw, wl, 0.01, 0.01)


def test_tensor_check_order():
    t = torch.zeros((2, 3), dtype=torch.float32)
    assert _checks.tensor(None, "x", torch.float32, optional=True) is None
    with pytest.raises(ValueError, match="required"):
        _checks.tensor(None, "x", torch.float32)
    with pytest.raises(ValueError, match="dtype"):
        _checks.tensor(t, "x", torch.float64, shape=(9,))
    with pytest.raises(ValueError, match="shape"):
        _checks.tensor(t, "x", torch.float32, shape=(3, 2))
    with pytest.raises(ValueError, match="elements"):
        _checks.tensor(t, "x", torch.float32, numel=5)
    with pytest.raises(ValueError, match="contiguous"):
        _checks.tensor(t.t(), "x", torch.float32, numel=6)
