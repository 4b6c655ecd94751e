"""hws_direct_predict_device_noalloc, the direct predict whose scratch is a
fixed per-device array instead of allocated buffers, called through the C
ABI: its output equals hws_direct_predict_device's bit for bit (same kernels,
same order of sums) on two calls in a row, the second with a smaller image
and fewer samples than the first, so a stale pixel list or count from the
first call would show; and it rejects images beyond its fixed capacity."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _ptr(t, offset=0):
    return C.c_void_p(t.data_ptr() + offset)


def _case(seed, width, height, nrow, nchan, n_images, density):
    rng = np.random.default_rng(seed)
    dl = 0.002
    wl = C_LIGHT / np.linspace(120e6, 150e6, nchan)
    umax = 0.4 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-40.0, 40.0, nrow)])
    model = rng.standard_normal((height, width)).astype(np.float32)
    model[rng.uniform(size=model.shape) > density] = 0.0
    return uvw, wl, model, dl


def _predict(L, fn, dev, case, shift, n_images, image):
    uvw, wl, model, dl = case
    nrow, nchan = uvw.shape[0], wl.shape[0]
    d_uvw = torch.from_numpy(uvw).to(dev)
    d_wl = torch.from_numpy(wl).to(dev)
    d_model = torch.from_numpy(model).to(dev)
    out = torch.full((nrow, nchan, n_images), complex(7.0, -7.0), dtype=torch.complex64,
                     device=dev)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    h, w = model.shape
    st = fn(0, stream, w, h, dl, dl, shift[0], shift[1], _ptr(d_model), nrow, nchan,
            _ptr(d_uvw), _ptr(d_wl), _ptr(out, image * 8), n_images)
    assert st == 0, L.hws_last_error(None)
    torch.cuda.synchronize()
    return out.cpu()


def test_noalloc_equals_allocating_direct_predict_on_consecutive_calls():
    from wsclean_amd import _lib
    L = _lib.load()
    dev = torch.device("cuda:0")
    cases = [  # the second call is smaller in image and samples
        (_case(1, 384, 256, 300, 4, 2, 0.3), (0.01, -0.02), 2, 1),
        (_case(2, 64, 96, 40, 2, 1, 0.5), (0.0, 0.0), 1, 0),
    ]
    for k, (case, shift, n_images, image) in enumerate(cases):
        print(f"call {k}: image {case[2].shape}, {case[0].shape[0]} rows x {case[1].shape[0]} "
              f"ch, {int((case[2] != 0).sum())} non-zero pixels", flush=True)
        ref = _predict(L, L.hws_direct_predict_device, dev, case, shift, n_images, image)
        got = _predict(L, L.hws_direct_predict_device_noalloc, dev, case, shift, n_images,
                       image)
        assert torch.equal(got.view(torch.float32).view(torch.int32),
                           ref.view(torch.float32).view(torch.int32)), k
        # the other image of the output is untouched
        if n_images > 1:
            assert torch.all(got[..., 1 - image] == complex(7.0, -7.0))
        # the sums are not trivially zero
        assert float(torch.abs(got[..., image]).max()) > 1.0


def test_noalloc_rejects_images_beyond_its_capacity():
    from wsclean_amd import _lib
    L = _lib.load()
    fake = C.c_void_p(1 << 20)  # never dereferenced: rejected by the size check
    st = L.hws_direct_predict_device_noalloc(0, None, 2048, 1024, 0.01, 0.01, 0.0, 0.0, fake,
                                             10, 1, fake, fake, fake, 1)
    assert st == 1
    assert b"at most 1048576" in L.hws_last_error(None)
