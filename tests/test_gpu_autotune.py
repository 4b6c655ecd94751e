"""The autotuner times every (tile size, work chunk) candidate on the given
samples and picks the fastest; its w-layer count is the reference's
suggestion (gridding/wsmsgridder.cpp:77-128), never less than asked."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


def test_autotune_picks_fastest_candidate():
    from wsclean_amd import synth
    from wsclean_amd.autotune import autotune, suggested_w_layers
    obs = synth.config2(scale=0.01, nchan=16)
    dev = torch.device("cuda:0")
    uvw = torch.from_numpy(obs.uvw_m).to(dev)
    wl = torch.from_numpy(obs.wavelengths).to(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4)
    vis = torch.randn((obs.nrow, obs.nchan, 2), dtype=torch.complex64, device=dev,
                      generator=gen)
    pix = 4.0 * np.pi / (180 * 3600)
    r = autotune(1024, 1024, pix, pix, uvw, wl, vis, n_images=2, complex_images=(1,),
                 chunks=(2048, 8192), repeats=1, min_layers=4)
    assert len(r["trials"]) == 4
    assert r["ms"] == min(t["ms"] for t in r["trials"])
    assert (r["tile_size"], r["work_chunk"]) in {(t["tile_size"], t["work_chunk"])
                                                 for t in r["trials"]}
    lo, hi = r["w_range"]
    assert r["n_w_layers"] == max(4, suggested_w_layers(1024, 1024, pix, pix, -hi, hi))
    assert r["passes"] == 1


def test_autotune_w_layers_meet_an_image_error_bound():
    """With max_image_error the tuner searches the w-layer count upwards from
    the suggestion: the count it returns meets the bound against the FP64
    direct transform of the sampled visibilities, the counts before it do
    not, and the error falls as the layers get closer."""
    from wsclean_amd import accuracy
    from wsclean_amd.autotune import autotune
    wk = accuracy.precision_workload(n_samples=20_000, seed=5)
    dev = torch.device("cuda:0")
    # one sample per row, uvw in wavelengths: the layout DeviceGridder takes
    uvw = torch.from_numpy(np.ascontiguousarray(wk["uvw_lambda"])).to(dev)
    wl = torch.ones(1, dtype=torch.float64, device=dev)
    vis = torch.from_numpy(wk["vis"]).to(dev).reshape(-1, 1, 1)
    W, dl = 512, wk["dl"]
    first = autotune(W, W, dl, dl, uvw, wl, vis, chunks=(8192,), tiles=(16,), repeats=1)
    # the error curve from the suggestion to eight times it (a bound of 0
    # is never met, so every step is tried); the bound lies halfway
    # between its first and its smallest value
    _, curve = accuracy.w_layers_for_accuracy(W, W, dl, dl, uvw, wl, vis, 0.0,
                                              n_w_start=first["n_w_layers"],
                                              n_w_max=8 * first["n_w_layers"])
    errs = [e for _, e in curve]
    bound = 0.5 * (errs[0] + min(errs))
    r = autotune(W, W, dl, dl, uvw, wl, vis, chunks=(8192,), tiles=(16,), repeats=1,
                 max_image_error=bound, accuracy_samples=20_000)
    tried = r["accuracy"]
    print(f"suggested {first['n_w_layers']} layers; error curve {curve}; bound {bound:.3g}: "
          f"{tried}")
    assert r["n_w_layers"] == tried[-1][0] and tried[-1][1] <= bound
    assert all(e > bound for _, e in tried[:-1])
    assert tried[0][0] == first["n_w_layers"] and len(tried) >= 2
    assert tried[-1][1] < tried[0][1]
