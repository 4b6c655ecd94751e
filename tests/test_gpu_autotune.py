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
