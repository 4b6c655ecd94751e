"""Benchmark: visibilities/s of grid + degrid (invert + predict) at 4096^2 on
one MI355X, with roofline numbers for the dominant kernel and a CPU baseline.

Workload (one "step" = one major-cycle pass of the hot path): the 4096^2,
64-w-layer observation of 62 stations on a 40 km log spiral (1891
baselines), 1800 x 2 s at dec +52 deg, 64 channels 115-165 MHz, four
correlations XX, XY, YX, YY (XX/YY real images, XY/YX complex images):
217.8M (row, channel) samples, 871.4M visibilities. A step bins the samples
by (w-layer, uv tile), inverts all four images (grid, 64 layer FFTs, w-
correction, kernel correction) and predicts all four (image -> layers, FFTs,
degrid), with every input and output resident in HBM.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("visibilities/s grid+degrid at 4096² (1 GPU) + RMS err vs FP64 DFT; "
          "CPU baseline")
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--precision", default="mixed", choices=["mixed", "reference", "plain"])
    p.add_argument("--scale", type=float, default=1.0,
                   help="fraction of the 1800 timesteps (1.0 = full workload)")
    p.add_argument("--tile", type=int, default=0)
    p.add_argument("--chunk", type=int, default=0, help="records per grid work item (0: auto)")
    p.add_argument("--autotune", action="store_true",
                   help="choose tile size and work chunk by timing candidates first")
    p.add_argument("--no-plain", action="store_true", help="skip the plain-form comparison")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--no-accuracy", action="store_true",
                   help="skip the error report against the FP64 direct transform")
    p.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                   help="first copy BYTES device-to-device (the known-size copy against which "
                        "hardware byte counters of a profile are scaled)")
    p.add_argument("--no-facets", action="store_true", help="skip the facet-batch leg")
    p.add_argument("--no-plugin", action="store_true",
                   help="skip the plugin leg (host rows staged over the bus, host outputs)")
    p.add_argument("--facets-only", action="store_true",
                   help="run only the facet-batch leg and print its JSON object")
    p.add_argument("--facet-scale", type=float, default=0.25,
                   help="fraction of the 1800 timesteps in the facet leg")
    p.add_argument("--padded-only", metavar="SIZE,LAYERS,OS",
                   help="run only this workload gridded at a padded size (e.g. 6144,128,1023, the "
                        "setting that meets 1e-5) and print its JSON object")
    p.add_argument("--verbose", action="store_true")
    return p.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Workload:
    W = H = 4096
    NW = 64
    PIXEL = 4.0 * np.pi / (180.0 * 3600.0)
    IMAGES = ("XX", "XY", "YX", "YY")
    COMPLEX = (1, 2)

    def __init__(self, scale, device):
        import torch
        from wsclean_amd import synth
        obs = synth.config2(scale=scale)
        self.obs = obs
        self.nrow, self.nchan = obs.nrow, obs.nchan
        self.nvis = self.nrow * self.nchan * 4
        self.uvw = torch.from_numpy(obs.uvw_m).to(device)
        self.wl = torch.from_numpy(obs.wavelengths).to(device)
        wl = obs.wavelengths
        aw = np.abs(obs.uvw_m[:, 2])
        self.min_w = float(aw.min() / wl.max())
        self.max_w = float(aw.max() / wl.min())
        g = torch.Generator(device=device)
        g.manual_seed(2024)
        # seeded complex Gaussian visibilities (kernel cost does not depend on values)
        self.vis = torch.randn((self.nrow, self.nchan, 4), dtype=torch.complex64, device=device,
                               generator=g)
        # seeded visibility weights, uniform in [0.5, 1.5) per visibility
        self.weights = torch.rand((self.nrow, self.nchan, 4), dtype=torch.float32, device=device,
                                  generator=g) + 0.5
        self.vis_out = torch.empty_like(self.vis)
        self.real = torch.zeros((4, self.H, self.W), dtype=torch.float32, device=device)
        self.imag = torch.zeros_like(self.real)


def image_weights(wl_, device):
    """Briggs robust 0 imaging weights of the workload's samples, gridded
    from the XX visibility weights (ImageWeights::Grid over one correlation)
    and finished."""
    from wsclean_amd.weights import DeviceImageWeights
    iw = DeviceImageWeights("briggs", wl_.W, wl_.H, wl_.PIXEL, wl_.PIXEL, robust=0.0,
                            device=device.index or 0)
    iw.grid(wl_.uvw, wl_.wl, wl_.weights_xx)
    iw.finish()
    return iw


def run_form(wl_, precision, steps, warmup, tile, device, chunk=0, cycle=True, weighted=True,
             n_w=None, oversampling=63, size=None):
    """The workload's steps in one precision; `size`: a padded gridded size
    (same pixels, the images of that size) instead of 4096."""
    import torch
    import wsclean_amd
    W = H = size or wl_.W
    real, imag = wl_.real, wl_.imag
    if W != wl_.W:
        real = torch.zeros((4, H, W), dtype=torch.float32, device=device)
        imag = torch.zeros_like(real)
    g = wsclean_amd.DeviceGridder(W, H, wl_.PIXEL, wl_.PIXEL, 7, oversampling,
                                  precision=precision, n_images=4, complex_images=wl_.COMPLEX,
                                  tile_size=tile, work_chunk=chunk)
    npass = g.prepare_w_layers(n_w or wl_.NW, wl_.min_w, wl_.max_w)
    phases = {"weights": 0.0, "bin": 0.0, "grid": 0.0, "fft_inv": 0.0, "project": 0.0,
              "img2layer": 0.0, "fft_pred": 0.0, "degrid": 0.0}
    stats_inv = stats_pred = None
    if weighted and not hasattr(wl_, "weights_xx"):
        wl_.weights_xx = wl_.weights[..., :1].contiguous()
    iw_keep = []

    def step(record):
        nonlocal stats_inv, stats_pred
        if weighted:
            # a step builds the imaging weights, plans (the imaging weight of
            # every sample and the per-image weight sums), inverts with each
            # image normalised by its weight sum, and predicts
            t = time.perf_counter()
            iw = image_weights(wl_, device)
            g.set_weighting(iw, wl_.weights, "normal")
            iw_keep[:] = [iw]  # the previous step's weights are released after re-attaching
            if record:
                torch.cuda.synchronize()
                phases["weights"] += (time.perf_counter() - t) * 1e3
        g.plan(wl_.uvw, wl_.wl)
        g.invert(wl_.vis, 1.0, real, imag)
        if record:
            s = g.stats()
            stats_inv = s
            phases["bin"] += s["ms_bin"]
            phases["grid"] += s["ms_grid"]
            phases["fft_inv"] += s["ms_fft"]
            phases["project"] += s["ms_project"]
        g.predict(real, imag, wl_.vis_out)
        if record:
            s = g.stats()
            stats_pred = s
            phases["img2layer"] += s["ms_project"]
            phases["fft_pred"] += s["ms_fft"]
            phases["degrid"] += s["ms_grid"]

    for _ in range(warmup):
        step(False)
    torch.cuda.synchronize()
    g.synchronize()
    # outputs of the first step, bit patterns included (NaN where a sample is
    # not predicted): every later step must reproduce them, exactly up to the
    # order of the summations done with atomics (the imaging-weight grid's
    # FP64 sums, the binning scatter's placement of a bin's records), which
    # changes results at rounding level (measured: 6e-7 of the image peak)
    ref = None
    if warmup > 0 and precision != "plain":
        ref = (real.clone(), imag.clone(), wl_.vis_out.view(torch.int32).clone())
    # timed region: K steps, host clock, device synchronised on both sides
    t0 = time.perf_counter()
    for _ in range(steps):
        step(False)
    g.synchronize()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    check = {"images_finite": bool(torch.isfinite(real).all() and
                                   torch.isfinite(imag[list(wl_.COMPLEX)]).all())}
    if ref is not None:
        exact = bool(torch.equal(real, ref[0]) and torch.equal(imag, ref[1]) and
                     torch.equal(wl_.vis_out.view(torch.int32), ref[2]))
        peak = float(ref[0].abs().max())
        d_img = float((real - ref[0]).abs().max()) / peak if peak > 0 else 0.0
        vis_ref = ref[2].view(torch.complex64)
        nan_same = bool(torch.equal(torch.isnan(wl_.vis_out), torch.isnan(vis_ref)))
        ok = ~torch.isnan(vis_ref)
        vpeak = float(vis_ref[ok].abs().max()) if bool(ok.any()) else 0.0
        d_vis = float((wl_.vis_out[ok] - vis_ref[ok]).abs().max()) / vpeak if vpeak > 0 else 0.0
        check["last_step_matches_first"] = nan_same and d_img < 1e-5 and d_vis < 1e-5
        info = {"last_step_equals_first_bitwise": exact,
                "last_step_max_rel_diff": [d_img, d_vis]}
        del ref, vis_ref
        torch.cuda.empty_cache()
    else:
        info = {}
    if not all(check.values()):
        raise RuntimeError(f"{precision}: output check failed: {check} {info}")
    check.update(info)
    # one more step with per-phase statistics (stats synchronise)
    step(True)
    # a major cycle reuses the plan (the uvw do not change between cycles):
    # invert + predict only, over the same K steps
    ms_cycle = None
    if cycle:
        g.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            g.invert(wl_.vis, 1.0, real, imag)
            g.predict(real, imag, wl_.vis_out)
        g.synchronize()
        torch.cuda.synchronize()
        ms_cycle = (time.perf_counter() - t0) / steps
    g.close()
    return elapsed / steps, npass, phases, stats_inv, stats_pred, ms_cycle, check


def cpu_baseline(scale_rows=2, threads=None):
    """The oracle's C++ restatement of WSMSGridder (float, layer-routed
    worker lanes, per-thread FFTs) on a bounded sample: Stokes I at 4096^2
    with 64 w-layers, invert + predict, on every n-th row of the workload.
    The fixed per-image cost (FFTs, w-correction) is measured with one row;
    the per-visibility cost from the sample; the full four-image workload
    time is 4 * fixed + per_vis * N_vis. The sample's per-visibility part
    must stand well above the run-to-run variation of the fixed part (about
    a second of 16 on a shared host): every 2nd row gives about 10 s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as po
    from wsclean_amd import synth
    threads = threads or min(16, os.cpu_count() or 1)
    obs = synth.config2(scale=1.0)
    W = 4096
    pix = 4.0 * np.pi / (180.0 * 3600.0)
    wl = obs.wavelengths
    s = po.ms_settings(W, W, pix, pix, nw=64, ncpus=threads, memory=64e9)
    rng = np.random.default_rng(5)

    def run(rows):
        uvw = obs.uvw_m[rows]
        data = (rng.standard_normal((len(rows), obs.nchan)) +
                1j * rng.standard_normal((len(rows), obs.nchan))).astype(np.complex64)
        t0 = time.perf_counter()
        img, _, st = po.ms_invert(s, uvw, wl, data)
        po.ms_predict(s, uvw, wl, img)
        return time.perf_counter() - t0, len(rows) * obs.nchan

    t_fixed, n0 = run(np.arange(0, 1))
    rows = np.arange(0, obs.nrow, scale_rows)
    t_sample, ns = run(rows)
    per_vis = max(t_sample - t_fixed, 1e-9) / max(ns - n0, 1)
    nvis_full = obs.nrow * obs.nchan * 4
    t_full = 4 * t_fixed + per_vis * nvis_full
    return {
        "value": nvis_full / t_full, "unit": "visibilities/s", "cores": threads, "kind": "port",
        "sample": (f"oracle C++ restatement of WSMSGridder (float, {threads} threads), Stokes I "
                   f"4096^2 nW=64 invert+predict: 1 row (fixed cost {t_fixed:.2f} s) and one row "
                   f"in {scale_rows} ({ns} vis, {t_sample:.2f} s); full 4-correlation time "
                   f"modelled as 4*fixed + per-vis*{nvis_full}"),
        "seconds": t_fixed + t_sample,
    }


def facet_leg(args, dev):
    """Facets of one major cycle: invert + predict of 16 uneven facets over
    the same samples, (a) captured into one HIP graph per direction, (b)
    launched facet by facet with the same kernels, (c) in the plain
    per-channel FP32 form, facet by facet. w-layers per facet are the
    reference's suggestion; tile size and work chunk come from the autotuner
    on the largest facet. Plans are built once per facet
    set (the samples do not change between major cycles) and timed apart."""
    import torch
    from wsclean_amd import synth
    from wsclean_amd.autotune import autotune
    from wsclean_amd.facets import FacetBatch
    import wsclean_amd
    obs = synth.config2(scale=args.facet_scale)
    pix = Workload.PIXEL
    nrow, nchan = obs.nrow, obs.nchan
    nvis = nrow * nchan * 4
    uvw = torch.from_numpy(obs.uvw_m).to(dev)
    wl = torch.from_numpy(obs.wavelengths).to(dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(4040)
    vis = torch.randn((nrow, nchan, 4), dtype=torch.complex64, device=dev, generator=gen)
    facets = synth.config4_facets()
    big = max(facets, key=lambda f: f.width)
    t = time.perf_counter()
    tuned = autotune(big.width, big.height, pix, pix, uvw, wl, vis, n_images=4,
                     complex_images=Workload.COMPLEX, l_shift=big.l_shift, m_shift=big.m_shift,
                     chunks=(0, 8192, 16384), repeats=1)
    log(f"facets: autotune on {big.width}^2 in {time.perf_counter() - t:.1f} s: tile "
        f"{tuned['tile_size']} chunk {tuned['work_chunk']}")
    torch.cuda.synchronize()
    t = time.perf_counter()
    fb = FacetBatch(facets, pix, pix, uvw, wl, n_images=4, complex_images=Workload.COMPLEX,
                    tile_size=tuned["tile_size"], work_chunk=tuned["work_chunk"])
    torch.cuda.synchronize()
    plan_ms = (time.perf_counter() - t) * 1e3
    log(f"facets: {len(facets)} planned in {plan_ms:.0f} ms; sizes "
        f"{[f.width for f in facets]}, w-layers {fb.n_w}, passes {fb.passes}")
    re, im = fb.images(), fb.images()
    outs = [torch.empty_like(vis) for _ in facets]

    def timed(step, k):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(k):
            step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / k * 1e3

    def eager():
        fb.invert(vis, 1.0, re, im)
        fb.predict(re, im, outs)

    eager()
    eager_ms = timed(eager, args.steps)
    ref_img = re[0].clone()
    ginv = fb.capture_invert(vis, 1.0, re, im)
    gpred = fb.capture_predict(re, im, outs)

    def graph():
        ginv.launch()
        gpred.launch()

    graph()
    graph_ms = timed(graph, args.steps)
    same = bool(torch.equal(re[0], ref_img))
    ginv.close()
    gpred.close()
    n_w = list(fb.n_w)
    fb.close()
    del re, im, outs
    torch.cuda.empty_cache()
    log(f"facets: graph {graph_ms:.1f} ms/step, eager {eager_ms:.1f} ms/step")

    plain = None
    if not args.no_plain:
        # the plain form's cost is linear in the channel count (one launch
        # per channel) plus a fixed per-facet part: time 4 and 8 channels
        # and extrapolate to all of them
        def plain_ms(c):
            v = vis[:, :c].contiguous()
            wlc = wl[:c].contiguous()
            total = 0.0
            for f, nw in zip(facets, n_w):
                g = wsclean_amd.DeviceGridder(f.width, f.height, pix, pix, 7, 63,
                                              precision="plain", n_images=4,
                                              complex_images=Workload.COMPLEX,
                                              l_shift=f.l_shift, m_shift=f.m_shift)
                g.prepare_w_layers(nw, fb.w_limits[0], fb.w_limits[1])
                r = torch.zeros((4, f.height, f.width), dtype=torch.float32, device=dev)
                i_ = torch.zeros_like(r)
                o = torch.empty_like(v)
                g.plan(uvw, wlc)
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                g.invert(v, 1.0, r, i_)
                g.predict(r, i_, o)
                g.synchronize()
                total += time.perf_counter() - t0
                g.close()
            return total * 1e3
        try:
            p4, p8 = plain_ms(4), plain_ms(8)
            p_full = p4 + (p8 - p4) / 4.0 * (nchan - 4)
            plain = {"ms_per_step": p_full, "value": nvis / (p_full * 1e-3),
                     "ms_at_4_channels": p4, "ms_at_8_channels": p8,
                     "method": f"plain per-channel FP32 form facet by facet, timed at 4 and 8 of "
                               f"{nchan} channels and extrapolated linearly"}
            log(f"facets: plain {p_full:.0f} ms/step (extrapolated from {p4:.0f}, {p8:.0f})")
        except Exception as e:  # reported, never hidden
            plain = {"error": str(e)}
    del vis
    torch.cuda.empty_cache()
    return {
        "workload": (f"{len(facets)} facets, boxes uniform in 512..2048 px at 4 arcsec, centres "
                     f"uniform within 5 deg; {obs.meta['n_baselines']} baselines x "
                     f"{obs.n_times} x 2 s x {nchan} ch x 4 corr = {nvis} visibilities; "
                     "invert + predict per step, plans built once"),
        "sizes": [f.width for f in facets], "w_layers": n_w,
        "tile_size": tuned["tile_size"], "work_chunk": tuned["work_chunk"],
        "plan_ms": plan_ms, "graph_ms_per_step": graph_ms, "eager_ms_per_step": eager_ms,
        "graph_value": nvis / (graph_ms * 1e-3), "eager_value": nvis / (eager_ms * 1e-3),
        "unit": "visibilities/s", "graph_equals_eager": same, "plain_form": plain,
        "speedup_graph_vs_plain": (plain["ms_per_step"] / graph_ms
                                   if plain and "ms_per_step" in plain else None),
    }


def plugin_leg(w, steps=1, batch_rows=65536):
    """The drop-in plugin's path at this workload, host memory at both ends:
    what HipWSMSGridder does per WSClean Invert and Predict call once its
    reader has produced the (pre-weighted) rows. Through the staged C ABI the
    rows go batch by batch into two pinned host buffers and cross the bus
    asynchronously while the next batch is copied (hws_stage_rows), are
    planned once per call (hws_stage_end), inverted or predicted on the
    device, and the four images or the predicted visibilities come back to
    host memory. One step = invert (stage uvw + visibilities, plan, invert,
    images to host) + predict (stage uvw, plan, models from host, predict,
    visibilities to host). The visibilities are the workload's, unweighted
    (the plugin weights its rows on the host before staging them)."""
    import ctypes as C
    import wsclean_amd
    uvw = np.ascontiguousarray(w.uvw.cpu().numpy())
    wl = np.ascontiguousarray(w.wl.cpu().numpy())
    vis = np.ascontiguousarray(w.vis.cpu().numpy())
    out = np.empty_like(vis)
    real = np.zeros((4, w.H, w.W), np.float32)
    imag = np.zeros_like(real)
    g = wsclean_amd.WStackingGridder(w.W, w.H, w.PIXEL, w.PIXEL, 7, 63, precision="mixed",
                                     n_images=4, complex_images=w.COMPLEX)
    g.prepare_w_layers(Workload.NW, 0.0, w.min_w, w.max_w)
    L, h = g.L, g.h
    pd, pf = C.POINTER(C.c_double), C.POINTER(C.c_float)
    nrow, nchan = w.nrow, w.nchan
    phases = {"invert_stage_s": 0.0, "invert_s": 0.0, "predict_stage_s": 0.0, "predict_s": 0.0}

    def stage(with_vis):
        g._check(L.hws_stage_begin(h, nrow, nchan, wl.ctypes.data_as(pd), int(with_vis)))
        for r0 in range(0, nrow, batch_rows):
            r1 = min(nrow, r0 + batch_rows)
            g._check(L.hws_stage_rows(h, r1 - r0, uvw[r0:r1].ctypes.data_as(pd),
                                      vis[r0:r1].ctypes.data_as(pf) if with_vis else None))
        g._check(L.hws_stage_end(h))

    def step(record):
        t = time.perf_counter()
        stage(True)
        t1 = time.perf_counter()
        g._check(L.hws_invert_staged(h, 1.0, w.W, w.H, w.W, w.H, real.ctypes.data_as(pf),
                                     imag.ctypes.data_as(pf)))
        t2 = time.perf_counter()
        stage(False)
        t3 = time.perf_counter()
        g._check(L.hws_predict_staged(h, real.ctypes.data_as(pf), imag.ctypes.data_as(pf),
                                      w.W, w.H, w.W, w.H, out.ctypes.data_as(pf)))
        t4 = time.perf_counter()
        if record:
            phases["invert_stage_s"] += t1 - t
            phases["invert_s"] += t2 - t1
            phases["predict_stage_s"] += t3 - t2
            phases["predict_s"] += t4 - t3

    step(False)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(True)
    elapsed = (time.perf_counter() - t0) / steps
    g.close()
    finite = bool(np.isfinite(real).all() and np.isfinite(imag[list(w.COMPLEX)]).all())
    return {"value": w.nvis / elapsed, "unit": "visibilities/s", "s_per_step": elapsed,
            "steps": steps, "phases_s": {k: round(v / steps, 3) for k, v in phases.items()},
            "bytes_to_device_per_step": int(2 * uvw.nbytes + vis.nbytes + real.nbytes * 2),
            "bytes_to_host_per_step": int(out.nbytes + real.nbytes * 2),
            "images_finite": finite,
            "what": "staged host rows -> plan -> invert -> host images; staged uvw -> plan -> "
                    "host models -> predict -> host visibilities (mixed precision, unweighted, "
                    "batches of %d rows)" % batch_rows}


def traffic_from_profiles():
    """HBM bytes per step of each kernel class, measured with rocprofv3 --pmc
    FETCH_SIZE and WRITE_SIZE over one profiled step and scaled by the same
    counters on a device copy of known size; read from the committed summary
    (profiles/traffic_4096_step.md: lines "traffic_bytes_per_step[<class>]: <bytes>"),
    or an empty dict when it is absent."""
    path = os.path.join(ROOT, "profiles", "traffic_4096_step.md")
    out = {}
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_step["):
                name = line[len("traffic_bytes_per_step["):line.index("]")]
                out[name] = int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (--gpus {args.gpus} requested); the multi-GPU "
              "layout is described in DESIGN.md and not measured here.", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(args.copy_calibration // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.facets_only:
        print(json.dumps({"facets": facet_leg(args, dev)}))
        return
    t0 = time.perf_counter()
    w = Workload(args.scale, dev)
    if args.padded_only:
        size, n_w, os_ = (int(v) for v in args.padded_only.split(","))
        ms, npass, ph, _, _, _, _ = run_form(w, args.precision, args.steps, args.warmup, args.tile,
                                             dev, args.chunk, cycle=False, n_w=n_w,
                                             oversampling=os_, size=size)
        print(json.dumps({"padded": {"gridded_size": size, "w_layers": n_w, "oversampling": os_,
                                     "value": w.nvis / ms, "unit": "visibilities/s",
                                     "ms_per_step": ms * 1e3, "passes": npass,
                                     "phases_ms": {k: round(v, 3) for k, v in ph.items()}}}))
        return
    log(f"workload ready in {time.perf_counter() - t0:.1f} s: {w.nrow} rows x {w.nchan} ch x 4 "
        f"= {w.nvis} vis, w in [{w.min_w:.1f}, {w.max_w:.1f}]")

    tuned = None
    if args.autotune:
        from wsclean_amd.autotune import autotune
        t = time.perf_counter()
        tuned = autotune(w.W, w.H, w.PIXEL, w.PIXEL, w.uvw, w.wl, w.vis, n_images=4,
                         complex_images=w.COMPLEX, precision=args.precision,
                         min_layers=Workload.NW, chunks=(4096, 8192), repeats=1)
        args.tile, args.chunk = tuned["tile_size"], tuned["work_chunk"]
        log(f"autotune in {time.perf_counter() - t:.1f} s: tile {args.tile}, chunk {args.chunk}, "
            f"suggested layers {tuned['n_w_layers']}; trials {tuned['trials']}")
    ms_step, npass, ph, st_inv, st_pred, ms_cycle, out_check = run_form(
        w, args.precision, args.steps, args.warmup, args.tile, dev, args.chunk)
    log(f"{args.precision}: {ms_step * 1e3:.1f} ms/step, passes={npass}, phases(ms)={ph}")
    value = w.nvis / ms_step

    # Rooflines of the step's kernels from the phase times (HIP events the
    # library records around each phase on the stream it launches on) and the
    # algorithmic bytes or flops of each phase, the sparse fractions counted
    # by the plan. Per image of kind k (XX, YY real; XY, YX complex), with nW
    # layers of HW cells:
    #   forward transform (fftRowsModelLayersKernel + fftColsBlockedSparseKernel):
    #     the model read once (4 B per pixel, 8 complex), 8 B written per cell
    #     of a needed tile column by the rows pass, the same read by the columns
    #     pass, 8 B written per cell of a needed tile;
    #   backward rows (fftRowsBlockedSparseKernel): 8 B read per cell of an
    #     occupied tile, 8 B written per cell of a live tile row;
    #   columns + projection (fftColsProjectKernel): 8 B read per cell of a live
    #     tile row; the accumulators read and written once (sum and
    #     compensation, 8 B each way; the imaginary ones for every image);
    #   grid (gridMfmaF16Kernel): 245 FP32 flops per gridded visibility (49
    #     taps, 5 flops), against the FP32 matrix peak (the kernel forms them as
    #     three FP16 products of split operands with FP32 sums);
    #   degrid (degridTilesKernel): 245 FP32 flops per predicted visibility,
    #     against the FP32 vector peak;
    #   binning (count + scatter passes, both kinds): the uvw read twice per
    #     kind, 8 B written per record.
    nw = Workload.NW
    HW = Workload.W * Workload.H
    need = st_pred["tiles_needed"] if st_pred else [0, 0]
    occ = st_pred["tiles_occupied"] if st_pred else [0, 0]
    ntiles = st_pred["tiles_total"] if st_pred else 0
    rlive = st_pred["tile_rows_live"] if st_pred else [0, 0]
    nrows_t = st_pred["tile_rows_total"] if st_pred else 0
    cneed = st_pred["tile_cols_needed"] if st_pred else [0, 0]
    ncols_t = st_pred["tile_cols_total"] if st_pred else 0
    fwd_bytes = rows_bytes = proj_bytes = 0.0
    for k, nimg in ((0, 2), (1, 2)):
        f_need = need[k] / ntiles if ntiles else 1.0
        f_occ = occ[k] / ntiles if ntiles else 1.0
        f_cols = cneed[k] / ncols_t if ncols_t else 1.0
        f_rows = rlive[k] / nrows_t if nrows_t else 1.0
        fwd_bytes += nimg * (nw * HW * (16 * f_cols + 8 * f_need) + HW * 4 * (1 + k))
        rows_bytes += nimg * nw * HW * 8 * (f_occ + f_rows)
        proj_bytes += nimg * (nw * HW * 8 * f_rows + HW * 32)
    gridded = st_inv["samples_gridded"] if st_inv else 0
    flops = gridded * 4 * 245  # four visibilities per gridded (row, channel) sample
    records = 2 * 1.89 * w.nrow * w.nchan  # both kinds (the plan's records per sample)
    bin_bytes = 2 * 2 * w.nrow * 24 + records * 8

    def entry(name, kernels, ms, bound, amount, peak):
        unit = "GB/s" if bound == "hbm" else "TFLOP/s"
        scale = 1e9 if bound == "hbm" else 1e12
        achieved = amount / (ms * 1e-3) / scale if ms > 0 else 0.0
        e = {"name": name, "kernels": kernels, "ms_per_step": round(ms, 3), "bound": bound,
             "achieved": round(achieved, 2), "peak": peak, "unit": unit,
             "frac": round(achieved / peak, 4)}
        e["bytes" if bound == "hbm" else "flops"] = amount
        return e

    rooflines = [
        entry("forward layer transform", "fftRowsModelLayersKernel + fftColsBlockedSparseKernel",
              ph["fft_pred"], "hbm", fwd_bytes, HBM_PEAK_GBS),
        entry("backward rows", "fftRowsBlockedSparseKernel", ph["fft_inv"], "hbm", rows_bytes,
              HBM_PEAK_GBS),
        entry("backward columns + w-projection", "fftColsProjectKernel", ph["project"], "hbm",
              proj_bytes, HBM_PEAK_GBS),
        entry("grid", "gridMfmaF16Kernel + reducePartialsKernel", ph["grid"], "mfma", flops,
              FP32_PEAK_TFLOPS),
        entry("degrid", "degridTilesKernel", ph["degrid"], "fp32", flops, FP32_PEAK_TFLOPS),
        entry("binning", "binKernel (count, scatter) + scans", ph["bin"], "hbm", bin_bytes,
              HBM_PEAK_GBS),
    ]
    traffic = traffic_from_profiles()
    # the dominant kernel class: the one with the largest share of the step
    dom = max(rooflines, key=lambda r: r["ms_per_step"])
    roofline = {"bound": dom["bound"] if dom["bound"] in ("hbm", "mfma") else "mfma",
                "kernel": f"{dom['name']} ({dom['kernels']})", "achieved": dom["achieved"],
                "peak": dom["peak"], "unit": dom["unit"], "frac": dom["frac"],
                "traffic": traffic.get(dom["name"]) if traffic else None,
                "ms_per_step": dom["ms_per_step"]}
    phase_ms = {k: round(v, 3) for k, v in ph.items()}
    dominant = dom["name"]
    grid_tflops = flops / (ph["grid"] * 1e-3) / 1e12 if ph["grid"] > 0 else 0.0

    plain = None
    if not args.no_plain and args.precision != "plain":
        try:
            ms_plain, _, ph_plain, _, _, _, _ = run_form(w, "plain", 1, 0, args.tile, dev,
                                                         cycle=False)
            plain = {"value": w.nvis / ms_plain, "ms_per_step": ms_plain * 1e3,
                     "phases_ms": {k: round(v, 3) for k, v in ph_plain.items()}}
            log(f"plain: {ms_plain * 1e3:.1f} ms/step")
        except Exception as e:  # reported, never hidden
            plain = {"error": str(e)}

    accuracy = None
    if not args.no_accuracy:
        from wsclean_amd import accuracy as acc
        t = time.perf_counter()
        try:
            accuracy = acc.report(device="cuda:0")
            # the w-layer count at which the error meets 1e-5 of the peak, and
            # what this workload runs at with its layers scaled alike (its 64
            # layers stand where the precision protocol's 32 stand)
            sw = acc.sweep(device="cuda:0", paddings=(1.0, 1.25, 1.5, 2.0))
            accuracy["w_sweep"] = sw
            ok = [r for r in sw["rows"] if r["image_rms_rel"] <= 1e-5]
            if ok:
                # the cheapest setting that meets the bound: layer transforms
                # cost n_w * size^2
                best = min(ok, key=lambda r: (r["n_w_layers"] * r["gridded_size"] ** 2,
                                              r["oversampling"]))
                n_w2 = int(np.ceil(Workload.NW * best["n_w_layers"] / 32))
                size2 = acc.padded_size(Workload.W, best["padding"])
                ms2, _, _, _, _, _, _ = run_form(w, args.precision, 2, 1, args.tile, dev,
                                                 args.chunk, cycle=False, n_w=n_w2,
                                                 oversampling=best["oversampling"],
                                                 size=size2)
                accuracy["meets_1e-5"] = dict(best, workload_w_layers=n_w2,
                                              workload_gridded_size=size2,
                                              workload_value=w.nvis / ms2,
                                              workload_ms_per_step=ms2 * 1e3)
            else:
                accuracy["meets_1e-5"] = None
        except Exception as e:  # reported, never hidden
            accuracy = {"error": f"{type(e).__name__}: {e}"}
        log(f"accuracy report in {time.perf_counter() - t:.1f} s: {accuracy}")

    plugin = None
    if not args.no_plugin:
        t = time.perf_counter()
        try:
            plugin = plugin_leg(w)
        except Exception as e:  # reported, never hidden
            plugin = {"error": f"{type(e).__name__}: {e}"}
        log(f"plugin leg in {time.perf_counter() - t:.1f} s: {plugin}")

    facets = None
    if not args.no_facets:
        del w.vis_out
        torch.cuda.empty_cache()
        t = time.perf_counter()
        try:
            facets = facet_leg(args, dev)
        except Exception as e:  # reported, never hidden
            facets = {"error": f"{type(e).__name__}: {e}"}
        log(f"facet leg in {time.perf_counter() - t:.1f} s")

    cpu = None
    if not args.no_cpu:
        t = time.perf_counter()
        cpu = cpu_baseline()
        log(f"cpu baseline {cpu['value']:.3e} vis/s in {time.perf_counter() - t:.1f} s")

    line = {
        "metric": METRIC, "value": value, "unit": "visibilities/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 (fp64 positions/phases; gridding products on fp16 matrix cores as "
                 "hi/lo-split operands with fp32 sums)"
        if args.precision == "mixed" else ("fp32 (fp64 positions/phases)"
                                           if args.precision == "reference" else "fp32"),
        "data": "synthetic: seeded uvw of the 62-station 40 km log-spiral array, seeded "
                "complex Gaussian visibility values",
        "config": {"workload": "4096^2 x 64 w-layers, 1891 baselines x 1800 x 64 ch x 4 corr "
                               "(XX,YY real; XY,YX complex), invert+predict",
                   "precision": args.precision, "n_vis": w.nvis, "passes": npass,
                   "scale": args.scale},
        "roofline": roofline,
        "rooflines": rooflines,
        "dominant_phase": dominant,
        "phases_ms_per_step": phase_ms,
        "grid_fp32_tflops": round(grid_tflops, 3),
        "grid_fp32_frac": round(grid_tflops / FP32_PEAK_TFLOPS, 4),
        "samples_gridded": gridded,
        "output_check": out_check,
        "samples_off_grid": st_inv["samples_off_grid"] if st_inv else None,
        "major_cycle": {"ms_per_cycle": ms_cycle * 1e3, "value": w.nvis / ms_cycle,
                        "unit": "visibilities/s",
                        "what": "invert + predict with the binning plan reused (uvw are fixed "
                                "across major cycles); value above includes the plan"},
        "plugin_path": plugin,
        "plain_form": plain,
        "speedup_vs_plain": (value / plain["value"]) if plain and "value" in plain else None,
        "rms_err_vs_fp64_dft": (accuracy.get(f"{args.precision}_os63", {}).get("image_rms_rel")
                                if accuracy else None),
        "accuracy": accuracy,
        "autotune": ({k: tuned[k] for k in ("tile_size", "work_chunk", "trials")}
                     if tuned else None),
        "facets": facets,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
