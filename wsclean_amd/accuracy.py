"""Accuracy of the w-stacking gridder against the FP64 direct Fourier
transform (the error statement DirectMSGridder<double> makes possible,
gridding/directmsgridder.cpp:16-114).

Protocol (the precision workload): the 64-antenna 3 km-disc observation
(2016 baselines, dec -30 deg, 100 x 10 s, 16 channels 120-168 MHz), Stokes I
of 20 point sources (log-uniform 0.01-10 Jy) plus complex Gaussian noise of
0.1 Jy; a seeded random subset of 1e5 (row, channel) samples; natural
weighting. The dirty image is gridded at 1024^2, 20 arcsec, 32 w-layers,
Kaiser-Bessel K=7, OS=63 (and OS=1023), and its central 512^2 cutout is
compared with the FP64 direct transform on those pixels. A model of the same 20 sources on
pixel centres is predicted for the same samples and compared with the FP64
direct predict. Errors are relative to the peak of the direct result:
  image:  RMS(I - I_dft) / max|I_dft| and max|I - I_dft| / max|I_dft|, over
          the cutout and over its central 128^2
  vis:    RMS|V - V_dft| / max|V_dft| and max|V - V_dft| / max|V_dft|;
and, for the FP32 modes, the RMS difference from the reference-precision
mode (FP64 positions and phases) relative to the same peaks: the share of
the error that the arithmetic, not the algorithm, contributes.
Everything runs on the GPU: the gridder through its C ABI, the direct
transform through hws_direct_*_device.
"""
from __future__ import annotations

import math

import numpy as np

from . import synth

ARCSEC = np.pi / (180.0 * 3600.0)


def precision_workload(n_samples=100_000, seed=303, n_sources=20, noise=0.1,
                       cutout=512, width=1024):
    """Seeded inputs of the protocol; uvw in wavelengths, one sample per row."""
    obs = synth.config1()
    rng = np.random.default_rng(seed)
    n_all = obs.nrow * obs.nchan
    idx = np.sort(rng.choice(n_all, size=min(n_samples, n_all), replace=False))
    rows, chans = idx // obs.nchan, idx % obs.nchan
    uvw_l = obs.uvw_m[rows] / obs.wavelengths[chans][:, None]
    dl = 20.0 * ARCSEC
    # sources on the pixel grid of the compared cutout (kept 8 px from its edge)
    half = cutout // 2 - 8
    px = rng.integers(-half, half, size=(n_sources, 2))
    l = -px[:, 0] * dl  # l = (W/2 - x) dl with x = W/2 + px
    m = px[:, 1] * dl   # m = (y - H/2) dm with y = H/2 + py
    lo, hi = np.log(0.01), np.log(10.0)
    flux = np.exp(rng.uniform(lo, hi, n_sources))
    vis = synth.source_visibilities(uvw_l, l, m, flux)
    vis = vis + noise * (rng.standard_normal(vis.shape) + 1j * rng.standard_normal(vis.shape))
    model = np.zeros((width, width), np.float32)
    np.add.at(model, (width // 2 + px[:, 1], width // 2 + px[:, 0]), flux.astype(np.float32))
    return dict(uvw_lambda=uvw_l, vis=vis.astype(np.complex64), model=model, dl=dl,
                width=width, cutout=cutout, n_w=32, flux=flux, pixels=px)


def direct_reference(wk, device="cuda:0"):
    """The FP64 direct transform of a precision workload: the dirty image on
    the central cutout (normalised by 1/n) and the predicted visibilities of
    its model, as numpy arrays."""
    import torch
    from .direct import direct_invert, direct_predict
    n = wk["uvw_lambda"].shape[0]
    dev = torch.device(device)
    uvw = torch.from_numpy(np.ascontiguousarray(wk["uvw_lambda"])).to(dev)
    one = torch.ones(1, dtype=torch.float64, device=dev)  # uvw already in wavelengths
    vis = torch.from_numpy(wk["vis"]).to(dev).reshape(n, 1)
    model = torch.from_numpy(wk["model"]).to(dev)
    cut, dl = wk["cutout"], wk["dl"]
    img = direct_invert(uvw, one, vis, cut, cut, dl, dl, multiplication_factor=1.0 / n)
    v = direct_predict(model, uvw, one, dl, dl)
    torch.cuda.synchronize()
    return img.cpu().numpy(), v.cpu().numpy().reshape(-1)


def errors(image, vis, dft_img, dft_vis, cutout):
    """(image RMS, vis RMS) relative to the direct peaks, image on the
    central cutout of a full-size image, as report() computes them."""
    lo = (image.shape[-1] - cutout) // 2
    im = np.asarray(image, np.float64)[lo:lo + cutout, lo:lo + cutout]
    pv = np.asarray(vis, np.complex128).reshape(-1)
    ok = np.isfinite(pv)
    return _rel(im, dft_img)[0], _rel(pv[ok], dft_vis[ok])[0]


def _rel(a, b):
    peak = float(np.max(np.abs(b)))
    d = np.abs(a - b)
    return float(np.sqrt(np.mean(d * d)) / peak), float(np.max(d) / peak), peak


def report(device="cuda:0", precisions=("mixed", "reference", "plain"), n_samples=100_000,
           seed=303, oversamplings=(63, 1023)):
    """Runs the protocol and returns a dict of errors per precision mode and
    kernel oversampling (63: the benchmark setting; 1023: the reference's
    default, which lowers the kernel's own approximation error)."""
    import torch
    from .gridder import DeviceGridder

    wk = precision_workload(n_samples=n_samples, seed=seed)
    W, dl, cut = wk["width"], wk["dl"], wk["cutout"]
    n = wk["uvw_lambda"].shape[0]
    dev = torch.device(device)
    uvw = torch.from_numpy(np.ascontiguousarray(wk["uvw_lambda"])).to(dev)
    one = torch.ones(1, dtype=torch.float64, device=dev)  # uvw already in wavelengths
    vis = torch.from_numpy(wk["vis"]).to(dev).reshape(n, 1, 1)
    model = torch.from_numpy(wk["model"]).to(dev)
    aw = np.abs(wk["uvw_lambda"][:, 2])
    w_lo, w_hi = float(aw.min()), float(aw.max())

    dft_img, dft_vis = direct_reference(wk, device)

    out = {"protocol": (f"{n} random (row, channel) samples of the 2016-baseline 3 km-disc "
                        "observation, 20 point sources + 0.1 Jy noise, natural weights; "
                        f"{W}^2 at 20 arcsec, {wk['n_w']} w-layers, KB K=7 (OS in the key); image "
                        f"errors on the central {cut}^2 vs the FP64 direct FT, visibility "
                        "errors vs the FP64 direct predict; relative to the direct peak"),
           "dft_image_peak": float(np.max(np.abs(dft_img))),
           "dft_vis_peak": float(np.max(np.abs(dft_vis)))}
    lo = (W - cut) // 2
    inner = slice(cut // 2 - cut // 8, cut // 2 + cut // 8)  # central (cut/4)^2
    order = ["reference"] + [p for p in precisions if p != "reference"]
    for os_ in oversamplings:
        base = None
        for prec in order:
            g = DeviceGridder(W, W, dl, dl, 7, os_, precision=prec, n_images=1)
            g.prepare_w_layers(wk["n_w"], w_lo, w_hi)
            img = torch.zeros((1, W, W), dtype=torch.float32, device=dev)
            g.plan(uvw, one)
            g.invert(vis, 1.0 / n, img)
            pred = torch.empty_like(vis)
            g.predict(model.reshape(1, W, W), None, pred)
            g.synchronize()
            g.close()
            im = img.cpu().numpy()[0, lo:lo + cut, lo:lo + cut].astype(np.float64)
            pv = pred.cpu().numpy().reshape(-1).astype(np.complex128)
            ok = np.isfinite(pv)
            i_rms, i_max, peak = _rel(im, dft_img)
            v_rms, v_max, vpeak = _rel(pv[ok], dft_vis[ok])
            d_in = im[inner, inner] - dft_img[inner, inner]
            rec = {"image_rms_rel": i_rms, "image_max_rel": i_max,
                   "image_inner_rms_rel": float(np.sqrt(np.mean(d_in * d_in)) / peak),
                   "vis_rms_rel": v_rms, "vis_max_rel": v_max,
                   "vis_not_predicted": int((~ok).sum())}
            if base is None:
                base = (im, pv)
            else:
                di = im - base[0]
                dv = np.abs(pv - base[1])
                rec["image_rms_rel_vs_reference_mode"] = float(np.sqrt(np.mean(di * di)) / peak)
                rec["vis_rms_rel_vs_reference_mode"] = float(
                    np.sqrt(np.nanmean(dv * dv)) / vpeak)
            if prec in precisions:
                out[f"{prec}_os{os_}"] = rec
    return out


def _image_error(W, H, dl, dm, uvw, one, vis, n, n_w, w_lo, w_hi, dft_img, cutout, os_,
                 precision, kernel_size, device, n_images=1):
    import torch
    from .gridder import DeviceGridder
    g = DeviceGridder(W, H, dl, dm, kernel_size, os_, precision=precision, n_images=n_images)
    g.prepare_w_layers(n_w, w_lo, w_hi)
    img = torch.zeros((n_images, H, W), dtype=torch.float32, device=torch.device(device))
    g.plan(uvw, one)
    g.invert(vis, 1.0 / n, img)
    g.synchronize()
    g.close()
    lo_y, lo_x = (H - cutout) // 2, (W - cutout) // 2
    im = img.cpu().numpy()[0, lo_y:lo_y + cutout, lo_x:lo_x + cutout].astype(np.float64)
    return _rel(im, dft_img)[0]


def padded_size(width, padding, multiple=16):
    """The smallest gridded size >= width * padding that is a multiple of
    `multiple` (the uv tile) and 7-smooth (the transforms' radices), as the
    reference pads an image before trimming it (WSClean's -padding)."""
    n = int(math.ceil(width * padding / multiple - 1e-9)) * multiple
    while True:
        m = n
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        if m == 1:
            return n
        n += multiple


def sweep(device="cuda:0", n_ws=(32, 48, 64, 96, 128, 192, 256), oversamplings=(63, 1023),
          precision="mixed", n_samples=100_000, seed=303, paddings=(1.0,),
          padded_n_ws=(32, 64, 128)):
    """Image and visibility errors of the precision protocol (the report's
    workload) against the FP64 direct transform, for every w-layer count in
    n_ws and every kernel oversampling: where the w-stacking approximation
    (each sample on its nearest layer) stops dominating the error, and what
    the extra layers cost. With paddings above 1 the image is also gridded
    at padded_size(1024, padding) with the same pixels and the same central
    cutout compared (the model embedded at the centre for the predict), for
    the layer counts in padded_n_ws: the kernel's aliasing at the image edge,
    which more layers do not remove, falls with the padding. The reference's
    suggestion for this workload is reported beside it (suggested_w_layers
    in autotune)."""
    import torch
    from .autotune import suggested_w_layers
    from .gridder import DeviceGridder

    wk = precision_workload(n_samples=n_samples, seed=seed)
    W, dl, cut = wk["width"], wk["dl"], wk["cutout"]
    n = wk["uvw_lambda"].shape[0]
    dev = torch.device(device)
    uvw = torch.from_numpy(np.ascontiguousarray(wk["uvw_lambda"])).to(dev)
    one = torch.ones(1, dtype=torch.float64, device=dev)
    vis = torch.from_numpy(wk["vis"]).to(dev).reshape(n, 1, 1)
    model = torch.from_numpy(wk["model"]).to(dev)
    aw = np.abs(wk["uvw_lambda"][:, 2])
    w_lo, w_hi = float(aw.min()), float(aw.max())
    dft_img, dft_vis = direct_reference(wk, device)
    rows = []
    for padding in paddings:
        Wp = W if padding <= 1.0 else padded_size(W, padding)
        model_p = model
        if Wp != W:
            model_p = torch.zeros((Wp, Wp), dtype=torch.float32, device=dev)
            o = (Wp - W) // 2
            model_p[o:o + W, o:o + W] = model
        for os_ in oversamplings:
            for n_w in (n_ws if Wp == W else padded_n_ws):
                g = DeviceGridder(Wp, Wp, dl, dl, 7, os_, precision=precision, n_images=1)
                g.prepare_w_layers(n_w, w_lo, w_hi)
                img = torch.zeros((1, Wp, Wp), dtype=torch.float32, device=dev)
                g.plan(uvw, one)
                g.invert(vis, 1.0 / n, img)
                pred = torch.empty_like(vis)
                g.predict(model_p.reshape(1, Wp, Wp), None, pred)
                g.synchronize()
                g.close()
                i_rms, v_rms = errors(img.cpu().numpy()[0], pred.cpu().numpy(), dft_img,
                                      dft_vis, cut)
                rows.append({"n_w_layers": n_w, "oversampling": os_, "padding": Wp / W,
                             "gridded_size": Wp, "image_rms_rel": i_rms, "vis_rms_rel": v_rms})
    return {"precision": precision,
            "suggested_w_layers": suggested_w_layers(W, W, dl, dl, w_lo, w_hi),
            "w_range": (w_lo, w_hi), "rows": rows}


def w_layers_for_accuracy(width, height, dl, dm, uvw_m, wavelengths, vis, bound, *,
                          n_w_start, n_w_max=1024, growth=1.5, oversampling=63,
                          kernel_size=7, precision="mixed", n_samples=20_000, cutout=None,
                          seed=0):
    """The smallest w-layer count from n_w_start up (growing by `growth`) at
    which the dirty image of a random subset of the caller's samples is
    within `bound` (RMS over the central cutout, relative to the peak) of
    the FP64 direct transform of the same samples. uvw_m [nrow, 3],
    wavelengths [nchan] and vis [nrow, nchan, n_images] are device tensors
    as DeviceGridder takes them; the first image's visibilities are used.
    Returns (n_w or None, [(n_w, error), ...])."""
    import torch
    from .direct import direct_invert
    dev = uvw_m.device
    nrow, nchan = uvw_m.shape[0], wavelengths.shape[0]
    rng = np.random.default_rng(seed)
    total = nrow * nchan
    pick = np.sort(rng.choice(total, size=min(n_samples, total), replace=False))
    idx = torch.from_numpy(pick).to(dev)
    rows, chans = idx // nchan, idx % nchan
    uvw = (uvw_m[rows] / wavelengths[chans][:, None]).contiguous()
    v = vis.reshape(nrow * nchan, -1)[idx, 0].contiguous().reshape(-1, 1, 1)
    n = uvw.shape[0]
    one = torch.ones(1, dtype=torch.float64, device=dev)
    cut = cutout or min(width, height) // 2
    dft = direct_invert(uvw, one, v.reshape(n, 1), cut, cut, dl, dm,
                        multiplication_factor=1.0 / n)
    torch.cuda.synchronize()
    dft_img = dft.cpu().numpy()
    aw = torch.abs(uvw[:, 2])
    w_lo, w_hi = float(aw.min()), float(aw.max())
    tried = []
    n_w = max(1, int(n_w_start))
    while n_w <= n_w_max:
        e = _image_error(width, height, dl, dm, uvw, one, v, n, n_w, w_lo, w_hi, dft_img, cut,
                         oversampling, precision, kernel_size, str(dev))
        tried.append((n_w, e))
        if e <= bound:
            return n_w, tried
        n_w = max(n_w + 1, int(math.ceil(n_w * growth)))
    return None, tried
