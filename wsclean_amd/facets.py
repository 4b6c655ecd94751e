"""Facets of one major cycle, batched into one graph launch.

A facet is imaged the way the reference images it: as an image of its own
size whose phase centre is shifted to the facet centre
(scheduling/griddingtaskmanager.cpp:75-81, main/wsclean.cpp:529-535 and
2119-2130), gridded with the w-stacking gridder and its visibility rotation
(gridding/wsmsgridder.cpp:185-190). Here every facet gets its own planned
context, and the invert (or predict) of all facets of a cycle is captured
into ONE HIP graph with one branch per facet, so a major cycle costs one
launch however many facets it holds.

- ``create_facet_grid`` splits an image into a regular grid of boxes as
  CreateFacetGrid does (structures/facetutil.cpp:22-48).
- ``facet_shift`` is the facet's l/m shift (main/wsclean.cpp:529-535 with the
  centre offset of main/wsclean.cpp:2127-2130).
- ``gridded_size`` picks a facet image size the gridder takes: at least the
  box times the padding, a multiple of the uv tile edge and 7-smooth (the
  low-prime rule of math/calculatefftsize.h:4-19). The reference's own rule
  for untrimmed facet boxes lives in schaapcommon, which the reference tree
  does not vendor.
- ``FacetBatch`` plans one context per facet, runs them eagerly or captures
  them into one graph.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .autotune import suggested_w_layers
from .gridder import DeviceGridder, Graph


@dataclass
class Facet:
    """One facet: its gridded image size and the l/m shift of its centre.
    `box` is the (x0, y0, x1, y1) bounding box in the main image, if any."""
    width: int
    height: int
    l_shift: float
    m_shift: float
    box: tuple = None
    label: str = ""


def create_facet_grid(image_width, image_height, grid_width, grid_height):
    """Boxes (x0, y0, x1, y1) and labels "gx, gy" of a grid_width x
    grid_height facet grid, row by row (structures/facetutil.cpp:22-48)."""
    out = []
    for gy in range(grid_height):
        y0 = gy * image_height // grid_height
        y1 = (gy + 1) * image_height // grid_height
        for gx in range(grid_width):
            x0 = gx * image_width // grid_width
            x1 = (gx + 1) * image_width // grid_width
            out.append(((x0, y0, x1, y1), f"{gx}, {gy}"))
    return out


def facet_shift(box, trimmed_width, trimmed_height, dl, dm, l_shift=0.0, m_shift=0.0):
    """The l/m shift of a facet whose box is (x0, y0, x1, y1) in the main
    image: the centre offset in pixels is centre - size/2 (integer centre
    and half size, main/wsclean.cpp:2127-2130); l decreases and m increases
    with it (main/wsclean.cpp:529-535)."""
    x0, y0, x1, y1 = box
    cx = (x0 + x1) // 2 - trimmed_width // 2
    cy = (y0 + y1) // 2 - trimmed_height // 2
    return l_shift - cx * dl, m_shift + cy * dm


def _low_prime(n):
    for p in (2, 3, 5, 7):
        while n % p == 0 and n > 1:
            n //= p
    return n == 1


def gridded_size(n, padding=1.0, multiple=32):
    """The smallest size >= n * padding that is a multiple of `multiple` and
    has no prime factor above 7."""
    m = -(-int(round(n * padding * 1e6)) // 1000000)
    m = -(-m // multiple) * multiple
    while not _low_prime(m):
        m += multiple
    return m


def _origin(box, width, height):
    x0, y0, x1, y1 = box
    return (x0 + x1) // 2 - width // 2, (y0 + y1) // 2 - height // 2


def cut_facet(image, box, width, height):
    """The facet's own width x height image of a main-image model: the box's
    pixels placed about the facet centre (facet pixel width/2 is the box
    centre, as facet_shift assumes), zero elsewhere. Works on numpy arrays
    and torch tensors [..., H, W]."""
    x0, y0, x1, y1 = box
    ox, oy = _origin(box, width, height)
    out = image.new_zeros(image.shape[:-2] + (height, width)) if hasattr(image, "new_zeros") \
        else np.zeros(image.shape[:-2] + (height, width), image.dtype)
    out[..., y0 - oy:y1 - oy, x0 - ox:x1 - ox] = image[..., y0:y1, x0:x1]
    return out


def paste_facet(image, facet_image, box):
    """Writes the box of a facet image (trimming the rest) into the main
    image, in place."""
    x0, y0, x1, y1 = box
    ox, oy = _origin(box, facet_image.shape[-1], facet_image.shape[-2])
    image[..., y0:y1, x0:x1] = facet_image[..., y0 - oy:y1 - oy, x0 - ox:x1 - ox]
    return image


def facets_of_grid(image_width, image_height, grid_width, grid_height, dl, dm, l_shift=0.0,
                   m_shift=0.0, padding=1.0, multiple=32):
    """Facets (gridded size and shift) of a regular facet grid."""
    out = []
    for box, label in create_facet_grid(image_width, image_height, grid_width, grid_height):
        x0, y0, x1, y1 = box
        ls, ms = facet_shift(box, image_width, image_height, dl, dm, l_shift, m_shift)
        out.append(Facet(gridded_size(x1 - x0, padding, multiple),
                         gridded_size(y1 - y0, padding, multiple), ls, ms, box, label))
    return out


class FacetBatch:
    """One planned DeviceGridder per facet over the same samples.

    uvw_m float64 [nrow, 3] and wavelengths float64 [nchan] are device
    tensors; vis complex64 [nrow, nchan, n_images] is shared by all facets
    for invert; images and predicted visibilities are one tensor per facet.

    w-layers: ``n_w_layers`` (an int for all facets, or one per facet), else
    the reference's suggestion for each facet (its size and shift),
    optionally capped at ``max_w_layers`` so that every facet fits in a single
    pass. Graph capture needs a single pass per facet.
    """

    def __init__(self, facets, dl, dm, uvw_m, wavelengths, *, n_images=1, complex_images=(),
                 precision="mixed", kernel_size=7, oversampling=63, n_w_layers=None,
                 max_w_layers=None, min_w_layers=1, tile_size=0, work_chunk=0,
                 w_limits=None, device=0, stream=None):
        import torch
        self.facets = list(facets)
        if not self.facets:
            raise ValueError("no facets")
        if w_limits is None:
            aw = torch.abs(uvw_m[:, 2])
            w_limits = (float(aw.min() / wavelengths.max()), float(aw.max() / wavelengths.min()))
        min_w, max_w = w_limits
        self.w_limits = (min_w, max_w)
        self.n_images = n_images
        self.complex_images = tuple(complex_images)
        self.gridders, self.n_w, self.suggested_w = [], [], []
        try:
            for i, f in enumerate(self.facets):
                lo = -max_w if complex_images else min_w
                # the suggestion counts the trimmed size (the box), as the
                # reference's does (gridding/wsmsgridder.cpp:77-86)
                tw = f.box[2] - f.box[0] if f.box else f.width
                th = f.box[3] - f.box[1] if f.box else f.height
                sug = suggested_w_layers(tw, th, dl, dm, lo, max_w, f.l_shift, f.m_shift)
                if n_w_layers is None:
                    nw = max(sug, int(min_w_layers))
                    if max_w_layers is not None:
                        nw = min(nw, int(max_w_layers))
                elif isinstance(n_w_layers, int):
                    nw = n_w_layers
                else:
                    nw = int(n_w_layers[i])
                g = DeviceGridder(f.width, f.height, dl, dm, kernel_size, oversampling,
                                  precision=precision, device=device, n_images=n_images,
                                  complex_images=complex_images, l_shift=f.l_shift,
                                  m_shift=f.m_shift, rotate_visibilities=True,
                                  tile_size=tile_size, work_chunk=work_chunk, stream=stream)
                self.gridders.append(g)
                g.prepare_w_layers(nw, min_w, max_w)
                g.plan(uvw_m, wavelengths)
                self.n_w.append(nw)
                self.suggested_w.append(sug)
        except Exception:
            self.close()
            raise
        self.passes = [g.n_passes for g in self.gridders]

    def images(self, dtype=None):
        """Zeroed output images, one [n_images, H, W] float32 tensor per facet."""
        import torch
        dev = self.gridders[0].device
        return [torch.zeros((self.n_images, f.height, f.width), dtype=dtype or torch.float32,
                            device=torch.device("cuda", dev)) for f in self.facets]

    def invert(self, vis, multiplication_factor, reals, imags=None):
        """Eager invert of every facet, one after the other."""
        self._check_count(reals, imags)
        for i, g in enumerate(self.gridders):
            g.invert(vis, multiplication_factor, reals[i], imags[i] if imags else None)

    def predict(self, reals, imags, vis_outs):
        """Eager predict of every facet into its own visibility tensor."""
        self._check_count(reals, imags, vis_outs)
        for i, g in enumerate(self.gridders):
            g.predict(reals[i], imags[i] if imags else None, vis_outs[i])

    def _check_count(self, *lists):
        n = len(self.gridders)
        for lst in lists:
            if lst is not None and len(lst) != n:
                raise ValueError(f"expected one tensor per facet ({n}), got {len(lst)}")

    def _arrays(self, tensors):
        arr = (C.c_void_p * len(self.gridders))()
        for i in range(len(self.gridders)):
            t = tensors[i] if tensors is not None else None
            arr[i] = t.data_ptr() if t is not None else None
        return arr

    def capture_invert(self, vis, multiplication_factor, reals, imags=None):
        """One HIP graph that inverts every facet (one branch per facet).
        Valid while the plans and the tensors stay alive."""
        n = len(self.gridders)
        self._check_count(reals, imags)
        for i, g in enumerate(self.gridders):
            g.check_invert_args(vis, reals[i], imags[i] if imags else None)
        hs = (C.c_void_p * n)(*[g.h.value for g in self.gridders])
        vs = (C.c_void_p * n)(*([vis.data_ptr()] * n))
        mult = (C.c_double * n)(*([float(multiplication_factor)] * n))
        gr = C.c_void_p()
        g0 = self.gridders[0]
        g0._check(g0.L.hws_graph_capture_invert_batch(
            hs, n, vs, mult, self._arrays(reals), self._arrays(imags) if imags else None,
            C.byref(gr)))
        return Graph(g0, gr, (vis, list(reals), list(imags) if imags else None))

    def capture_predict(self, reals, imags, vis_outs):
        """One HIP graph that predicts every facet into its own output."""
        n = len(self.gridders)
        self._check_count(reals, imags, vis_outs)
        for i, g in enumerate(self.gridders):
            g.check_predict_args(reals[i], imags[i] if imags else None, vis_outs[i])
        hs = (C.c_void_p * n)(*[g.h.value for g in self.gridders])
        gr = C.c_void_p()
        g0 = self.gridders[0]
        g0._check(g0.L.hws_graph_capture_predict_batch(
            hs, n, self._arrays(reals), self._arrays(imags) if imags else None,
            self._arrays(vis_outs), C.byref(gr)))
        return Graph(g0, gr, (list(reals), list(imags) if imags else None, list(vis_outs)))

    def synchronize(self):
        for g in self.gridders:
            g.synchronize()

    def close(self):
        for g in self.gridders:
            g.close()
        self.gridders = []


__all__ = ["Facet", "FacetBatch", "create_facet_grid", "cut_facet", "facet_shift",
           "facets_of_grid", "gridded_size", "paste_facet"]
