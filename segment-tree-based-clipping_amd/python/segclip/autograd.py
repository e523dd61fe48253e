"""torch.autograd binding of the batched intersection area / IoU (libsegclip).

    area, iou = clip_area_iou(base_xy, base_sizes, ov_xy, ov_sizes)

base_xy / ov_xy are float64 CUDA tensors of shape (n, 2) holding rings back
to back (no repeated closing vertex), base_sizes / ov_sizes their vertex
counts.  Paired mode (default): pair k clips base ring k against overlay ring
k.  One-to-many (paired=False): one base ring against every overlay.  The
forward runs the whole clip on the GPU and returns per-pair intersection area
and IoU; the backward is libsegclip's gradient kernel (combinatorics of the
forward held fixed).  recompute=True keeps nothing but the input rings
between forward and backward and clips again in backward.

Tensors cross the C ABI as raw device pointers.  The library runs on a torch
side stream of its own (one per device, default priority), ordered after the
caller's current stream on entry and before it on exit (stream waits, no
device-wide synchronisation); tensors used on both streams are recorded on
the other one so that torch's caching allocator does not reuse them early.
Inside forward and backward the library allocates from that allocator.  The
tuning settings persisted for a workload of the batch's shape
(segclip.tune.settings_for) are applied; tuning= overrides them.
"""
import numpy as np
import torch

from . import Clipper, torch_allocator
from . import tune

_side = {}


def _side_stream(dev):
    s = _side.get(dev.index)
    if s is None:
        s = _side[dev.index] = torch.cuda.Stream(dev)
    return s


class AreaIoU(torch.autograd.Function):
    @staticmethod
    def forward(ctx, base_xy, ov_xy, base_sizes, ov_sizes, paired, recompute, tuning):
        if base_xy.dtype != torch.float64 or ov_xy.dtype != torch.float64:
            raise TypeError("clip_area_iou: float64 coordinates required")
        if not (base_xy.is_cuda and ov_xy.is_cuda):
            raise ValueError("clip_area_iou: the rings must be CUDA tensors (the clip runs only on the GPU)")
        base_xy = base_xy.contiguous()
        ov_xy = ov_xy.contiguous()
        bsz = np.asarray(base_sizes, dtype=np.int32)
        csz = np.asarray(ov_sizes, dtype=np.int32)
        # the library's buffers come from torch's pool: the release after the
        # forward and the rebuild in the backward then cost no driver calls
        with torch_allocator():
            return AreaIoU._forward(ctx, base_xy, ov_xy, bsz, csz, paired, recompute, tuning)

    @staticmethod
    def _forward(ctx, base_xy, ov_xy, bsz, csz, paired, recompute, tuning):
        dev = base_xy.device
        cur = torch.cuda.current_stream(dev)
        side = _side_stream(dev)
        side.wait_stream(cur)
        base_xy.record_stream(side)
        ov_xy.record_stream(side)
        with torch.cuda.stream(side):
            cl = Clipper(device=dev.index or 0, stream=side.cuda_stream)
            if tuning is None:
                tuning = tune.settings_for(paired, int(max(bsz.max(), csz.max())))
            cl.apply_tuning(tuning)
            if paired:
                cl.set_pairs_device(base_xy.data_ptr(), bsz, ov_xy.data_ptr(), csz)
            else:
                if len(bsz) != 1:
                    raise ValueError("clip_area_iou: one base ring in one-to-many mode")
                cl.set_polygons_device(base_xy.data_ptr(), int(bsz[0]), ov_xy.data_ptr(), csz)
            cl.set_provenance(not recompute)
            cl.clip_batch()
            J = len(csz)
            out = torch.empty((4, J), dtype=torch.float64, device=dev)
            cl.batch_area_device(out.data_ptr())
            if recompute:
                cl.release()
        cur.wait_stream(side)
        out.record_stream(cur)
        ctx.cl = cl
        ctx.recompute = recompute
        ctx.shapes = (base_xy.shape, ov_xy.shape, dev)
        return out[0].clone(), out[3].clone()

    @staticmethod
    def backward(ctx, g_area, g_iou):
        with torch_allocator():
            return AreaIoU._backward(ctx, g_area, g_iou)

    @staticmethod
    def _backward(ctx, g_area, g_iou):
        cl = ctx.cl
        bshape, cshape, dev = ctx.shapes
        cur = torch.cuda.current_stream(dev)
        side = _side_stream(dev)
        ga = g_area.contiguous() if g_area is not None else None
        gi = g_iou.contiguous() if g_iou is not None else None
        side.wait_stream(cur)
        for t in (ga, gi):
            if t is not None:
                t.record_stream(side)
        with torch.cuda.stream(side):
            if ctx.recompute:
                cl.set_provenance(True)
                cl.clip_batch()
            gb = torch.empty(bshape, dtype=torch.float64, device=dev)
            gc = torch.empty(cshape, dtype=torch.float64, device=dev)
            cl.batch_area_backward_device(ga.data_ptr() if ga is not None else 0,
                                          gi.data_ptr() if gi is not None else 0, gb.data_ptr(), gc.data_ptr())
            if ctx.recompute:
                cl.release()
        cur.wait_stream(side)
        gb.record_stream(cur)
        gc.record_stream(cur)
        return gb, gc, None, None, None, None, None


def clip_area_iou(base_xy, base_sizes, ov_xy, ov_sizes, paired=True, recompute=False, tuning=None):
    """Per pair (or overlay): intersection area and IoU, differentiable w.r.t. the vertices."""
    return AreaIoU.apply(base_xy, ov_xy, list(base_sizes), list(ov_sizes), paired, recompute, tuning)
