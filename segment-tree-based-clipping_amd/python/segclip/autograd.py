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

Tensors cross the C ABI as raw device pointers; the library works on its own
stream, so both directions synchronise the device around the call.  Inside
forward and backward the library allocates from torch's caching allocator.
"""
import numpy as np
import torch

from . import Clipper, torch_allocator


def _sync():
    torch.cuda.synchronize()


class AreaIoU(torch.autograd.Function):
    @staticmethod
    def forward(ctx, base_xy, ov_xy, base_sizes, ov_sizes, paired, recompute):
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
            return AreaIoU._forward(ctx, base_xy, ov_xy, bsz, csz, paired, recompute)

    @staticmethod
    def _forward(ctx, base_xy, ov_xy, bsz, csz, paired, recompute):
        cl = Clipper(device=base_xy.device.index or 0)
        _sync()
        if paired:
            cl.set_pairs_device(base_xy.data_ptr(), bsz, ov_xy.data_ptr(), csz)
        else:
            if len(bsz) != 1:
                raise ValueError("clip_area_iou: one base ring in one-to-many mode")
            cl.set_polygons_device(base_xy.data_ptr(), int(bsz[0]), ov_xy.data_ptr(), csz)
        cl.set_provenance(not recompute)
        cl.clip_batch()
        J = len(csz)
        out = torch.empty((4, J), dtype=torch.float64, device=base_xy.device)
        cl.batch_area_device(out.data_ptr())
        if recompute:
            cl.release()
        ctx.cl = cl
        ctx.recompute = recompute
        ctx.shapes = (base_xy.shape, ov_xy.shape, base_xy.device)
        return out[0].clone(), out[3].clone()

    @staticmethod
    def backward(ctx, g_area, g_iou):
        with torch_allocator():
            return AreaIoU._backward(ctx, g_area, g_iou)

    @staticmethod
    def _backward(ctx, g_area, g_iou):
        cl = ctx.cl
        bshape, cshape, dev = ctx.shapes
        if ctx.recompute:
            cl.set_provenance(True)
            cl.clip_batch()
        ga = g_area.contiguous() if g_area is not None else None
        gi = g_iou.contiguous() if g_iou is not None else None
        gb = torch.empty(bshape, dtype=torch.float64, device=dev)
        gc = torch.empty(cshape, dtype=torch.float64, device=dev)
        _sync()
        cl.batch_area_backward_device(ga.data_ptr() if ga is not None else 0,
                                      gi.data_ptr() if gi is not None else 0, gb.data_ptr(), gc.data_ptr())
        if ctx.recompute:
            cl.release()
        return gb, gc, None, None, None, None


def clip_area_iou(base_xy, base_sizes, ov_xy, ov_sizes, paired=True, recompute=False):
    """Per pair (or overlay): intersection area and IoU, differentiable w.r.t. the vertices."""
    return AreaIoU.apply(base_xy, ov_xy, list(base_sizes), list(ov_sizes), paired, recompute)
