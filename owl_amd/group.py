"""Ragged batches of forward convolutions in one persistent GPU launch.

``ConvGroup`` wraps the ``owl_amd_conv_group_*`` entry points: build a plan
for many conv2d / conv3d problems of different shapes, optionally autotune
per shape class a plan -- the tile edge and split-K factor in the
persistent launch, or a launch of the single-problem kernels of its own
(the halo kernel wins for large stride-1 layers) -- then run all of them
with one persistent kernel launch plus the split-K reductions and those own
launches, or capture that run in a hipGraph and replay it.  Each problem's
geometry is derived exactly like ``owl_amd.ndarray.dilated_conv2d``.
"""
import ctypes

import torch

from . import _lib
from .ndarray import _TORCH_DT, conv_problem
from .shapes import SAME


def _ptrs(ts):
    return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])


class ConvGroup:
    def __init__(self, problems, in_dtype=torch.bfloat16, out_dtype=torch.bfloat16):
        """problems: list of (x_shape, k_shape, stride, rate, padding)."""
        self.geoms = [conv_problem(xs, ks, st, rt, pad) for xs, ks, st, rt, pad in problems]
        self.in_dtype, self.out_dtype = in_dtype, out_dtype
        n = len(self.geoms)
        descs = (_lib.ConvDesc * max(n, 1))(*[g.desc() for g in self.geoms])
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().owl_amd_conv_group_create(ctypes.byref(h), n, descs, _TORCH_DT[in_dtype],
                                                        _TORCH_DT[out_dtype]))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            _lib.lib().owl_amd_conv_group_destroy(h)
            self._h = None

    def output_shapes(self):
        return [[g.B, *g.sp_out, g.Co] for g in self.geoms]

    def empty_outputs(self, device="cuda"):
        return [torch.empty(s, dtype=self.out_dtype, device=device) for s in self.output_shapes()]

    @staticmethod
    def _stream(stream):
        s = stream if stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def run(self, xs, ws, ys=None, stream=None):
        ys = ys if ys is not None else self.empty_outputs(xs[0].device)
        _lib.check(_lib.lib().owl_amd_conv_group_run(self._h, _ptrs(xs), _ptrs(ws), _ptrs(ys), self._stream(stream)))
        return ys

    def autotune(self, xs, ws, ys, stream=None):
        _lib.check(_lib.lib().owl_amd_conv_group_autotune(self._h, _ptrs(xs), _ptrs(ws), _ptrs(ys),
                                                          self._stream(stream)))

    def capture(self, xs, ws, ys, stream):
        _lib.check(_lib.lib().owl_amd_conv_group_capture(self._h, _ptrs(xs), _ptrs(ws), _ptrs(ys),
                                                         self._stream(stream)))

    def replay(self, stream):
        _lib.check(_lib.lib().owl_amd_conv_group_replay(self._h, self._stream(stream)))

    def plan(self):
        """(per-problem kind: 0 full tile, 1 half tile, 2 own launch;
        effective split-K count)."""
        n = len(self.geoms)
        k, s = (ctypes.c_int * max(n, 1))(), (ctypes.c_int * max(n, 1))()
        _lib.check(_lib.lib().owl_amd_conv_group_plan(self._h, k, s))
        return list(k)[:n], list(s)[:n]

    def set_plan(self, kinds, splits):
        """Impose per-problem kinds and requested split counts (1..16; 1
        with kind 2); drops a captured graph."""
        n = len(self.geoms)
        k = (ctypes.c_int * max(n, 1))(*kinds)
        s = (ctypes.c_int * max(n, 1))(*splits)
        _lib.check(_lib.lib().owl_amd_conv_group_set_plan(self._h, k, s))

    def tile_edges(self):
        """(per-problem tile edge in the persistent launch, 0 for a problem
        with its own launch; tiles of the persistent launch)."""
        n = len(self.geoms)
        e = (ctypes.c_int * max(n, 1))()
        t = ctypes.c_int64()
        _lib.check(_lib.lib().owl_amd_conv_group_info(self._h, e, ctypes.byref(t)))
        return list(e)[:n], t.value


def ragged_problems(seed, n=64):
    """n conv2d problems drawn as the ragged-batch workload specifies: batch
    U{1..32}, H = W in {7,14,28,56,112}, Ci, Co in {16,...,512}, kernel
    {1,3,5}, stride {1,2}, dilation {1,2}, SAME / VALID; draws whose VALID
    output would be empty are rejected."""
    import numpy as np
    from .shapes import VALID, conv_out_dim, k_eff
    rng = np.random.default_rng(seed)
    out = []
    while len(out) < n:
        B = int(rng.integers(1, 33))
        H = int(rng.choice([7, 14, 28, 56, 112]))
        ci, co = (int(rng.choice([16, 32, 64, 128, 256, 512])) for _ in range(2))
        k = int(rng.choice([1, 3, 5]))
        s = int(rng.choice([1, 2]))
        r = int(rng.choice([1, 2]))
        pad = [SAME, VALID][int(rng.integers(0, 2))]
        if conv_out_dim(pad, H, k_eff(k, r), s) <= 0:
            continue
        out.append(([B, H, H, ci], [k, k, ci, co], [s, s], [r, r], pad))
    return out
