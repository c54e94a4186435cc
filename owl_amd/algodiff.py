"""Reverse-mode differentiation through the GPU conv / pool kernels.

Mirrors the part of Owl's Algodiff that this path feeds (src/base/algodiff):

* ``build_piso`` / ``build_siso`` (owl_algodiff_ops_builder.ml:267-330)
  register a primal function and its adjoint rule; the conv and pool ops
  below are built with them exactly as owl_algodiff_ops.ml:2328-2424 and
  :2754-2886 build ``NN.conv2d`` / ``NN.conv3d`` / ``NN.max_pool2d`` /
  ``NN.avg_pool3d``: the adjoint of a conv is
  ``(conv*_backward_input a b s ca, conv*_backward_kernel a b s ca)``.
* ``reverse_prop`` (owl_algodiff_reverse.ml:49-53) pushes adjoints from the
  output back to the leaves; ``grad`` returns the leaves' adjoints.

Values stay on the device (torch tensors); every primal and adjoint is a
call into libowl_amd.  Only reverse mode is provided: Owl's conv tangents
are not JVPs (owl_algodiff_ops.ml:2357-2361).
"""
import torch

from . import ndarray as nd
from .shapes import SAME


class Node:
    """A value on the tape (Owl's ``DR`` constructor): primal, parents and
    the closure that turns this node's adjoint into the parents' adjoints."""

    __slots__ = ("primal", "parents", "adjoint_fn", "adjoint", "name")

    def __init__(self, primal, parents=(), adjoint_fn=None, name="leaf"):
        self.primal = primal
        self.parents = parents
        self.adjoint_fn = adjoint_fn
        self.adjoint = None
        self.name = name


def make_reverse(x):
    return Node(x)


def primal(n):
    return n.primal if isinstance(n, Node) else n


def build_siso(name, ff, dr):
    """Single input, single output op: ``ff a`` forward, ``dr a cp ca`` adjoint."""
    def op(a):
        cp = ff(primal(a))
        out = Node(cp, (a,), None, name)
        out.adjoint_fn = lambda ca: (dr(primal(a), cp, ca),)
        return out
    return op


def build_piso(name, ff, dr_ab):
    """Pair input, single output op: ``ff a b`` forward, ``dr_ab a b cp ca`` ->
    (adjoint of a, adjoint of b)."""
    def op(a, b):
        cp = ff(primal(a), primal(b))
        out = Node(cp, (a, b), None, name)
        out.adjoint_fn = lambda ca: dr_ab(primal(a), primal(b), cp, ca)
        return out
    return op


def build_aiso(name, ff, dr):
    """Array input, single output op (owl_algodiff_ops_builder.ml:384-430):
    ``ff primals`` forward; ``dr idxs primals cp ca`` returns the adjoints of
    the inputs at ``idxs`` (those on the tape), in that order."""
    def op(*xs):
        ps = [primal(x) for x in xs]
        cp = ff(ps)
        idxs = [i for i, x in enumerate(xs) if isinstance(x, Node)]
        out = Node(cp, tuple(xs[i] for i in idxs), None, name)
        out.adjoint_fn = lambda ca: dr(idxs, ps, cp, ca)
        return out
    return op


def _topo(root):
    order, seen = [], set()

    def visit(n):
        if not isinstance(n, Node) or id(n) in seen:
            return
        seen.add(id(n))
        for p in n.parents:
            visit(p)
        order.append(n)
    visit(root)
    return order


def reverse_prop(seed, root):
    """Accumulate adjoints from ``root`` (seeded with ``seed``) to the leaves."""
    for n in _topo(root):
        n.adjoint = None
    root.adjoint = seed
    for n in reversed(_topo(root)):
        if n.adjoint is None or n.adjoint_fn is None:
            continue
        for p, g in zip(n.parents, n.adjoint_fn(n.adjoint)):
            if isinstance(p, Node) and g is not None:
                p.adjoint = g if p.adjoint is None else p.adjoint + g


def grad(f, *xs):
    """Adjoints of the scalar ``f(*xs)`` with respect to each of ``xs``."""
    leaves = [make_reverse(x) for x in xs]
    y = f(*leaves)
    reverse_prop(torch.ones_like(y.primal), y)
    return [l.adjoint for l in leaves]


# --------------------------------------------------------------------- NN ops
def conv2d(x, w, stride, padding=SAME):
    """owl_algodiff_ops.ml:2328-2375."""
    return build_piso(
        "conv2d",
        lambda a, b: nd.conv2d(a, b, stride, padding),
        # both gradients (dr_ab) in one paired call: conv2d_backward_input
        # and conv2d_backward_kernel, the reduction beside the first
        lambda a, b, cp, ca: nd.conv2d_backward(a, b, stride, ca, kernel_out_dtype=b.dtype),
    )(x, w)


def conv3d(x, w, stride, padding=SAME):
    """owl_algodiff_ops.ml:2378-2424."""
    return build_piso(
        "conv3d",
        lambda a, b: nd.conv3d(a, b, stride, padding),
        lambda a, b, cp, ca: nd.conv3d_backward(a, b, stride, ca, kernel_out_dtype=b.dtype),
    )(x, w)


def max_pool2d(x, kernel, stride, padding=SAME):
    """owl_algodiff_ops.ml:2754-2778."""
    return build_siso("max_pool2d", lambda a: nd.max_pool2d(a, kernel, stride, padding),
                      lambda a, cp, ca: nd.max_pool2d_backward(padding, a, kernel, stride, ca))(x)


def avg_pool3d(x, kernel, stride, padding=SAME):
    """owl_algodiff_ops.ml:2862-2886.  Pooling runs in the input's dtype when
    the library provides it (FP32 / FP64), else in FP32."""
    def ff(a):
        if a.dtype in (torch.float32, torch.float64):
            return nd.avg_pool3d(a, kernel, stride, padding)
        return nd.avg_pool3d(a.float(), kernel, stride, padding).to(a.dtype)

    def dr(a, cp, ca):
        if a.dtype in (torch.float32, torch.float64):
            return nd.avg_pool3d_backward(padding, a, kernel, stride, ca)
        return nd.avg_pool3d_backward(padding, a.float(), kernel, stride, ca.float()).to(a.dtype)
    return build_siso("avg_pool3d", ff, dr)(x)


def sum_sqr(x):
    """sum of squares (reduced in FP64 on the device); adjoint 2 x ca."""
    return build_siso("sum_sqr", lambda a: (a.double() * a.double()).sum(),
                      lambda a, cp, ca: (2.0 * a.double() * ca).to(a.dtype))(x)


# --------------------------------------------------------------------- checking
def finite_difference(loss, params, index, eps=1e-4):
    """Central difference d loss / d params[index[0]][index[1:]] with FP64
    parameters (Owl_algodiff_check.Reverse, owl_algodiff_check.ml:23-103,
    uses central differences with the same role)."""
    k, pos = index[0], tuple(index[1:])
    p = [q.clone() for q in params]
    p[k][pos] += eps
    up = float(loss(*p))
    p[k][pos] -= 2 * eps
    dn = float(loss(*p))
    return (up - dn) / (2 * eps)


def conv2d_bias_relu(x, w, b, stride, padding=SAME):
    """Owl's conv layer with a relu, ``relu (conv2d ~padding x w stride + b)``
    (owl_neural_neuron.ml:1253 and owl_algodiff_ops.ml:762-778), as one
    ``build_aiso`` op: the forward is the fused GPU conv (bias and relu in the
    epilogue, pre-activation kept for the adjoint); the adjoint is
    ``bias_act_backward`` then the two conv backward passes.  The bias is
    float32 (float64 for float64 layers) as the fused kernels take it."""
    saved = {}

    def ff(ps):
        a, k, bb = ps
        y, pre = nd.conv2d_bias_act(a, k, bb, stride, padding, act="relu", keep_pre=True)
        saved["pre"] = pre
        return y

    def dr(idxs, ps, cp, ca):
        a, k, bb = ps
        dz, db = nd.bias_act_backward(saved["pre"], ca, act="relu", bias_grad=2 in idxs)
        out = []
        for i in idxs:
            if i == 0:
                out.append(nd.conv2d_backward_input(a, k, stride, dz))
            elif i == 1:
                out.append(nd.conv2d_backward_kernel(a, k, stride, dz, out_dtype=k.dtype))
            else:
                out.append(db.to(bb.dtype))
        return out
    return build_aiso("conv2d_bias_relu", ff, dr)(x, w, b)


def conv2d_bias_relu_maxpool(x, w, b, stride, padding=SAME):
    """The conv layer with a relu followed by Owl's 2x2 stride-2 VALID
    max-pool layer (owl_neural_neuron.ml:1253, :2024), as one ``build_aiso``
    op: the forward pools in the fused conv's epilogue (y and the
    pre-activation kept for the adjoint); the adjoint is
    ``maxpool_bias_act_backward`` (max-pool backward, relu factor and bias
    gradient in one pass) then the two conv backward passes."""
    saved = {}

    def ff(ps):
        a, k, bb = ps
        y, pooled, pre = nd.conv2d_bias_act_maxpool(a, k, bb, stride, padding, act="relu", keep_pre=True)
        saved["y"], saved["pre"] = y, pre
        return pooled

    def dr(idxs, ps, cp, ca):
        a, k, bb = ps
        dz, db = nd.maxpool_bias_act_backward(saved["y"], saved["pre"], ca, bias_grad=2 in idxs)
        out = []
        for i in idxs:
            if i == 0:
                out.append(nd.conv2d_backward_input(a, k, stride, dz))
            elif i == 1:
                out.append(nd.conv2d_backward_kernel(a, k, stride, dz, out_dtype=k.dtype))
            else:
                out.append(db.to(bb.dtype))
        return out
    return build_aiso("conv2d_bias_relu_maxpool", ff, dr)(x, w, b)
