"""A GPU engine for Owl's computation graphs (CGraph) on the convolution /
pooling path.

Owl's CGraph (``src/base/compute``) builds a symbolic graph whose nodes carry
an operator (owl_computation_type.ml) and a shape inferred when the node is
made (owl_computation_shape.ml), lets an optimiser rewrite the graph
(owl_computation_optimiser.ml), plans memory blocks -- a node may take over a
parent's block when the parent dies at that node and the operator allows it
(``split_parents``, owl_computation_cpu_init.ml) -- and evaluates every node
through the in-place ``*_ ~out`` ndarray wrappers
(owl_computation_cpu_eval.ml:180-281).  This module is that engine for the
operators on the path, on device tensors:

* construction: ``var_arr``, ``const_arr`` and the operators below, with
  Owl's names and argument order; shapes are inferred as each node is made;
* ``make_graph(input, output, name)``: the nodes in evaluation order and a
  block plan -- each result gets a device block, a block returns to the pool
  after its result's last consumer and a later result that fits takes it over
  (the smallest free block that holds it), ``relu`` and ``add`` write into their first parent's block
  when that parent dies there; one split-K workspace, sized for the largest
  backward-kernel node, is owned by the graph;
* ``optimise(graph)``: ``relu(add(conv2d(x, w), b))`` and ``add(conv2d(x, w),
  b)`` (a bias over the channels) become one node run by the fused conv
  kernel, bias and relu in its epilogue (in FP32 and FP64 this is the
  unfused result bit for bit; 16-bit results round once instead of twice);
* ``eval_graph(graph)``: evaluates on the current stream through the
  in-place ndarray wrappers, i.e. the library's HIP entry points (``relu``
  and ``add`` left unfused run as torch elementwise kernels);
* ``capture_graph`` / ``replay_graph``: one evaluation recorded into a HIP
  graph (after an eager warm-up evaluation) and replayed; ``assign_arr``
  copies into a variable's block, so new inputs reach the next replay.

Nothing is allocated while a graph evaluates: blocks, the workspace and the
32-bit bias copies of fused nodes are allocated once, on the first
evaluation.
"""
import numpy as np

from . import ndarray as nd
from .shapes import SAME, VALID, InvalidArgument, conv_out_dim, k_eff, transpose_out_dim

try:  # torch is plumbing for device memory, streams and graph capture
    import torch
except Exception:  # pragma: no cover
    torch = None

__all__ = [
    "var_arr", "const_arr", "conv2d", "conv3d", "dilated_conv2d", "transpose_conv2d",
    "conv2d_backward_input", "conv2d_backward_kernel", "conv3d_backward_input", "conv3d_backward_kernel",
    "max_pool2d", "avg_pool2d", "max_pool3d", "avg_pool3d", "max_pool2d_backward", "avg_pool2d_backward",
    "upsampling2d", "relu", "add", "make_graph", "optimise", "eval_graph", "assign_arr", "unpack_arr",
    "capture_graph", "replay_graph", "SAME", "VALID",
]


def _verify(cond, msg):
    if not cond:
        raise InvalidArgument(msg)


class Node:
    """A graph node: operator name, parents, operator parameters, and the
    inferred shape and element type of its result."""
    _count = 0

    def __init__(self, op, parents, params, shape, dtype, name=""):
        Node._count += 1
        self.id = Node._count
        self.op, self.parents, self.params = op, list(parents), dict(params)
        self.shape, self.dtype = [int(v) for v in shape], dtype
        self.name = name or f"{op}_{self.id}"
        self.value = None  # constants: the data; variables: set by assign_arr

    def nbytes(self):
        return int(np.prod(self.shape)) * torch.empty(0, dtype=self.dtype).element_size()

    def __repr__(self):
        return f"<{self.name}: {self.op} {self.shape}>"


# ----------------------------------------------------------------------------- construction
def var_arr(name, shape, dtype=None):
    """A variable (graph input) of the given shape (``var_arr ?shape name``,
    owl_computation_symbol_sig.ml:74); ``dtype`` a torch dtype, float32 by
    default."""
    return Node("var", [], {}, shape, dtype or torch.float32, name)


def const_arr(name, value):
    """A constant holding a device tensor (``const_arr name arr``,
    owl_computation_symbol_sig.ml:80)."""
    _verify(nd._is_torch(value) and value.is_cuda, "const_arr: the value must be a device tensor")
    n = Node("const", [], {}, value.shape, value.dtype, name)
    n.value = value.contiguous()
    return n


def _same_dtype(op, *ns):
    _verify(len({n.dtype for n in ns}) == 1, f"{op}: operands of different element types")
    return ns[0].dtype


def conv2d(input, kernel, stride, padding=SAME):
    """Conv2d node (owl_computation_cpu_eval.ml: ``A.conv2d_ ~out ~padding``)."""
    return _conv("conv2d", 2, input, kernel, stride, None, padding)


def conv3d(input, kernel, stride, padding=SAME):
    return _conv("conv3d", 3, input, kernel, stride, None, padding)


def dilated_conv2d(input, kernel, stride, rate, padding=SAME):
    return _conv("dilated_conv2d", 2, input, kernel, stride, rate, padding)


def _conv(op, ndim, x, k, stride, rate, padding):
    dt = _same_dtype(op, x, k)
    _verify(len(x.shape) == ndim + 2 and len(k.shape) == ndim + 2 and len(stride) == ndim,
            f"{op}: input / kernel / stride dimensions")
    _verify(x.shape[-1] == k.shape[-2], f"{op}: input channels {x.shape[-1]} != kernel's {k.shape[-2]}")
    r = list(rate) if rate is not None else [1] * ndim
    sp = [conv_out_dim(padding, n, k_eff(kk, rr), s) for n, kk, s, rr in zip(x.shape[1:-1], k.shape[:-2], stride, r)]
    _verify(all(v > 0 for v in sp), f"{op}: empty output shape {sp}")
    return Node(op, [x, k], {"stride": list(stride), "rate": rate, "padding": padding},
                [x.shape[0], *sp, k.shape[-1]], dt)


def transpose_conv2d(input, kernel, stride, padding=SAME):
    dt = _same_dtype("transpose_conv2d", input, kernel)
    sp = [transpose_out_dim(padding, n, kk, s) for n, kk, s in zip(input.shape[1:-1], kernel.shape[:-2], stride)]
    return Node("transpose_conv2d", [input, kernel], {"stride": list(stride), "padding": padding},
                [input.shape[0], *sp, kernel.shape[-1]], dt)


def conv2d_backward_input(input, kernel, stride, output_grad):
    return _conv_bwd("conv2d_backward_input", input, kernel, stride, output_grad, input.shape)


def conv2d_backward_kernel(input, kernel, stride, output_grad):
    return _conv_bwd("conv2d_backward_kernel", input, kernel, stride, output_grad, kernel.shape)


def conv3d_backward_input(input, kernel, stride, output_grad):
    return _conv_bwd("conv3d_backward_input", input, kernel, stride, output_grad, input.shape)


def conv3d_backward_kernel(input, kernel, stride, output_grad):
    return _conv_bwd("conv3d_backward_kernel", input, kernel, stride, output_grad, kernel.shape)


def _conv_bwd(op, x, k, stride, dy, shape):
    dt = _same_dtype(op, x, k, dy)
    return Node(op, [x, k, dy], {"stride": list(stride)}, shape, dt)


def _pool(op, ndim, x, kernel, stride, padding):
    _verify(len(x.shape) == ndim + 2 and len(kernel) == ndim and len(stride) == ndim, f"{op}: dimensions")
    sp = [conv_out_dim(padding, n, k, s) for n, k, s in zip(x.shape[1:-1], kernel, stride)]
    return Node(op, [x], {"kernel": list(kernel), "stride": list(stride), "padding": padding},
                [x.shape[0], *sp, x.shape[-1]], x.dtype)


def max_pool2d(input, kernel, stride, padding=SAME):
    return _pool("max_pool2d", 2, input, kernel, stride, padding)


def avg_pool2d(input, kernel, stride, padding=SAME):
    return _pool("avg_pool2d", 2, input, kernel, stride, padding)


def max_pool3d(input, kernel, stride, padding=SAME):
    return _pool("max_pool3d", 3, input, kernel, stride, padding)


def avg_pool3d(input, kernel, stride, padding=SAME):
    return _pool("avg_pool3d", 3, input, kernel, stride, padding)


def max_pool2d_backward(padding, input, kernel, stride, output_grad):
    return Node("max_pool2d_backward", [input, output_grad],
                {"kernel": list(kernel), "stride": list(stride), "padding": padding}, input.shape,
                _same_dtype("max_pool2d_backward", input, output_grad))


def avg_pool2d_backward(padding, input, kernel, stride, output_grad):
    return Node("avg_pool2d_backward", [input, output_grad],
                {"kernel": list(kernel), "stride": list(stride), "padding": padding}, input.shape,
                _same_dtype("avg_pool2d_backward", input, output_grad))


def upsampling2d(input, size):
    B, H, W, C = input.shape
    return Node("upsampling2d", [input], {"size": list(size)}, [B, H * size[0], W * size[1], C], input.dtype)


def relu(x):
    """Owl's relu, ``fmax(x, 0)`` (NaN -> 0)."""
    return Node("relu", [x], {}, x.shape, x.dtype)


def add(x, y):
    """x + y, y either x's shape or a vector over x's last axis (a bias)."""
    dt = _same_dtype("add", x, y)
    _verify(y.shape == x.shape or y.shape == [x.shape[-1]], f"add: shapes {x.shape} and {y.shape}")
    return Node("add", [x, y], {}, x.shape, dt)


# ----------------------------------------------------------------------------- graph
class Graph:
    def __init__(self, inputs, outputs, name):
        self.name, self.inputs, self.outputs = name, list(inputs), list(outputs)
        self.order = []
        self.plan = None          # node id -> block index
        self.block_bytes = []     # bytes of each block
        self.ws_bytes = 0
        self.blocks = None        # device tensors, on first evaluation
        self.views = None
        self.workspace = None
        self.hip_graph = None
        self._rebuild()

    def _rebuild(self):
        order, seen = [], set()

        def visit(n):
            if n.id in seen:
                return
            seen.add(n.id)
            for p in n.parents:
                visit(p)
            order.append(n)
        for o in self.outputs:
            visit(o)
        vars_ = [n for n in order if n.op == "var"]
        ids = {n.id for n in self.inputs}
        missing = [n.name for n in vars_ if n.id not in ids]
        _verify(not missing, f"make_graph: variables {missing} are not among the inputs")
        self.order = order
        self._plan()

    def _plan(self):
        """Assign blocks in evaluation order; see the module docstring."""
        last = {}
        for i, n in enumerate(self.order):
            for p in n.parents:
                last[p.id] = i
        keep = {n.id for n in self.outputs} | {n.id for n in self.order if n.op in ("var", "const")}
        plan, alive, sizes, free = {}, {}, [], []
        for i, n in enumerate(self.order):
            if n.op == "const":
                continue  # a constant keeps its own tensor
            nb = n.nbytes()
            p0 = n.parents[0] if n.parents else None
            if n.op == "var":
                # written before the run starts: a block of its own for the
                # whole evaluation, never one freed by an earlier node
                sizes.append(nb)
                alive[n.id] = len(sizes) - 1
            elif (n.op in ("relu", "add") and p0 is not None and p0.id in alive and p0.id not in keep
                    and last.get(p0.id) == i and p0.nbytes() == nb
                    and all(q.id != p0.id for q in n.parents[1:])):
                alive[n.id] = alive.pop(p0.id)  # overwrite the parent that dies here
            else:
                fit = [b for b in free if sizes[b] >= nb]
                if fit:  # best fit: the smallest free block that holds the result
                    b = min(fit, key=lambda j: (sizes[j], j))
                    free.remove(b)
                    alive[n.id] = b
                else:
                    sizes.append(nb)
                    alive[n.id] = len(sizes) - 1
            plan[n.id] = alive[n.id]
            # parents whose last consumer this is give their blocks back
            for p in {q.id: q for q in n.parents}.values():
                if last.get(p.id) == i and p.id not in keep and p.id in alive:
                    free.append(alive.pop(p.id))
        self.plan, self.block_bytes = plan, sizes
        ws = 0
        for n in self.order:
            if n.op in ("conv2d_backward_kernel", "conv3d_backward_kernel"):
                ndim = 2 if n.op.startswith("conv2d") else 3
                g = nd._bwd_geom(ndim, n.parents[0].shape, n.parents[1].shape, n.params["stride"], None,
                                 n.parents[2].shape)
                ws = max(ws, nd._ws_bytes(g, torch.empty(0, dtype=n.dtype)))
        self.ws_bytes = ws

    def planned_bytes(self):
        return sum(self.block_bytes) + self.ws_bytes

    def _materialise(self, device):
        if self.blocks is not None:
            return
        self.blocks = [torch.empty(b, dtype=torch.uint8, device=device) for b in self.block_bytes]
        self.views = {}
        for n in self.order:
            if n.op == "const":
                self.views[n.id] = n.value
            else:
                blk = self.blocks[self.plan[n.id]]
                self.views[n.id] = blk[: n.nbytes()].view(n.dtype).view(n.shape)
            if n.op in ("conv2d_bias_act", "conv2d_bias_act_maxpool") and n.parents[2] is not None:
                acc = torch.float64 if n.dtype == torch.float64 else torch.float32
                n.params["bias32"] = torch.empty(n.parents[2].shape, dtype=acc, device=device)
        self.workspace = torch.empty(max(self.ws_bytes, 1), dtype=torch.uint8, device=device)
        self.zeros = {n.dtype: torch.zeros((), dtype=n.dtype, device=device) for n in self.order}


def make_graph(input, output, name):
    """The graph of ``output``'s nodes, with ``input`` its variables
    (``make_graph ~input ~output name``, owl_computation_graph_sig.ml:39)."""
    return Graph(input, output, name)


def optimise(graph):
    """Fuse conv2d -> add bias -> relu (and conv2d -> add bias) into one node
    run by the fused conv kernel, and such a node followed by a 2x2 stride-2
    VALID max_pool2d into one that writes only the pooled layer.  A conv, add
    or fused layer whose result has another consumer, or is a graph output,
    is left alone."""
    consumers = {}
    for n in graph.order:
        for p in n.parents:
            consumers.setdefault(p.id, []).append(n)
    outs = {n.id for n in graph.outputs}

    def only_use(n, user):
        return n.id not in outs and len(consumers.get(n.id, [])) == 1 and consumers[n.id][0] is user

    changed = False
    for n in list(graph.order):
        if n.op == "relu" and n.parents[0].op == "add":
            a, act = n.parents[0], "relu"
        elif n.op == "add" and not (len(consumers.get(n.id, [])) == 1 and consumers[n.id][0].op == "relu"
                                    and only_use(n, consumers[n.id][0])):
            a, act = n, None
        else:
            continue
        c, b = a.parents
        if (c.op != "conv2d" or b.shape != [c.shape[-1]] or not only_use(c, a)
                or (act is not None and not only_use(a, n))):
            continue
        # n becomes the fused node in place, so its consumers keep their parent
        n.op, n.params = "conv2d_bias_act", {"stride": c.params["stride"], "padding": c.params["padding"],
                                             "act": act}
        n.parents = [c.parents[0], c.parents[1], b]
        changed = True
    if changed:
        graph._rebuild()
    consumers = {}
    for n in graph.order:
        for p in n.parents:
            consumers.setdefault(p.id, []).append(n)
    pooled = False
    for n in list(graph.order):
        c = n.parents[0] if n.parents else None
        if (n.op != "max_pool2d" or list(n.params["kernel"]) != [2, 2] or list(n.params["stride"]) != [2, 2]
                or n.params["padding"] != VALID or c.op != "conv2d_bias_act" or c.parents[2] is None
                or not only_use(c, n)):
            continue
        n.op, n.params = "conv2d_bias_act_maxpool", dict(c.params)
        n.parents = list(c.parents)
        pooled = True
    if pooled:
        graph._rebuild()
    return graph


def assign_arr(node, value):
    """Give a variable its value (``assign_arr``, owl_computation_symbol_sig.ml:212):
    kept as the node's value and copied into its block when the graph runs."""
    _verify(node.op == "var", "assign_arr: not a variable")
    _verify(list(value.shape) == node.shape and value.dtype == node.dtype,
            f"assign_arr: {node.name} expects {node.shape} {node.dtype}")
    node.value = value


def unpack_arr(graph, node):
    """A graph output's (or variable's) current value, a view of its block."""
    _verify(node in graph.outputs or node.op == "var", "unpack_arr: only outputs and variables keep their value")
    _verify(graph.views is not None, "unpack_arr: the graph has not been evaluated")
    return graph.views[node.id]


def _copy_vars(graph):
    for n in graph.order:
        if n.op == "var":
            _verify(n.value is not None, f"eval_graph: variable {n.name} has no value")
            out = graph.views[n.id]
            if n.value.data_ptr() != out.data_ptr():
                out.copy_(n.value)


def _run(graph):
    """Every operator node, in order, into its block.  Variables are read from
    their blocks (filled by _copy_vars), so a captured run has no copy from a
    caller's tensor in it."""
    v = graph.views
    for n in graph.order:
        op, P, out = n.op, n.params, v[n.id]
        a = [v[p.id] for p in n.parents]
        if op in ("var", "const"):
            pass
        elif op == "conv2d":
            nd.conv2d_(out, a[0], a[1], P["stride"], P["padding"])
        elif op == "conv3d":
            nd.conv3d_(out, a[0], a[1], P["stride"], P["padding"])
        elif op == "dilated_conv2d":
            nd.dilated_conv2d_(out, a[0], a[1], P["stride"], P["rate"], P["padding"])
        elif op == "transpose_conv2d":
            nd.transpose_conv2d_(out, a[0], a[1], P["stride"], P["padding"])
        elif op == "conv2d_backward_input":
            nd.conv2d_backward_input_(out, a[0], a[1], P["stride"], a[2])
        elif op == "conv3d_backward_input":
            nd.conv3d_backward_input_(out, a[0], a[1], P["stride"], a[2])
        elif op in ("conv2d_backward_kernel", "conv3d_backward_kernel"):
            ndim = 2 if op.startswith("conv2d") else 3
            variant = "spatial_backward_kernel" if ndim == 2 else "cuboid_backward_kernel_im2col"
            nd._conv_bwd(op, ndim, "kernel", a[0], a[1], P["stride"], None, a[2], variant, out=out,
                         workspace=graph.workspace)
        elif op in ("max_pool2d", "avg_pool2d", "max_pool3d", "avg_pool3d"):
            getattr(nd, op + "_")(out, a[0], P["kernel"], P["stride"], P["padding"])
        elif op in ("max_pool2d_backward", "avg_pool2d_backward"):
            getattr(nd, op + "_")(out, P["padding"], a[0], P["kernel"], P["stride"], a[1])
        elif op == "upsampling2d":
            nd.upsampling2d_(out, a[0], P["size"])
        elif op == "relu":
            torch.fmax(a[0], graph.zeros[n.dtype], out=out)
        elif op == "add":
            torch.add(a[0], a[1], out=out)
        elif op == "conv2d_bias_act":
            b32 = P["bias32"]
            b32.copy_(a[2])
            nd.conv2d_bias_act(a[0], a[1], b32, P["stride"], P["padding"], act=P["act"], out=out)
        elif op == "conv2d_bias_act_maxpool":
            b32 = P["bias32"]
            b32.copy_(a[2])
            nd.conv2d_bias_act_maxpool(a[0], a[1], b32, P["stride"], P["padding"], act=P["act"], keep_y=False,
                                       out=out)
        else:  # pragma: no cover
            raise InvalidArgument(f"eval_graph: unknown operator {op}")


def eval_graph(graph):
    """Evaluate the graph on the current stream (``eval_graph``,
    owl_computation_cpu_engine.ml:31).  Variables must have values."""
    dev = next((n.value.device for n in graph.order if n.op in ("var", "const") and n.value is not None), None)
    _verify(dev is not None, "eval_graph: no variable has a value")
    graph._materialise(dev)
    _copy_vars(graph)
    _run(graph)
    return [graph.views[o.id] for o in graph.outputs]


def capture_graph(graph):
    """Record one evaluation into a HIP graph: an eager evaluation on a side
    stream first (the library's per-device state and every block exist before
    capture), then the capture.  ``replay_graph`` runs it."""
    eval_graph(graph)
    main = torch.cuda.current_stream()
    side = torch.cuda.Stream()
    side.wait_stream(main)
    with torch.cuda.stream(side):
        _run(graph)
    main.wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _run(graph)
    graph.hip_graph = g
    return graph


def replay_graph(graph):
    """Replay the captured evaluation on the current stream.  Variables are
    read from their blocks: ``assign_arr`` values are copied in first."""
    _verify(graph.hip_graph is not None, "replay_graph: capture_graph first")
    _copy_vars(graph)
    graph.hip_graph.replay()
    return [graph.views[o.id] for o in graph.outputs]
