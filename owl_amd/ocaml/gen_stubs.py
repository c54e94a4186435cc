"""Generate owl_amd_stubs.c: the OCaml C stubs that Owl's externals bind
(``external ... = "stub_<kind>_ndarray_<op>_bytecode" "..._native"``,
src/owl/core/owl_ndarray_conv.ml, owl_ndarray_pool.ml,
owl_ndarray_upsampling.ml), each forwarding to the libowl_amd drop-in of the
same name.

Every stub is the same mechanical shape, written out from one table:
  * a Bigarray ``value`` becomes its data pointer (``Caml_ba_data_val``), an
    OCaml ``int`` is untagged (``Long_val``) to int64_t;
  * all ``value``s are read before the runtime lock is released
    (``caml_enter_blocking_section``): the library touches only Bigarray data,
    which lives outside the OCaml heap, so other OCaml threads run meanwhile;
  * a non-zero library status raises ``Failure`` with the library's message
    (``caml_failwith``) after the lock is taken back -- where Owl's C exits on
    an allocation failure (owl_ndarray_conv_impl.h:299);
  * the ``_bytecode`` form forwards ``argv[0 .. n-1]`` to the ``_native`` one
    (Owl's own stubs do the same, e.g. owl_ndarray_conv_impl.h:456-462).

Complex max pooling keeps Owl's CPU stubs (``OWL_KEEP``): the library reports
it unsupported (Owl's complex max starts from -inf + 0i, owl_core.h:125, and
is degenerate).  Run ``python -m owl_amd.ocaml.gen_stubs`` after changing the
table; ``tests/test_ocaml_stubs.py`` checks the committed file against it and
against the reference's externals.
"""
import os

KINDS = [("float32", "float"), ("float64", "double"), ("complex32", "owl_amd_complex32"),
         ("complex64", "owl_amd_complex64")]

# argument families: leading Bigarrays (element type: T, or int64_t for "i64"),
# then OCaml ints; header macros in include/owl_amd.h name the same lists
FAMILIES = {
    "CONV2D_FWD": (["T", "T", "T"], 14),
    "CONV2D_BWD": (["T", "T", "T"], 13),
    "CONV3D_FWD": (["T", "T", "T"], 16),
    "CONV3D_BWD": (["T", "T", "T"], 15),
    "DCONV3D_FWD": (["T", "T", "T"], 19),
    "DCONV3D_BWD": (["T", "T", "T"], 18),
    "POOL2D_FWD": (["T", "T"], 13),
    "POOL2D_BWD": (["T", "T", "T"], 12),
    "POOL2D_ARG": (["T", "T", "i64"], 12),
    "POOL3D_FWD": (["T", "T"], 15),
    "POOL3D_BWD": (["T", "T", "T"], 15),
    "UPSAMPLE": (["T", "T"], 8),
}

# (op, family, reference definition)
OPS = [
    ("conv_spatial", "CONV2D_FWD", "owl_ndarray_conv_impl.h:246"),
    ("conv_spatial_backward_input", "CONV2D_BWD", "owl_ndarray_conv_impl.h:465"),
    ("conv_spatial_backward_kernel", "CONV2D_BWD", "owl_ndarray_conv_impl.h:675"),
    ("conv_spatial_im2col", "CONV2D_FWD", "owl_ndarray_conv_impl.h:903"),
    ("conv_spatial_backward_kernel_im2col", "CONV2D_BWD", "owl_ndarray_conv_impl.h:1006"),
    ("conv_spatial_backward_input_im2col", "CONV2D_BWD", "owl_ndarray_conv_impl.h:1116"),
    ("conv_cuboid_im2col", "CONV3D_FWD", "owl_ndarray_conv_impl.h:1211"),
    ("conv_cuboid_backward_kernel_im2col", "CONV3D_BWD", "owl_ndarray_conv_impl.h:1331"),
    ("conv_cuboid_backward_input_im2col", "CONV3D_BWD", "owl_ndarray_conv_impl.h:1456"),
    ("conv_spatial_mec", "CONV2D_FWD", "owl_ndarray_conv_impl.h:1570"),
    ("conv_spatial_backward_kernel_mec", "CONV2D_BWD", "owl_ndarray_conv_impl.h:1699"),
    ("conv_spatial_backward_input_mec", "CONV2D_BWD", "owl_ndarray_conv_impl.h:1828"),
    ("conv_cuboid_mec", "CONV3D_FWD", "owl_ndarray_conv_impl.h:1957"),
    ("conv_cuboid_backward_kernel_mec", "CONV3D_BWD", "owl_ndarray_conv_impl.h:2118"),
    ("conv_cuboid_backward_input_mec", "CONV3D_BWD", "owl_ndarray_conv_impl.h:2278"),
    ("conv_spatial_naive", "CONV2D_FWD", "owl_ndarray_conv_impl.h:2440"),
    ("conv_spatial_backward_kernel_naive", "CONV2D_BWD", "owl_ndarray_conv_impl.h:2546"),
    ("conv_spatial_backward_input_naive", "CONV2D_BWD", "owl_ndarray_conv_impl.h:2642"),
    ("conv_cuboid_naive", "CONV3D_FWD", "owl_ndarray_conv_impl.h:2735"),
    ("conv_cuboid_backward_kernel_naive", "CONV3D_BWD", "owl_ndarray_conv_impl.h:2860"),
    ("conv_cuboid_backward_input_naive", "CONV3D_BWD", "owl_ndarray_conv_impl.h:2977"),
    ("conv_dilated_spatial_im2col", "CONV2D_FWD", "owl_ndarray_conv_impl.h:3097"),
    ("conv_dilated_spatial_backward_kernel_im2col", "CONV2D_BWD", "owl_ndarray_conv_impl.h:3203"),
    ("conv_dilated_spatial_backward_input_im2col", "CONV2D_BWD", "owl_ndarray_conv_impl.h:3317"),
    ("conv_dilated_cuboid_im2col", "DCONV3D_FWD", "owl_ndarray_conv_impl.h:3417"),
    ("conv_dilated_cuboid_backward_kernel_im2col", "DCONV3D_BWD", "owl_ndarray_conv_impl.h:3545"),
    ("conv_dilated_cuboid_backward_input_im2col", "DCONV3D_BWD", "owl_ndarray_conv_impl.h:3678"),
    ("maxpool_spatial", "POOL2D_FWD", "owl_ndarray_pool_impl.h:9"),
    ("avgpool_spatial", "POOL2D_FWD", "owl_ndarray_pool_impl.h:9"),
    ("maxpool_spatial_backward", "POOL2D_BWD", "owl_ndarray_pool_impl.h:102"),
    ("avgpool_spatial_backward", "POOL2D_BWD", "owl_ndarray_pool_impl.h:102"),
    ("maxpool_cuboid", "POOL3D_FWD", "owl_ndarray_pool_impl.h:212"),
    ("avgpool_cuboid", "POOL3D_FWD", "owl_ndarray_pool_impl.h:212"),
    ("maxpool_cuboid_backward", "POOL3D_BWD", "owl_ndarray_pool_impl.h:324"),
    ("avgpool_cuboid_backward", "POOL3D_BWD", "owl_ndarray_pool_impl.h:324"),
    ("maxpool_spatial_arg", "POOL2D_ARG", "owl_ndarray_pool_impl.h:459"),
    ("upsampling_spatial_backward", "UPSAMPLE", "owl_ndarray_upsampling_impl.h:9"),
]

# (kind, op) pairs whose Owl CPU stubs stay in place
OWL_KEEP = {(k, op) for k in ("complex32", "complex64")
            for op in ("maxpool_spatial", "maxpool_spatial_backward", "maxpool_cuboid", "maxpool_cuboid_backward",
                       "maxpool_spatial_arg")}

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "owl_amd_stubs.c")

HEADER = """\
/* Generated by owl_amd/ocaml/gen_stubs.py -- edit the table there, not this file.
 *
 * OCaml C stubs for Owl's convolution / pooling / upsampling externals
 * (src/owl/core/owl_ndarray_conv.ml, owl_ndarray_pool.ml,
 * owl_ndarray_upsampling.ml), forwarding to libowl_amd's drop-ins
 * (include/owl_amd.h).  Compiled with Owl's stubs in place of the
 * instantiations in owl_ndarray_conv_stub.c, owl_ndarray_pool_stub.c and
 * owl_ndarray_upsampling_stub.c (complex max pooling excepted: its Owl stubs
 * stay).  A Bigarray value becomes its data pointer and an OCaml int an
 * int64_t, read before the runtime lock is released; a non-zero status
 * raises Failure with the library's message. */
#include <stdint.h>

#include <caml/bigarray.h>
#include <caml/fail.h>
#include <caml/mlvalues.h>
#include <caml/signals.h>

#include "owl_amd.h"

static void owl_amd_stub_check(int rc) {
  if (rc != OWL_AMD_OK) caml_failwith(owl_amd_last_error());
}
"""


def stub(kind, ctype, op, fam, where):
    arrs, nint = FAMILIES[fam]
    n = len(arrs) + nint
    name = f"stub_{kind}_ndarray_{op}"
    lines = [f"/* {where} */",
             f"CAMLprim value {name}_native(" + ", ".join(f"value v{i}" for i in range(n)) + ") {"]
    args = []
    for i, t in enumerate(arrs):
        ct = "int64_t" if t == "i64" else ctype
        lines.append(f"  {ct} *a{i} = ({ct} *)Caml_ba_data_val(v{i});")
        args.append(f"a{i}")
    for i in range(len(arrs), n):
        lines.append(f"  int64_t i{i} = (int64_t)Long_val(v{i});")
        args.append(f"i{i}")
    lines += ["  int rc;",
              "  caml_enter_blocking_section();",
              f"  rc = owl_amd_{kind}_ndarray_{op}(" + ", ".join(args) + ");",
              "  caml_leave_blocking_section();",
              "  owl_amd_stub_check(rc);",
              "  return Val_unit;",
              "}",
              "",
              f"CAMLprim value {name}_bytecode(value *argv, int argn) {{",
              "  (void)argn;",
              f"  return {name}_native(" + ", ".join(f"argv[{i}]" for i in range(n)) + ");",
              "}",
              ""]
    return "\n".join(lines)


def generate():
    parts = [HEADER]
    for kind, ctype in KINDS:
        parts.append(f"/* ---------------------------------------------------------------- {kind} */\n")
        for op, fam, where in OPS:
            if (kind, op) in OWL_KEEP:
                continue
            parts.append(stub(kind, ctype, op, fam, where))
    return "\n".join(parts)


def stubs():
    """[(native name, bytecode name, library entry point, arg kinds)] of the generated file."""
    rows = []
    for kind, _ in KINDS:
        for op, fam, _ in OPS:
            if (kind, op) in OWL_KEEP:
                continue
            arrs, nint = FAMILIES[fam]
            rows.append((f"stub_{kind}_ndarray_{op}_native", f"stub_{kind}_ndarray_{op}_bytecode",
                         f"owl_amd_{kind}_ndarray_{op}", ["arr"] * len(arrs) + ["int"] * nint,
                         ["i64" if t == "i64" else "T" for t in arrs]))
    return rows


if __name__ == "__main__":
    with open(OUT, "w") as f:
        f.write(generate())
    print(f"wrote {OUT}")
