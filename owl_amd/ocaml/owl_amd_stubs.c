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

/* ---------------------------------------------------------------- float32 */

/* owl_ndarray_conv_impl.h:246 */
CAMLprim value stub_float32_ndarray_conv_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:465 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_input(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_input_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:675 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_kernel(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_kernel_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:903 */
CAMLprim value stub_float32_ndarray_conv_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1006 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1116 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1211 */
CAMLprim value stub_float32_ndarray_conv_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:1331 */
CAMLprim value stub_float32_ndarray_conv_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1456 */
CAMLprim value stub_float32_ndarray_conv_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1570 */
CAMLprim value stub_float32_ndarray_conv_spatial_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1699 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1828 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1957 */
CAMLprim value stub_float32_ndarray_conv_cuboid_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2118 */
CAMLprim value stub_float32_ndarray_conv_cuboid_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2278 */
CAMLprim value stub_float32_ndarray_conv_cuboid_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2440 */
CAMLprim value stub_float32_ndarray_conv_spatial_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:2546 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2642 */
CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_spatial_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_spatial_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_spatial_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2735 */
CAMLprim value stub_float32_ndarray_conv_cuboid_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2860 */
CAMLprim value stub_float32_ndarray_conv_cuboid_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2977 */
CAMLprim value stub_float32_ndarray_conv_cuboid_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_cuboid_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_cuboid_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_cuboid_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:3097 */
CAMLprim value stub_float32_ndarray_conv_dilated_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_dilated_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_dilated_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_dilated_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:3203 */
CAMLprim value stub_float32_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_dilated_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_dilated_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3317 */
CAMLprim value stub_float32_ndarray_conv_dilated_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_dilated_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_dilated_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_dilated_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3417 */
CAMLprim value stub_float32_ndarray_conv_dilated_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20, value v21) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int64_t i21 = (int64_t)Long_val(v21);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_dilated_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20, i21);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_dilated_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_dilated_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20], argv[21]);
}

/* owl_ndarray_conv_impl.h:3545 */
CAMLprim value stub_float32_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_dilated_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_dilated_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_conv_impl.h:3678 */
CAMLprim value stub_float32_ndarray_conv_dilated_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_conv_dilated_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_conv_dilated_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_conv_dilated_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_pool_impl.h:9 */
CAMLprim value stub_float32_ndarray_maxpool_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_maxpool_spatial(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_maxpool_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_maxpool_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:9 */
CAMLprim value stub_float32_ndarray_avgpool_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_avgpool_spatial(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_avgpool_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_avgpool_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:102 */
CAMLprim value stub_float32_ndarray_maxpool_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_maxpool_spatial_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_maxpool_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_maxpool_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:102 */
CAMLprim value stub_float32_ndarray_avgpool_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_avgpool_spatial_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_avgpool_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_avgpool_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:212 */
CAMLprim value stub_float32_ndarray_maxpool_cuboid_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_maxpool_cuboid(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_maxpool_cuboid_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_maxpool_cuboid_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_pool_impl.h:212 */
CAMLprim value stub_float32_ndarray_avgpool_cuboid_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_avgpool_cuboid(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_avgpool_cuboid_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_avgpool_cuboid_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_pool_impl.h:324 */
CAMLprim value stub_float32_ndarray_maxpool_cuboid_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_maxpool_cuboid_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_maxpool_cuboid_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_maxpool_cuboid_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_pool_impl.h:324 */
CAMLprim value stub_float32_ndarray_avgpool_cuboid_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  float *a2 = (float *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_avgpool_cuboid_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_avgpool_cuboid_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_avgpool_cuboid_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_pool_impl.h:459 */
CAMLprim value stub_float32_ndarray_maxpool_spatial_arg_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  int64_t *a2 = (int64_t *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_maxpool_spatial_arg(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_maxpool_spatial_arg_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_maxpool_spatial_arg_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_upsampling_impl.h:9 */
CAMLprim value stub_float32_ndarray_upsampling_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9) {
  float *a0 = (float *)Caml_ba_data_val(v0);
  float *a1 = (float *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float32_ndarray_upsampling_spatial_backward(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float32_ndarray_upsampling_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float32_ndarray_upsampling_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9]);
}

/* ---------------------------------------------------------------- float64 */

/* owl_ndarray_conv_impl.h:246 */
CAMLprim value stub_float64_ndarray_conv_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:465 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_input(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_input_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:675 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_kernel(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_kernel_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:903 */
CAMLprim value stub_float64_ndarray_conv_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1006 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1116 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1211 */
CAMLprim value stub_float64_ndarray_conv_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:1331 */
CAMLprim value stub_float64_ndarray_conv_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1456 */
CAMLprim value stub_float64_ndarray_conv_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1570 */
CAMLprim value stub_float64_ndarray_conv_spatial_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1699 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1828 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1957 */
CAMLprim value stub_float64_ndarray_conv_cuboid_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2118 */
CAMLprim value stub_float64_ndarray_conv_cuboid_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2278 */
CAMLprim value stub_float64_ndarray_conv_cuboid_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2440 */
CAMLprim value stub_float64_ndarray_conv_spatial_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:2546 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2642 */
CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_spatial_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_spatial_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_spatial_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2735 */
CAMLprim value stub_float64_ndarray_conv_cuboid_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2860 */
CAMLprim value stub_float64_ndarray_conv_cuboid_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2977 */
CAMLprim value stub_float64_ndarray_conv_cuboid_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_cuboid_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_cuboid_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_cuboid_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:3097 */
CAMLprim value stub_float64_ndarray_conv_dilated_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_dilated_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_dilated_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_dilated_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:3203 */
CAMLprim value stub_float64_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_dilated_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_dilated_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3317 */
CAMLprim value stub_float64_ndarray_conv_dilated_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_dilated_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_dilated_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_dilated_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3417 */
CAMLprim value stub_float64_ndarray_conv_dilated_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20, value v21) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int64_t i21 = (int64_t)Long_val(v21);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_dilated_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20, i21);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_dilated_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_dilated_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20], argv[21]);
}

/* owl_ndarray_conv_impl.h:3545 */
CAMLprim value stub_float64_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_dilated_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_dilated_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_conv_impl.h:3678 */
CAMLprim value stub_float64_ndarray_conv_dilated_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_conv_dilated_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_conv_dilated_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_conv_dilated_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_pool_impl.h:9 */
CAMLprim value stub_float64_ndarray_maxpool_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_maxpool_spatial(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_maxpool_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_maxpool_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:9 */
CAMLprim value stub_float64_ndarray_avgpool_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_avgpool_spatial(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_avgpool_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_avgpool_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:102 */
CAMLprim value stub_float64_ndarray_maxpool_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_maxpool_spatial_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_maxpool_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_maxpool_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:102 */
CAMLprim value stub_float64_ndarray_avgpool_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_avgpool_spatial_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_avgpool_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_avgpool_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:212 */
CAMLprim value stub_float64_ndarray_maxpool_cuboid_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_maxpool_cuboid(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_maxpool_cuboid_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_maxpool_cuboid_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_pool_impl.h:212 */
CAMLprim value stub_float64_ndarray_avgpool_cuboid_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_avgpool_cuboid(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_avgpool_cuboid_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_avgpool_cuboid_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_pool_impl.h:324 */
CAMLprim value stub_float64_ndarray_maxpool_cuboid_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_maxpool_cuboid_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_maxpool_cuboid_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_maxpool_cuboid_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_pool_impl.h:324 */
CAMLprim value stub_float64_ndarray_avgpool_cuboid_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  double *a2 = (double *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_avgpool_cuboid_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_avgpool_cuboid_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_avgpool_cuboid_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_pool_impl.h:459 */
CAMLprim value stub_float64_ndarray_maxpool_spatial_arg_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  int64_t *a2 = (int64_t *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_maxpool_spatial_arg(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_maxpool_spatial_arg_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_maxpool_spatial_arg_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_upsampling_impl.h:9 */
CAMLprim value stub_float64_ndarray_upsampling_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9) {
  double *a0 = (double *)Caml_ba_data_val(v0);
  double *a1 = (double *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_float64_ndarray_upsampling_spatial_backward(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_float64_ndarray_upsampling_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_float64_ndarray_upsampling_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9]);
}

/* ---------------------------------------------------------------- complex32 */

/* owl_ndarray_conv_impl.h:246 */
CAMLprim value stub_complex32_ndarray_conv_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:465 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_input(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_input_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:675 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_kernel(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_kernel_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:903 */
CAMLprim value stub_complex32_ndarray_conv_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1006 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1116 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1211 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:1331 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1456 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1570 */
CAMLprim value stub_complex32_ndarray_conv_spatial_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1699 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1828 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1957 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2118 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2278 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2440 */
CAMLprim value stub_complex32_ndarray_conv_spatial_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:2546 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2642 */
CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_spatial_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_spatial_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_spatial_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2735 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2860 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2977 */
CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_cuboid_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_cuboid_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_cuboid_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:3097 */
CAMLprim value stub_complex32_ndarray_conv_dilated_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_dilated_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_dilated_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_dilated_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:3203 */
CAMLprim value stub_complex32_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_dilated_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_dilated_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3317 */
CAMLprim value stub_complex32_ndarray_conv_dilated_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_dilated_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_dilated_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_dilated_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3417 */
CAMLprim value stub_complex32_ndarray_conv_dilated_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20, value v21) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int64_t i21 = (int64_t)Long_val(v21);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_dilated_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20, i21);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_dilated_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_dilated_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20], argv[21]);
}

/* owl_ndarray_conv_impl.h:3545 */
CAMLprim value stub_complex32_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_dilated_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_dilated_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_conv_impl.h:3678 */
CAMLprim value stub_complex32_ndarray_conv_dilated_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_conv_dilated_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_conv_dilated_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_conv_dilated_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_pool_impl.h:9 */
CAMLprim value stub_complex32_ndarray_avgpool_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_avgpool_spatial(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_avgpool_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_avgpool_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:102 */
CAMLprim value stub_complex32_ndarray_avgpool_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_avgpool_spatial_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_avgpool_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_avgpool_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:212 */
CAMLprim value stub_complex32_ndarray_avgpool_cuboid_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_avgpool_cuboid(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_avgpool_cuboid_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_avgpool_cuboid_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_pool_impl.h:324 */
CAMLprim value stub_complex32_ndarray_avgpool_cuboid_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  owl_amd_complex32 *a2 = (owl_amd_complex32 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_avgpool_cuboid_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_avgpool_cuboid_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_avgpool_cuboid_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_upsampling_impl.h:9 */
CAMLprim value stub_complex32_ndarray_upsampling_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9) {
  owl_amd_complex32 *a0 = (owl_amd_complex32 *)Caml_ba_data_val(v0);
  owl_amd_complex32 *a1 = (owl_amd_complex32 *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex32_ndarray_upsampling_spatial_backward(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex32_ndarray_upsampling_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex32_ndarray_upsampling_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9]);
}

/* ---------------------------------------------------------------- complex64 */

/* owl_ndarray_conv_impl.h:246 */
CAMLprim value stub_complex64_ndarray_conv_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:465 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_input(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_input_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:675 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_kernel(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_kernel_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:903 */
CAMLprim value stub_complex64_ndarray_conv_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1006 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1116 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1211 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:1331 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1456 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:1570 */
CAMLprim value stub_complex64_ndarray_conv_spatial_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:1699 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1828 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:1957 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2118 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_kernel_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_backward_kernel_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_kernel_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_backward_kernel_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2278 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_input_mec_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_backward_input_mec(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_input_mec_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_backward_input_mec_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2440 */
CAMLprim value stub_complex64_ndarray_conv_spatial_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:2546 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2642 */
CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_spatial_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_spatial_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_spatial_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:2735 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18]);
}

/* owl_ndarray_conv_impl.h:2860 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_kernel_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_backward_kernel_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_kernel_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_backward_kernel_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:2977 */
CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_input_naive_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_cuboid_backward_input_naive(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_cuboid_backward_input_naive_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_cuboid_backward_input_naive_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_conv_impl.h:3097 */
CAMLprim value stub_complex64_ndarray_conv_dilated_spatial_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_dilated_spatial_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_dilated_spatial_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_dilated_spatial_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_conv_impl.h:3203 */
CAMLprim value stub_complex64_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_dilated_spatial_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_dilated_spatial_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_dilated_spatial_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3317 */
CAMLprim value stub_complex64_ndarray_conv_dilated_spatial_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_dilated_spatial_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_dilated_spatial_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_dilated_spatial_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15]);
}

/* owl_ndarray_conv_impl.h:3417 */
CAMLprim value stub_complex64_ndarray_conv_dilated_cuboid_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20, value v21) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int64_t i21 = (int64_t)Long_val(v21);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_dilated_cuboid_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20, i21);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_dilated_cuboid_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_dilated_cuboid_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20], argv[21]);
}

/* owl_ndarray_conv_impl.h:3545 */
CAMLprim value stub_complex64_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_dilated_cuboid_backward_kernel_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_dilated_cuboid_backward_kernel_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_dilated_cuboid_backward_kernel_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_conv_impl.h:3678 */
CAMLprim value stub_complex64_ndarray_conv_dilated_cuboid_backward_input_im2col_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17, value v18, value v19, value v20) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int64_t i18 = (int64_t)Long_val(v18);
  int64_t i19 = (int64_t)Long_val(v19);
  int64_t i20 = (int64_t)Long_val(v20);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_conv_dilated_cuboid_backward_input_im2col(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17, i18, i19, i20);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_conv_dilated_cuboid_backward_input_im2col_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_conv_dilated_cuboid_backward_input_im2col_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17], argv[18], argv[19], argv[20]);
}

/* owl_ndarray_pool_impl.h:9 */
CAMLprim value stub_complex64_ndarray_avgpool_spatial_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_avgpool_spatial(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_avgpool_spatial_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_avgpool_spatial_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:102 */
CAMLprim value stub_complex64_ndarray_avgpool_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_avgpool_spatial_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_avgpool_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_avgpool_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14]);
}

/* owl_ndarray_pool_impl.h:212 */
CAMLprim value stub_complex64_ndarray_avgpool_cuboid_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_avgpool_cuboid(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_avgpool_cuboid_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_avgpool_cuboid_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16]);
}

/* owl_ndarray_pool_impl.h:324 */
CAMLprim value stub_complex64_ndarray_avgpool_cuboid_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9, value v10, value v11, value v12, value v13, value v14, value v15, value v16, value v17) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  owl_amd_complex64 *a2 = (owl_amd_complex64 *)Caml_ba_data_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int64_t i10 = (int64_t)Long_val(v10);
  int64_t i11 = (int64_t)Long_val(v11);
  int64_t i12 = (int64_t)Long_val(v12);
  int64_t i13 = (int64_t)Long_val(v13);
  int64_t i14 = (int64_t)Long_val(v14);
  int64_t i15 = (int64_t)Long_val(v15);
  int64_t i16 = (int64_t)Long_val(v16);
  int64_t i17 = (int64_t)Long_val(v17);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_avgpool_cuboid_backward(a0, a1, a2, i3, i4, i5, i6, i7, i8, i9, i10, i11, i12, i13, i14, i15, i16, i17);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_avgpool_cuboid_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_avgpool_cuboid_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9], argv[10], argv[11], argv[12], argv[13], argv[14], argv[15], argv[16], argv[17]);
}

/* owl_ndarray_upsampling_impl.h:9 */
CAMLprim value stub_complex64_ndarray_upsampling_spatial_backward_native(value v0, value v1, value v2, value v3, value v4, value v5, value v6, value v7, value v8, value v9) {
  owl_amd_complex64 *a0 = (owl_amd_complex64 *)Caml_ba_data_val(v0);
  owl_amd_complex64 *a1 = (owl_amd_complex64 *)Caml_ba_data_val(v1);
  int64_t i2 = (int64_t)Long_val(v2);
  int64_t i3 = (int64_t)Long_val(v3);
  int64_t i4 = (int64_t)Long_val(v4);
  int64_t i5 = (int64_t)Long_val(v5);
  int64_t i6 = (int64_t)Long_val(v6);
  int64_t i7 = (int64_t)Long_val(v7);
  int64_t i8 = (int64_t)Long_val(v8);
  int64_t i9 = (int64_t)Long_val(v9);
  int rc;
  caml_enter_blocking_section();
  rc = owl_amd_complex64_ndarray_upsampling_spatial_backward(a0, a1, i2, i3, i4, i5, i6, i7, i8, i9);
  caml_leave_blocking_section();
  owl_amd_stub_check(rc);
  return Val_unit;
}

CAMLprim value stub_complex64_ndarray_upsampling_spatial_backward_bytecode(value *argv, int argn) {
  (void)argn;
  return stub_complex64_ndarray_upsampling_spatial_backward_native(argv[0], argv[1], argv[2], argv[3], argv[4], argv[5], argv[6], argv[7], argv[8], argv[9]);
}
