/* Test-only stand-in for OCaml's <caml/bigarray.h>: a Bigarray is a custom
 * block whose field 0 is its operations pointer and whose data starts at
 * field 1 with struct caml_ba_array (data pointer first). */
#pragma once
#include "mlvalues.h"
struct caml_ba_array {
  void *data;
  intnat num_dims;
  intnat flags;
  void *proxy;
  intnat dim[1];
};
#define Data_custom_val(v) ((void *)&Field((v), 1))
#define Caml_ba_array_val(v) ((struct caml_ba_array *)Data_custom_val(v))
#define Caml_ba_data_val(v) (Caml_ba_array_val(v)->data)
