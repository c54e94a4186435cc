/* Test-only stand-in for the parts of OCaml's <caml/mlvalues.h> that
 * owl_amd/ocaml/owl_amd_stubs.c uses, with the runtime's encodings: a value
 * is a machine word, an int is tagged as 2 n + 1, a block's fields follow
 * its address.  Used by tests/test_ocaml_stubs.py only. */
#pragma once
#include <stdint.h>
typedef intptr_t intnat;
typedef intnat value;
#define CAMLprim
#define Long_val(x) ((x) >> 1)
#define Val_long(x) ((value)(((uintptr_t)(intnat)(x) << 1)) + 1)
#define Val_unit Val_long(0)
#define Field(x, i) (((value *)(x))[i])
