/* Test-only stand-in for OCaml's <caml/fail.h>. */
#pragma once
void caml_failwith(const char *msg);
