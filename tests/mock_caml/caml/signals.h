/* Test-only stand-in for OCaml's <caml/signals.h>. */
#pragma once
void caml_enter_blocking_section(void);
void caml_leave_blocking_section(void);
