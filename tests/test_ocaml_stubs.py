"""The OCaml stub layer (owl_amd/ocaml/owl_amd_stubs.c) compiled against a
test-only mock of the OCaml runtime headers (tests/mock_caml) and a recorder
standing in for libowl_amd, then driven through ctypes:

  * the committed file is what the generator's table gives;
  * the generated stubs plus the complex max-pooling stubs Owl keeps are
    exactly the stubs Owl's externals bind (tests/golden/owl_externals.json,
    extracted from owl_ndarray_{conv,pool,upsampling}.ml), with the same
    argument kinds in the same order;
  * each library entry point the stubs call is exported by libowl_amd.so and
    matches its prototype in include/owl_amd.h (the file compiles with
    -Werror against the header);
  * ``_native`` hands the library the Bigarrays' data pointers and the
    untagged ints in order, inside caml_enter/leave_blocking_section;
  * ``_bytecode`` forwards argv in order to the same call;
  * a non-zero library status raises through caml_failwith with the
    library's message, after the runtime lock is taken back."""
import ctypes
import json
import os
import subprocess

import pytest

from owl_amd.ocaml import gen_stubs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HERE = os.path.join(ROOT, "tests")

MOCK_RUNTIME = r"""
#include <string.h>
int mock_blocking = 0, mock_enter = 0, mock_leave = 0, mock_fail_count = 0, mock_fail_blocking = -1;
char mock_fail_msg[256];
void caml_enter_blocking_section(void) { mock_blocking = 1; mock_enter++; }
void caml_leave_blocking_section(void) { mock_blocking = 0; mock_leave++; }
void caml_failwith(const char *m) {
  strncpy(mock_fail_msg, m, 255);
  mock_fail_count++;
  mock_fail_blocking = mock_blocking;
}
"""


def recorder_source():
    out = ["#include <stdint.h>", "extern int mock_blocking;",
           "int64_t rec_args[32]; int rec_n = 0, rec_blocking = -1, rec_rc = 0, rec_calls = 0;",
           "char rec_name[128];",
           'const char *owl_amd_last_error(void) { return "mock library error"; }']
    for native, _, entry, kinds, _ in gen_stubs.stubs():
        params = ", ".join(f"void *p{i}" if k == "arr" else f"int64_t p{i}" for i, k in enumerate(kinds))
        body = [f"rec_args[{i}] = (int64_t)(intptr_t)p{i};" if k == "arr" else f"rec_args[{i}] = p{i};"
                for i, k in enumerate(kinds)]
        out.append(f"int {entry}({params}) {{ " + " ".join(body) +
                   f' rec_n = {len(kinds)}; rec_blocking = mock_blocking; rec_calls++;'
                   f' {{ const char *s = "{entry}"; int i = 0; for (; s[i]; ++i) rec_name[i] = s[i]; rec_name[i] = 0; }}'
                   " return rec_rc; }")
    return "\n".join(out) + "\n"


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    d = tmp_path_factory.mktemp("shim")
    (d / "mock_runtime.c").write_text(MOCK_RUNTIME)
    (d / "recorder.c").write_text(recorder_source())
    so = d / "libshim.so"
    # the stubs against the real header, warnings as errors: every call
    # matches the library's prototype
    subprocess.run(["gcc", "-std=c11", "-O1", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter", "-fPIC",
                    "-c", "-I", os.path.join(HERE, "mock_caml"), "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "owl_amd", "ocaml", "owl_amd_stubs.c"), "-o", str(d / "stubs.o")],
                   check=True)
    subprocess.run(["gcc", "-shared", "-fPIC", "-O1", str(d / "stubs.o"), str(d / "recorder.c"),
                    str(d / "mock_runtime.c"), "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


def test_committed_file_is_generated():
    with open(gen_stubs.OUT) as f:
        assert f.read() == gen_stubs.generate(), "run python -m owl_amd.ocaml.gen_stubs"


def test_stub_set_matches_owl_externals():
    ext = json.load(open(os.path.join(HERE, "golden", "owl_externals.json")))
    ref = {e["native"]: e for e in ext}
    assert len(ref) == len(ext) == 148
    gen = {n: (b, kinds) for n, b, _, kinds, _ in gen_stubs.stubs()}
    kept = {f"stub_{k}_ndarray_{op}_native" for k, op in gen_stubs.OWL_KEEP}
    assert set(gen) | kept == set(ref) and not (set(gen) & kept)
    for n, (b, kinds) in gen.items():
        assert ref[n]["bytecode"] == b
        assert ref[n]["args"] == kinds, n


def test_library_exports_every_entry_point():
    from owl_amd import _lib
    L = _lib.lib()
    for _, _, entry, _, _ in gen_stubs.stubs():
        assert hasattr(L, entry), entry


def _arr_value(keep, data_addr):
    # a Bigarray custom block: field 0 the ops pointer, then caml_ba_array
    blk = (ctypes.c_int64 * 8)(0x1234, data_addr, 4, 0, 0, 2, 3, 4)
    keep.append(blk)
    return ctypes.addressof(blk)


def _int_values(n_arr, n):
    vals = [(-3 if i % 5 == 0 else (7 * i + 1) * (1 if i % 2 else 1 << 33)) for i in range(n_arr, n)]
    return vals, [(v << 1) | 1 for v in vals]


def test_native_and_bytecode_forward_every_argument(shim):
    rec_args = (ctypes.c_int64 * 32).in_dll(shim, "rec_args")
    for native, bytecode, entry, kinds, _ in gen_stubs.stubs():
        n = len(kinds)
        n_arr = kinds.count("arr")
        keep = []
        datas = [0x10000 * (i + 1) + 16 * len(native) for i in range(n_arr)]
        vals = [_arr_value(keep, a) for a in datas]
        ints, tagged = _int_values(n_arr, n)
        vals += tagged
        expect = datas + ints
        fn = getattr(shim, native)
        fn.argtypes = [ctypes.c_int64] * n
        fn.restype = ctypes.c_int64
        bc = getattr(shim, bytecode)
        bc.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int]
        bc.restype = ctypes.c_int64
        for call in (lambda: fn(*vals), lambda: bc((ctypes.c_int64 * n)(*vals), n)):
            for i in range(32):
                rec_args[i] = -999
            before = ctypes.c_int.in_dll(shim, "rec_calls").value
            assert call() == 1                                    # Val_unit
            assert ctypes.c_int.in_dll(shim, "rec_calls").value == before + 1
            assert ctypes.c_char_p(ctypes.addressof(ctypes.c_char.in_dll(shim, "rec_name"))).value.decode() \
                == entry
            assert ctypes.c_int.in_dll(shim, "rec_n").value == n
            assert list(rec_args[:n]) == expect, native
            assert ctypes.c_int.in_dll(shim, "rec_blocking").value == 1   # runtime lock released
            assert ctypes.c_int.in_dll(shim, "mock_blocking").value == 0  # and taken back
    assert ctypes.c_int.in_dll(shim, "mock_enter").value == ctypes.c_int.in_dll(shim, "mock_leave").value


def test_library_error_raises_failure(shim):
    rc = ctypes.c_int.in_dll(shim, "rec_rc")
    native, _, _, kinds, _ = gen_stubs.stubs()[0]
    keep = []
    vals = [_arr_value(keep, 0x4000 + i) for i in range(kinds.count("arr"))]
    vals += _int_values(len(vals), len(kinds))[1]
    fn = getattr(shim, native)
    fn.argtypes = [ctypes.c_int64] * len(kinds)
    fn.restype = ctypes.c_int64
    n0 = ctypes.c_int.in_dll(shim, "mock_fail_count").value
    rc.value = 0
    fn(*vals)
    assert ctypes.c_int.in_dll(shim, "mock_fail_count").value == n0
    rc.value = 3
    try:
        fn(*vals)
    finally:
        rc.value = 0
    assert ctypes.c_int.in_dll(shim, "mock_fail_count").value == n0 + 1
    assert ctypes.c_char_p(ctypes.addressof(ctypes.c_char.in_dll(shim, "mock_fail_msg"))).value == \
        b"mock library error"
    assert ctypes.c_int.in_dll(shim, "mock_fail_blocking").value == 0
