#!/usr/bin/env python3
"""Extract Owl's conv/pool golden vectors into a JSON fixture.

Reads the reference's alcotest files as *text* (no OCaml is run or needed) and
writes ``owl_vectors.json`` next to this script.  Each record holds the call the
test makes (helper name, shapes, stride or rate, padding) and the expected
values the test asserts; how the helper builds its inputs (``sequential``,
``ones``, ``div_scalar``) is restated in ``tests/owl_cases.py``.

Sources (paths relative to the reference root):
  test/unit_conv2d_generic.ml, unit_conv3d_generic.ml, unit_pool2d_generic.ml,
  unit_pool3d_generic.ml, unit_dilated_conv2d.ml, unit_dilated_conv3d.ml,
  unit_transpose_conv2d.ml, unit_transpose_conv3d.ml; and
  unit_upsampling_generic.ml into owl_upsampling_vectors.json

Usage:  python tests/golden/extract_owl_vectors.py [/root/reference]
"""
import json
import os
import re
import sys

FILES = [
    "unit_conv2d_generic.ml",
    "unit_conv3d_generic.ml",
    "unit_pool2d_generic.ml",
    "unit_pool3d_generic.ml",
    "unit_dilated_conv2d.ml",
    "unit_dilated_conv3d.ml",
    "unit_transpose_conv2d.ml",
    "unit_transpose_conv3d.ml",
]

ARR = r"\[\|(.*?)\|\]"


def parse_array(txt):
    vals = [t.strip() for t in txt.replace("\n", " ").split(";")]
    return [float(v.rstrip(".") if v.endswith(".") and v.count(".") == 1 and v[:-1].lstrip("-").isdigit() else v)
            for v in vals if v]


def strip_comments(src):
    # OCaml comments (* ... *) are not nested in these files
    return re.sub(r"\(\*.*?\*\)", lambda m: " " * len(m.group(0)), src, flags=re.S)


def line_of(src, pos):
    return src.count("\n", 0, pos) + 1


def extract_file(path, fname):
    raw = open(path).read()
    src = strip_comments(raw)
    mods = [(m.start(), m.group(1)) for m in re.finditer(r"module\s+(To_test_\w+)\s*=\s*struct", src)]
    # registered wrappers: "(To_test_xxx.funNN ())" inside Alcotest checks that
    # appear in the test_set list
    wrappers = dict(re.findall(r"let\s+(fun_\w+)\s*\(\)\s*=\s*Alcotest\.\(check bool\)\s*\"[^\"]*\"\s*true\s*"
                               r"\((To_test_\w+\.fun\d+)\s*\(\)\)", src, flags=re.S))
    ts = re.search(r"let\s+test_set\s*=\s*\[(.*?)\]", src, flags=re.S)
    registered = set()
    if ts:
        for w in re.findall(r"(fun_\w+)\s*\)?\s*(?:;|$)", ts.group(1)):
            pass
        for w in re.findall(r"`\w+\s*,\s*(fun_\w+)", ts.group(1)):
            if w in wrappers:
                registered.add(wrappers[w])
    out = []
    for i, (mpos, mname) in enumerate(mods):
        mend = mods[i + 1][0] if i + 1 < len(mods) else len(src)
        body = src[mpos:mend]
        funs = [(m.start(), m.group(1)) for m in re.finditer(r"let\s+(fun\d+)\s*\(\)\s*=", body)]
        for j, (fpos, fname_) in enumerate(funs):
            fend = funs[j + 1][0] if j + 1 < len(funs) else len(body)
            fb = body[fpos:fend]
            rec = {"file": "test/" + fname, "line": line_of(src, mpos + fpos), "module": mname,
                   "fun": fname_, "registered": f"{mname}.{fname_}" in registered}
            e = re.search(r"let\s+expected\s*=\s*" + ARR, fb, flags=re.S)
            if e:
                rec["expected"] = parse_array(e.group(1))
                call = re.search(r"verify_value\s+(.*?)\s+expected\b", fb[e.end():], flags=re.S)
                if not call:
                    raise RuntimeError(f"{fname}:{mname}.{fname_}: no verify_value call")
                args = call.group(1)
                hm = re.match(r"\(\s*(\w+)\s*~stride:\s*" + ARR + r"\s*\)", args, flags=re.S)
                if hm:
                    rec["helper"] = hm.group(1)
                    rec["stride"] = [int(v) for v in parse_array(hm.group(2))]
                    args = args[hm.end():]
                else:
                    hm = re.match(r"(\w+)", args)
                    rec["helper"] = hm.group(1)
                    args = args[hm.end():]
                rec["seq"] = bool(re.search(r"~seq:\s*true", args))
                args = re.sub(r"~seq:\s*\w+", "", args)
                arrs = [[int(v) for v in parse_array(a)] for a in re.findall(ARR, args, flags=re.S)]
                pad = re.search(r"\b(SAME|VALID)\b", args).group(1)
                rec["input_shape"], rec["kernel_shape"], rec["third"] = arrs[0], arrs[1], arrs[2]
                rec["padding"] = pad
            else:
                t = re.search(r"test_forward\s+" + ARR + r"\s*" + ARR + r"\s*" + ARR + r"\s*(SAME|VALID)\s*" + ARR,
                              fb, flags=re.S)
                if not t:
                    raise RuntimeError(f"{fname}:{mname}.{fname_}: unrecognised test body")
                rec["helper"] = "test_forward"
                rec["input_shape"] = [int(v) for v in parse_array(t.group(1))]
                rec["kernel_shape"] = [int(v) for v in parse_array(t.group(2))]
                rec["third"] = [int(v) for v in parse_array(t.group(3))]
                rec["padding"] = t.group(4)
                rec["magic"] = parse_array(t.group(5))
                rec["seq"] = False
            out.append(rec)
    return out


def extract_upsampling(path, fname):
    """test/unit_upsampling_generic.ml: forward tests (``test_upsampling2d_forward
    shape size``, compared with ``repeat``) and backward golden vectors
    (``verify_value shape size expected``, the backward of an all-ones input
    with a ``sequential`` output gradient)."""
    src = strip_comments(open(path).read())
    registered = set(re.findall(r"\((To_test_\w+\.fun\d+)\s*\(\)\)", src))
    out = []
    for m in re.finditer(r"let\s+(fun\d+)\s*\(\)\s*=(.*?)(?=let\s+fun\d+\s*\(\)|\bend\b)", src, flags=re.S):
        name, body = m.group(1), m.group(2)
        rec = {"file": "test/" + fname, "line": line_of(src, m.start()), "module": "To_test_upsampling2d",
               "fun": name, "registered": f"To_test_upsampling2d.{name}" in registered}
        f = re.search(r"test_upsampling2d_forward\s+" + ARR + r"\s*" + ARR, body, flags=re.S)
        if f:
            rec["helper"] = "test_upsampling2d_forward"
            rec["input_shape"] = [int(v) for v in parse_array(f.group(1))]
            rec["size"] = [int(v) for v in parse_array(f.group(2))]
        else:
            e = re.search(r"let\s+expected\s*=\s*" + ARR, body, flags=re.S)
            v = re.search(r"verify_value\s+" + ARR + r"\s*" + ARR + r"\s*expected", body, flags=re.S)
            if not (e and v):
                raise RuntimeError(f"{fname}:{name}: unrecognised test body")
            rec["helper"] = "upsampling2d_backward_verify_value"
            rec["expected"] = parse_array(e.group(1))
            rec["input_shape"] = [int(x) for x in parse_array(v.group(1))]
            rec["size"] = [int(x) for x in parse_array(v.group(2))]
        out.append(rec)
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    recs = []
    for f in FILES:
        recs.extend(extract_file(os.path.join(ref, "test", f), f))
    here = os.path.dirname(os.path.abspath(__file__))
    dst = os.path.join(here, "owl_vectors.json")
    with open(dst, "w") as fh:
        json.dump({"source": "owlbarn/owl test/*.ml (golden vectors, transcribed by extract_owl_vectors.py)",
                   "cases": recs}, fh, indent=1)
    print(f"wrote {len(recs)} cases to {dst}")
    up = extract_upsampling(os.path.join(ref, "test", "unit_upsampling_generic.ml"), "unit_upsampling_generic.ml")
    dst = os.path.join(here, "owl_upsampling_vectors.json")
    with open(dst, "w") as fh:
        json.dump({"source": "owlbarn/owl test/unit_upsampling_generic.ml (transcribed by extract_owl_vectors.py)",
                   "cases": up}, fh, indent=1)
    print(f"wrote {len(up)} cases to {dst}")


if __name__ == "__main__":
    main()
