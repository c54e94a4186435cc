"""Write tests/golden/owl_externals.json: the C stubs Owl's OCaml externals
bind for the conv / pool / upsampling path, with each external's argument
kinds ("arr" for a Bigarray, "int" for an OCaml int) in order.

Reads, as text, src/owl/core/owl_ndarray_conv.ml, owl_ndarray_pool.ml and
owl_ndarray_upsampling.ml of the reference checkout (default
/root/reference).  The JSON holds only names, argument kinds and the file:line
of each declaration -- no source text.  Run:
    python tests/golden/extract_owl_externals.py [/root/reference]
"""
import json
import os
import re
import sys

FILES = ["src/owl/core/owl_ndarray_conv.ml", "src/owl/core/owl_ndarray_pool.ml",
         "src/owl/core/owl_ndarray_upsampling.ml"]
DECL = re.compile(r"external\s+(\w+)\s*:(.*?)=\s*\"(\w+)\"\s*\"(\w+)\"", re.S)


def kinds(sig):
    parts = [p.strip() for p in sig.split("->")]
    assert parts[-1] == "unit", parts
    out = []
    for p in parts[:-1]:
        if p.endswith("owl_arr"):
            out.append("arr")
        elif p == "int":
            out.append("int")
        else:
            raise ValueError(f"unexpected argument type {p!r}")
    return out


def main(root):
    rows = []
    for rel in FILES:
        text = open(os.path.join(root, rel)).read()
        for m in DECL.finditer(text):
            line = text.count("\n", 0, m.start()) + 1
            rows.append({"external": m.group(1), "bytecode": m.group(3), "native": m.group(4),
                         "args": kinds(m.group(2)), "where": f"{rel}:{line}"})
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "owl_externals.json"), "w") as f:
        json.dump(rows, f, indent=0)
        f.write("\n")
    print(f"{len(rows)} externals")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
