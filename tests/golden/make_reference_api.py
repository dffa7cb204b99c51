"""Writes reference_api.txt: the public C prototypes of the reference's
include/pastix.h and include/pastix/order.h, one per line as
"name|return type|parameter types" (parameter names dropped, whitespace
normalized).  Run in a tree that holds the reference at /root/reference;
the file it writes is the fixture tests/test_abi.py checks the library's
headers and exports against.
"""
import os
import re
import sys

REF = "/root/reference"
HEADERS = ["include/pastix.h", "include/pastix/order.h"]


def normalize_type(t):
    t = re.sub(r"\s+", " ", t.strip())
    t = re.sub(r"\s*\*\s*", "*", t)
    t = re.sub(r"\bconst\b\s*", "const ", t).strip()
    return t


def param_type(p):
    p = re.sub(r"\s+", " ", p.strip())
    if p in ("void", ""):
        return p
    # drop the parameter name (the last identifier), keep the stars
    m = re.match(r"^(.*?[\s\*])([A-Za-z_]\w*)\s*(\[\s*\])?$", p)
    if m:
        p = m.group(1) + ("*" if m.group(3) else "")
    return normalize_type(p)


def prototypes(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    text = re.sub(r"^\s*#.*$", "", text, flags=re.M)
    text = re.sub(r"\b(BEGIN_C_DECLS|END_C_DECLS)\b", "", text)
    out = {}
    for m in re.finditer(r"([A-Za-z_][\w\s\*]*?)\b([A-Za-z_]\w*)\s*\(([^;{}()]*)\)\s*;", text):
        ret, name, params = m.group(1), m.group(2), m.group(3)
        ret = normalize_type(ret.replace("extern", ""))
        if not ret or ret in ("return", "typedef") or name in ("if", "while"):
            continue
        ps = [param_type(p) for p in params.split(",")] if params.strip() else ["void"]
        out[name] = "%s|%s|%s" % (name, ret, ",".join(ps))
    return out


def main():
    protos = {}
    for h in HEADERS:
        protos.update(prototypes(open(os.path.join(REF, h)).read()))
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "reference_api.txt"), "w") as f:
        for name in sorted(protos):
            f.write(protos[name] + "\n")
    print("%d prototypes" % len(protos), file=sys.stderr)


if __name__ == "__main__":
    main()
