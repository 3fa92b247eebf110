"""Writes reference_interface.json: the NAMES of the virtual member functions of
the reference's Grid, Hardware and DataAcquisition classes and of the data
members of its Config (src/grid.h, src/hardware.h, src/data_Acquisition.h,
src/config.h), read as text from a checkout of the reference.  Names only: the
binding test checks that every `override` of the MI355X backend names a virtual
of the reference and that the backend reads no Config member the reference
lacks.

    python tests/golden/make_reference_interface.py /root/reference/src
"""
import json
import os
import re
import sys


def _strip_comments(s):
    s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
    return re.sub(r"//[^\n]*", "", s)


def virtuals(text):
    """names of the functions declared `virtual` (destructors included as ~X)"""
    out = set()
    for m in re.finditer(r"\bvirtual\b[^;{(]*?(~?\w+)\s*\(", _strip_comments(text)):
        out.add(m.group(1))
    return sorted(out)


def non_virtual_destructor(text, cls):
    t = _strip_comments(text)
    m = re.search(r"(\bvirtual\s+)?~" + cls + r"\s*\(", t)
    return bool(m) and not m.group(1)


def config_members(text):
    """data members of class Config: declarations 'type name[, name];' in its body"""
    t = _strip_comments(text)
    body = t[t.index("class Config"):]
    body = body[body.index("{") + 1:]
    depth, end = 1, 0
    for i, ch in enumerate(body):
        depth += ch == "{"
        depth -= ch == "}"
        if depth == 0:
            end = i
            break
    body = body[:end]
    names = set()
    for stmt in body.split(";"):
        s = stmt.strip()
        if not s or "(" in s or s.startswith(("public", "private", "protected")):
            continue
        s = re.sub(r"^(public|private|protected)\s*:\s*", "", s)
        s = re.sub(r"\[[^\]]*\]", "", s)
        s = re.sub(r"=[^,]*", "", s)
        parts = [p.strip() for p in s.split(",")]
        first = parts[0].split()
        if len(first) < 2:
            continue
        names.add(first[-1].lstrip("*&"))
        for p in parts[1:]:
            names.add(p.split()[-1].lstrip("*&"))
    return sorted(n for n in names if re.fullmatch(r"\w+", n))


def main(src):
    rd = lambda f: open(os.path.join(src, f)).read()
    out = {
        "source": "vetienne74/hpcscan src/grid.h, src/hardware.h, src/data_Acquisition.h, src/config.h",
        "Grid": {"virtuals": virtuals(rd("grid.h")), "destructor_virtual": not non_virtual_destructor(rd("grid.h"), "Grid")},
        "Hardware": {"virtuals": virtuals(rd("hardware.h")),
                     "destructor_virtual": not non_virtual_destructor(rd("hardware.h"), "Hardware")},
        "DataAcquisition": {"virtuals": virtuals(rd("data_Acquisition.h")),
                            "destructor_virtual": not non_virtual_destructor(rd("data_Acquisition.h"),
                                                                             "DataAcquisition")},
        "Config": {"members": config_members(rd("config.h"))},
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_interface.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {path}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
