"""Writes driver_symbols.json: what the reference's C driver files
(src/bluebottle.c, domain.c, particle.c, precursor.c, recorder.c,
seeder.c, vtk.c, rng.c -- the objects of its Makefile:49-56) call and
define, read from their source text:

  * "calls":   every cuda_* function they name;
  * "globals": every file-scope variable they define;
  * "functions": every function they define.

The link of the unchanged driver against libbluebottle_hip.so resolves
exactly these names, so tests/test_abi.py checks the library against this
list without the reference present.  Run in a container holding the
reference sources:  python tests/golden/make_driver_symbols.py
"""
import json
import os
import re

SRC = "/root/reference/src"
FILES = ["bluebottle.c", "domain.c", "particle.c", "precursor.c",
         "recorder.c", "seeder.c", "vtk.c", "rng.c"]
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                   "driver_symbols.json")


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def file_scope(text):
    """The text outside every brace pair (file-scope declarations)."""
    out, depth = [], 0
    for ch in text:
        if ch == "{":
            depth += 1
            out.append(";")
        elif ch == "}":
            depth -= 1
        elif depth == 0:
            out.append(ch)
    return "".join(out)


def main():
    calls, glob, funcs = set(), set(), set()
    decl = re.compile(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?[\s\*]"
                      r"(\**)([A-Za-z_]\w*)\s*(\[[^\]]*\])?\s*(=[^;]*)?$")
    for name in FILES:
        with open(os.path.join(SRC, name)) as f:
            text = strip_comments(f.read())
        text = re.sub(r"^\s*#[^\n]*", " ", text, flags=re.M)
        calls.update(re.findall(r"\b(cuda_\w+)\s*\(", text))
        top = file_scope(text)
        for stmt in top.split(";"):
            stmt = stmt.strip()
            if not stmt or stmt.startswith(("typedef", "extern", "struct ")):
                continue
            m = re.match(r"^[\w\s\*]+?\b([A-Za-z_]\w*)\s*\([^)]*\)\s*$", stmt,
                         flags=re.S)
            if m:
                funcs.add(m.group(1))
                continue
            for part in stmt.split(","):
                m = decl.match(part.strip()) or re.match(
                    r"^\s*(\**)([A-Za-z_]\w*)\s*(\[[^\]]*\])?\s*$", part)
                if m:
                    glob.add(m.group(2))
    glob -= funcs
    with open(OUT, "w") as f:
        json.dump({"source": "src/{%s}" % ",".join(FILES),
                   "calls": sorted(calls), "globals": sorted(glob),
                   "functions": sorted(funcs)}, f, indent=1)
        f.write("\n")
    print("calls %d, globals %d, functions %d" % (len(calls), len(glob),
                                                  len(funcs)))


if __name__ == "__main__":
    main()
