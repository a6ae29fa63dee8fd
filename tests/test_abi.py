"""The C ABI without a GPU: struct layouts match the reference's byte for
byte (as compiled from include/bluebottle.h by gcc and as mirrored in
ctypes), and the built libraries load and export every function the
header declares."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

from bluebottle_amd import abi, runtime

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bluebottle.h")

# src/domain.h / src/bluebottle.h / src/particle.h layouts on x86-64 with
# real == double (sizes and the offsets the raw restart files depend on)
LAYOUT = {
    "grid_info": 192, "dom_struct": 912, "BC": 720, "part_struct": 928,
    "g_struct": 72, "gradP_struct": 72, "cage_struct": 72,
}
OFFSETS = {
    ("part_struct", "nodes"): 488, ("part_struct", "rho"): 592,
    ("part_struct", "cage"): 616, ("part_struct", "rs"): 696,
}


def _compile_probe(tmp_path):
    src = tmp_path / "probe.c"
    lines = ['#include <stdio.h>', '#include <stddef.h>',
             '#include "bluebottle.h"', 'int main(void) {']
    for name in LAYOUT:
        lines.append('printf("%s %%zu\\n", sizeof(%s));' % (name, name))
    for (st, field) in OFFSETS:
        lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (
            st, field, st, field))
    lines.append("return 0; }")
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"),
                    str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True,
                         text=True).stdout
    return dict(line.split() for line in out.strip().splitlines())


def test_header_layout_gcc(tmp_path):
    got = _compile_probe(tmp_path)
    for name, size in LAYOUT.items():
        assert int(got[name]) == size, name
    for (st, field), off in OFFSETS.items():
        assert int(got["%s.%s" % (st, field)]) == off


def test_ctypes_mirror_layout():
    abi.check_sizes()
    for name, size in LAYOUT.items():
        assert C.sizeof(getattr(abi, name)) == size, name
    for (st, field), off in OFFSETS.items():
        assert getattr(getattr(abi, st), field).offset == off


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(
        r"^\s*(?:const\s+)?(?:void|real|int|char|double)\s*\*?\s*"
        r"([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(names))


def defined_symbols(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], check=True,
                         capture_output=True, text=True).stdout
    return {line.split()[-1] for line in out.splitlines() if line.strip()}


def test_header_declares_reference_entry_points():
    names = declared_functions()
    for ep in ("cuda_U_star_2", "cuda_ustar_helmholtz", "cuda_PP_bicgstab",
               "cuda_project", "cuda_quad_interp", "cuda_Lamb", "cuda_part_BC",
               "cuda_dom_malloc", "cuda_find_dt", "cuda_lamb_err",
               "cuda_build_cages", "cuda_move_parts", "recorder_bicgstab"):
        assert ep in names


def test_libraries_export_every_declared_function():
    assert runtime.libs_present(), "build the library first"
    hip = defined_symbols(runtime.HIP_LIB)
    host = defined_symbols(runtime.HOST_LIB)
    missing = []
    for name in declared_functions():
        # recorder_bicgstab is the driver's (src/recorder.c); the HIP
        # library only references it
        where = host if name == "recorder_bicgstab" else hip
        if name not in where:
            missing.append(name)
    assert not missing, missing


def test_hip_library_leaves_driver_globals_undefined():
    """The globals are the driver's (src/bluebottle.c:29-187,
    src/particle.c:29-75): the HIP library must reference, not define,
    them, so the reference driver's definitions bind."""
    hip = defined_symbols(runtime.HIP_LIB)
    for g in ("dom", "Dom", "_dom", "u", "_u", "parts", "_parts", "dt", "nu",
              "bc", "pp_residual", "coeff_stride", "nparts", "binDom"):
        assert g not in hip, g
    host = defined_symbols(runtime.HOST_LIB)
    for g in ("dom", "Dom", "u", "_u", "parts", "dt", "bc", "nparts"):
        assert g in host, g


def test_library_loads_without_gpu():
    host, hip = runtime.load()
    for name in runtime.ENTRY_POINTS:
        assert getattr(hip, name) is not None
    # the ctypes mirror agrees with the library's view of the globals
    assert C.sizeof(runtime.dom()) == LAYOUT["dom_struct"]


def test_outside_path_entry_points_exit_with_a_message():
    """The -DTEST checks and unused helpers are exported so the reference
    driver links (src/bluebottle.c); calling one ends the process with a
    message naming it, without touching a GPU."""
    code = ("import sys; sys.path.insert(0, %r)\n"
            "from bluebottle_amd import runtime\n"
            "host, hip = runtime.load()\n"
            "hip.cuda_collisions()\n"
            "print('returned')\n") % ROOT
    p = subprocess.run([sys.executable, "-c", code], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 1, (p.returncode, p.stdout, p.stderr)
    assert "cuda_collisions" in p.stderr
    assert "outside this library's path" in p.stderr
    assert "returned" not in p.stdout


def _undefined_unversioned(path):
    out = subprocess.run(["nm", "-D", "-u", path], check=True,
                         capture_output=True, text=True).stdout
    names = set()
    for line in out.splitlines():
        parts = line.split()
        if parts and "@" not in parts[-1]:
            names.add(parts[-1])
    return names - {"__gmon_start__"}


def test_reference_driver_links_against_the_library():
    """The link of the reference's unchanged C driver (the objects of its
    Makefile:49-56) against libbluebottle_hip.so, checked symbol by symbol
    from tests/golden/driver_symbols.json (written from the driver's source
    text by tests/golden/make_driver_symbols.py): every cuda_* function the
    driver calls is exported, and every symbol the library leaves for the
    executable to define is a global or a function the driver defines.  (A
    compile of the driver itself needs cgnslib.h, which this image lacks.)"""
    import json
    with open(os.path.join(ROOT, "tests", "golden",
                           "driver_symbols.json")) as f:
        drv = json.load(f)
    hip = defined_symbols(runtime.HIP_LIB)
    missing = [c for c in drv["calls"] if c not in hip]
    assert not missing, "driver calls not exported: %s" % missing
    provided = set(drv["globals"]) | set(drv["functions"])
    unresolved = sorted(_undefined_unversioned(runtime.HIP_LIB) - provided)
    assert not unresolved, "library needs symbols the driver does not " \
        "define: %s" % unresolved
