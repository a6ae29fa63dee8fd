"""The gfx950 machine code of the ghost-cell BC kernel, replayed on the CPU
(tests/isa) over every launch of one wall-bounded step captured from the
host emulation: no access outside an allocation, no load based on an
unwritten scalar register, results equal to the emulation's.

An earlier k_bc_axis, which selected kernel-argument fields by a run-time
axis index, compiled to code that loaded an extent through a scalar
register pair its axis-2 path never wrote; this replay is how that was
found, and it runs here so that such code cannot return unnoticed.  The
precursor kernels, the precursor domain's forcing and the particle list,
bin and error kernels are replayed the same way.  (The full replay of
every kernel of the step takes tens of minutes and is run by hand:
tests/isa/capture.py, tests/isa/replay.py.)"""
import glob
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EMU = os.path.join(ROOT, "tests", "emu")
ISA = os.path.join(ROOT, "tests", "isa")


def test_bc_kernel_machine_code_replays_clean(tmp_path):
    from test_emu import _asan
    from bluebottle_amd import runtime
    asan = _asan()
    if not os.path.isabs(asan):
        pytest.skip("AddressSanitizer runtime not available")
    assert runtime.libs_present(), "build the library first"
    subprocess.run(["make", "-s", "-j8", "-C", EMU], check=True)
    cap = tmp_path / "cap"
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               BB_EMU_ISA_KERNELS="k_bc_axis", BB_EMU_ISA_MAX="60")
    r = subprocess.run([sys.executable, os.path.join(ISA, "capture.py"),
                        "channel_free", str(cap), "1"], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    # set-up and one step: dom_BC twice, BC_star twice, BC_phi, dom_BC,
    # BC_p and dom_BC, three axis launches each
    launches = glob.glob(str(cap / "*.json"))
    assert len(launches) == 24, len(launches)
    assert all(json.load(open(f))["kernel"] == "k_bc_axis" for f in launches)
    env = dict(os.environ)
    env["BB_ISA_OBJDIR"] = str(tmp_path / "co")
    sys.path.insert(0, ISA)
    import replay
    replay.extract_code_objects(str(tmp_path / "co"))
    r = subprocess.run([sys.executable, os.path.join(ISA, "replay.py"),
                        str(cap)], env=env, capture_output=True, text=True,
                       timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "replayed 24 launches, 0 with findings" in r.stdout


def test_precursor_kernels_machine_code_replays_clean(tmp_path):
    """The PRECURSOR ghost fill (k_bc_axis's plane path) and the plane
    extraction (k_yank) replayed the same way, on the all-faces case of
    tests/test_gpu_precursor.py: two BC calls of three axis launches and
    fifteen extractions."""
    from test_emu import _asan
    from bluebottle_amd import runtime
    asan = _asan()
    if not os.path.isabs(asan):
        pytest.skip("AddressSanitizer runtime not available")
    assert runtime.libs_present(), "build the library first"
    subprocess.run(["make", "-s", "-j8", "-C", EMU], check=True)
    cap = tmp_path / "cap"
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               BB_EMU_ISA_KERNELS="k_bc_axis,k_yank", BB_EMU_ISA_MAX="60")
    r = subprocess.run([sys.executable, os.path.join(ISA, "capture.py"),
                        "precursor", str(cap)], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    launches = glob.glob(str(cap / "*.json"))
    assert len(launches) == 21, len(launches)
    env = dict(os.environ)
    env["BB_ISA_OBJDIR"] = str(tmp_path / "co")
    sys.path.insert(0, ISA)
    import replay
    replay.extract_code_objects(str(tmp_path / "co"))
    r = subprocess.run([sys.executable, os.path.join(ISA, "replay.py"),
                        str(cap)], env=env, capture_output=True, text=True,
                       timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "replayed 21 launches, 0 with findings" in r.stdout


def test_precursor_domain_kernels_machine_code_replays_clean(tmp_path):
    """The precursor domain's linear forcing (k_turb_sums, k_turb_final,
    k_turb_apply), cuda_compute_energy's reduction and the three
    colocations, replayed the same way."""
    from test_emu import _asan
    from bluebottle_amd import runtime
    asan = _asan()
    if not os.path.isabs(asan):
        pytest.skip("AddressSanitizer runtime not available")
    assert runtime.libs_present(), "build the library first"
    subprocess.run(["make", "-s", "-j8", "-C", EMU], check=True)
    cap = tmp_path / "cap"
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               BB_EMU_ISA_KERNELS="k_turb_sums,k_turb_final,k_turb_apply,"
               "k_colocate", BB_EMU_ISA_MAX="8")
    r = subprocess.run([sys.executable, os.path.join(ISA, "capture.py"),
                        "turb_forcing", str(cap)], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    launches = glob.glob(str(cap / "*.json"))
    # forcing: sums, final, apply; energy: sums, final; three colocations
    assert len(launches) == 8, len(launches)
    env = dict(os.environ)
    env["BB_ISA_OBJDIR"] = str(tmp_path / "co")
    sys.path.insert(0, ISA)
    import replay
    replay.extract_code_objects(str(tmp_path / "co"))
    r = subprocess.run([sys.executable, os.path.join(ISA, "replay.py"),
                        str(cap)], env=env, capture_output=True, text=True,
                       timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "replayed 8 launches, 0 with findings" in r.stdout


def test_particle_list_and_bin_kernels_machine_code_replays_clean(tmp_path):
    """The face lists (k_face_count, k_face_scan, k_face_write: per-block
    counts, offsets, and each block's faces written in index order), the bin
    placement (k_bin_fill's slots, k_bin_scatter, k_bin_place's rank within
    the bin) and the wavefront-wide k_lamb_err, on three spheres with a pair
    0.3 r apart."""
    from test_emu import _asan
    from bluebottle_amd import runtime
    asan = _asan()
    if not os.path.isabs(asan):
        pytest.skip("AddressSanitizer runtime not available")
    assert runtime.libs_present(), "build the library first"
    subprocess.run(["make", "-s", "-j8", "-C", EMU], check=True)
    cap = tmp_path / "cap"
    env = dict(os.environ)
    env.update(LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0",
               BB_EMU_ISA_KERNELS="k_face_count,k_face_scan,k_face_write,"
               "k_bin_fill,k_bin_scatter,k_bin_place,k_lamb_err",
               BB_EMU_ISA_MAX="4")
    r = subprocess.run([sys.executable, os.path.join(ISA, "capture.py"),
                        "particle_kernels", str(cap)], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    launches = glob.glob(str(cap / "*.json"))
    # the face lists' three kernels, one bin pass of three, one k_lamb_err
    assert len(launches) == 7, len(launches)
    env = dict(os.environ)
    env["BB_ISA_OBJDIR"] = str(tmp_path / "co")
    sys.path.insert(0, ISA)
    import replay
    replay.extract_code_objects(str(tmp_path / "co"))
    r = subprocess.run([sys.executable, os.path.join(ISA, "replay.py"),
                        str(cap)], env=env, capture_output=True, text=True,
                       timeout=1200)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "replayed 7 launches, 0 with findings" in r.stdout
