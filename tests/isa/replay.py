"""Replay captured launches (capture.py) on the gfx950 machine code of the
library (gcn.py) and compare with what the host emulation computed
(test infrastructure).

    python tests/isa/replay.py <capture dir> [kernel-substring ...]

For each launch: the kernel's code object entry is found by matching its
demangled parameter list, the kernarg segment is rebuilt (explicit
arguments as captured, hidden arguments from the launch shape), the
machine code runs over the captured memory image, and then
  * every global, scalar and LDS access outside an allocation, the kernarg
    segment or the kernel's LDS is reported (instruction address and lane);
  * every allocation is compared with the emulation's result: integer and
    byte data must match exactly, FP64 data to 1e-12 relative (contraction
    of multiply-adds differs between the two builds).
"""
import glob
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, HERE)
import gcn  # noqa: E402

CXXFILT = "c++filt"
KERNARG_BASE = 0x1000


def demangle(names):
    out = subprocess.run([CXXFILT], input="\n".join(names), capture_output=True,
                         text=True, check=True).stdout.splitlines()
    return dict(zip(names, out))


def params_of(demangled):
    i = demangled.index("(")
    depth = 0
    for j in range(i, len(demangled)):
        if demangled[j] == "(":
            depth += 1
        elif demangled[j] == ")":
            depth -= 1
            if depth == 0:
                return demangled[i + 1:j].replace(" ", "")
    return demangled[i + 1:].replace(" ", "")


class Library:
    def __init__(self, objdir):
        self.kernels = {}
        for co in sorted(glob.glob(os.path.join(objdir, "*.co"))):
            self.kernels.update(gcn.load_code_object(co))
        dm = demangle(list(self.kernels))
        self.by_sig = {}
        for mangled, d in dm.items():
            if "(" not in d:
                continue   # a name this c++filt cannot demangle (bf16 vectors)
            base = d.split("(")[0].split("<")[0]
            if base.startswith("void "):
                base = base[5:]   # template instances print their return type
            self.by_sig.setdefault(base, []).append((params_of(d), mangled))

    def find(self, name, ftype):
        base = name.split("<")[0]
        cands = self.by_sig.get(base, [])
        if len(cands) == 1:
            return self.kernels[cands[0][1]]
        # typeid name of a function pointer type, e.g. "PFv6PPGeom...E"
        sig = subprocess.run([CXXFILT, "-t"], input=ftype, capture_output=True,
                             text=True).stdout.strip()
        want = params_of(sig.replace("(*)", "F")) if "(" in sig else None
        for p, mangled in cands:
            if want is not None and p == want:
                return self.kernels[mangled]
        raise KeyError("no unique code object kernel for %s (%s)" % (name, sig))


def hidden_args(k, grid, block, buf):
    for off, size, kind in k.args:
        if not kind.startswith("hidden"):
            continue
        val = 0
        if kind == "hidden_block_count_x":
            val = grid[0]
        elif kind == "hidden_block_count_y":
            val = grid[1]
        elif kind == "hidden_block_count_z":
            val = grid[2]
        elif kind == "hidden_group_size_x":
            val = block[0]
        elif kind == "hidden_group_size_y":
            val = block[1]
        elif kind == "hidden_group_size_z":
            val = block[2]
        elif kind == "hidden_grid_dims":
            val = 3 if grid[2] * block[2] > 1 else (2 if grid[1] * block[1] > 1 else 1)
        buf[off:off + size] = int(val).to_bytes(size, "little")


def compare(before, after_emu, after_isa, tol=1e-12):
    """Worst FP64 relative difference and list of non-FP mismatches; `tol`
    is the FP64 tolerance (FP32 kernels pass a looser one).  Differences
    where either side reads as a subnormal double are integer data and
    must match exactly."""
    bad = []
    worst = 0.0
    if after_emu == after_isa:
        return worst, bad
    a = np.frombuffer(after_emu, dtype=np.uint8)
    b = np.frombuffer(after_isa, dtype=np.uint8)
    diff = np.nonzero(a != b)[0]
    if len(after_emu) % 8 == 0:
        fa = np.frombuffer(after_emu, dtype=np.float64)
        fb = np.frombuffer(after_isa, dtype=np.float64)
        idx = np.unique(diff // 8)
        tiny = np.finfo(np.float64).tiny
        with np.errstate(all="ignore"):
            sub = ((fa[idx] != 0) & (np.abs(fa[idx]) < tiny)) | \
                ((fb[idx] != 0) & (np.abs(fb[idx]) < tiny))
        if np.any(sub):
            bad.append(("int", int(diff[0]), len(diff), 0.0))
            return worst, bad
        with np.errstate(all="ignore"):
            sc = np.max(np.abs(fa[np.isfinite(fa)])) if np.any(np.isfinite(fa)) else 1.0
            sc = max(sc, 1e-300)
            rel = np.abs(fa[idx] - fb[idx]) / sc
        finite = np.isfinite(fa[idx]) & np.isfinite(fb[idx])
        worst = float(np.max(rel[finite])) if np.any(finite) else 0.0
        if np.any(~finite) or worst > tol:
            # an FP32 buffer: compare as FP32, to FP32 rounding (the two
            # builds contract multiply-adds differently)
            ga = np.frombuffer(after_emu, dtype=np.float32)
            gb = np.frombuffer(after_isa, dtype=np.float32)
            with np.errstate(all="ignore"):
                s32 = max(float(np.max(np.abs(ga[np.isfinite(ga)]))), 1e-30) \
                    if np.any(np.isfinite(ga)) else 1.0
                r32 = np.abs(ga.astype(np.float64) - gb.astype(np.float64)) / s32
            ok32 = np.all(np.isfinite(ga) == np.isfinite(gb)) and \
                float(np.max(r32[np.isfinite(r32)], initial=0.0)) <= 1e-5
            if not ok32:
                bad.append(("fp64", int(idx[0]), len(idx), worst))
            else:
                worst = 0.0
    else:
        bad.append(("bytes", int(diff[0]), len(diff), 0.0))
    return worst, bad


def replay_one(lib, stem, verbose=True):
    meta = json.load(open(stem + ".json"))
    k = lib.find(meta["kernel"], meta["ftype"])
    args = bytes.fromhex(meta["args"])
    kbuf = bytearray(max(k.kernarg_size, len(args)))
    kbuf[:len(args)] = args
    expl = [(o, s) for o, s, kd in k.args if not kd.startswith("hidden")]
    if expl:
        end = max(o + s for o, s in expl)
        if end != len(args):
            raise RuntimeError("%s: kernarg layout %d B captured, %d B in the code "
                               "object" % (k.name, len(args), end))
    hidden_args(k, meta["grid"], meta["block"], kbuf)
    mem = gcn.Memory()
    mem.add(KERNARG_BASE, kbuf)
    raw_before = open(stem + ".before", "rb").read()
    raw_after = open(stem + ".after", "rb").read()
    bufs = []
    pos = 0
    for base, size in meta["allocs"]:
        b = bytearray(raw_before[pos:pos + size])
        mem.add(base, b)
        bufs.append((base, size, pos, b))
        pos += size
    ex = gcn.Exec(k, mem, KERNARG_BASE, tuple(meta["grid"]), tuple(meta["block"]))
    ex.run()
    worst = 0.0
    mism = []
    # kernels instantiated for FP32 vectors: FP32 rounding differences
    # between the two builds reach the FP64 dot partials too
    tol = 1e-6 if "IfE" in k.name or "Pf" in k.name else 1e-12
    for base, size, p, b in bufs:
        w, bad = compare(raw_before[p:p + size], raw_after[p:p + size], bytes(b),
                         tol)
        worst = max(worst, w)
        for x in bad:
            mism.append((hex(base), size) + x)
    if verbose:
        status = "ok" if not mem.violations and not mism else "FAIL"
        print("%-28s %-6s grid %-14s accesses-outside %d  mismatches %d  "
              "worst fp64 rel %.1e" % (os.path.basename(stem), status,
                                      str(tuple(meta["grid"])), len(mem.violations),
                                      len(mism), worst), flush=True)
        if os.environ.get("BB_ISA_COUNT") and ex.wave_instructions:
            print("    wave instructions: longest %d, all waves %d; global "
                  "atomic lane operations %d" % (
                      max(ex.wave_instructions), sum(ex.wave_instructions),
                      getattr(ex, "atomic_lanes", 0)), flush=True)
        for v in mem.violations[:5]:
            print("    outside:", v, flush=True)
        for x in mism[:5]:
            print("    mismatch:", x, flush=True)
    return mem.violations, mism


BUNDLER = "/opt/rocm/llvm/bin/clang-offload-bundler"
OBJCOPY = "/opt/rocm/llvm/bin/llvm-objcopy"
READELF = "/opt/rocm/llvm/bin/llvm-readelf"
OBJ_DIR = os.path.join(ROOT, "bluebottle_amd", "lib")


def extract_code_objects(out_dir, obj_dir=OBJ_DIR, arch="gfx950"):
    """Unbundle the gfx950 code object of every object file of the HIP
    library build into out_dir (one <name>.co per .o)."""
    os.makedirs(out_dir, exist_ok=True)
    for o in sorted(glob.glob(os.path.join(obj_dir, "bb_*.o"))):
        stem = os.path.join(out_dir, os.path.basename(o)[:-2])
        secs = subprocess.run([READELF, "-S", "-W", o], capture_output=True,
                              text=True, check=True).stdout
        if ".hip_fatbin" not in secs:
            continue   # an object with no kernels
        # the host object carries the offload bundle in its .hip_fatbin section
        subprocess.run([OBJCOPY, "--dump-section", ".hip_fatbin=" + stem + ".fatbin",
                        o, stem + ".tmp.o"], check=True)
        os.remove(stem + ".tmp.o")
        subprocess.run([BUNDLER, "--unbundle", "--type=o",
                        "--input=" + stem + ".fatbin",
                        "--targets=hipv4-amdgcn-amd-amdhsa--" + arch,
                        "--output=" + stem + ".co"], check=True)
        os.remove(stem + ".fatbin")
    return out_dir


def main():
    cap = sys.argv[1]
    filt = sys.argv[2:]
    objdir = os.environ.get("BB_ISA_OBJDIR")
    if not objdir:
        objdir = extract_code_objects(os.path.join(ROOT, "tests", "isa", "_co"))
    lib = Library(objdir)
    stems = sorted(p[:-5] for p in glob.glob(os.path.join(cap, "*.json")))
    fails = 0
    for st in stems:
        if filt and not any(f in os.path.basename(st) for f in filt):
            continue
        v, m = replay_one(lib, st)
        fails += bool(v or m)
    print("replayed %d launches, %d with findings" % (len(stems), fails))
    return 1 if fails else 0


if __name__ == "__main__":
    sys.exit(main())
