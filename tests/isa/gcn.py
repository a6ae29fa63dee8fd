"""Wave64 interpreter for the gfx950 machine code of this library's kernels
(test infrastructure).

It executes the instructions `llvm-objdump -d --mcpu=gfx950` prints for the
code objects built from bluebottle_amd/csrc, one workgroup at a time, its
waves in turn with real barriers, over a memory image of the device
allocations.  Every global and scalar memory access of every active lane is
checked against the allocations and the kernarg segment before it is made;
an access outside them is recorded with the kernel, the instruction and the
lane instead of being performed.  Values follow IEEE double/single arithmetic
(v_rcp_f64 is exact and the division scale/fixup steps pass values through,
so quotients come out correctly rounded); integer and bit operations are
exact.

Only the instructions the library's kernels use are implemented; anything
else raises NotImplementedError naming the opcode.
"""
import math
import re
import struct

import numpy as np

M32 = 0xFFFFFFFF
M64 = (1 << 64) - 1
LANES = 64
LANE = np.arange(LANES, dtype=np.uint64)
FULL = (1 << 64) - 1

INLINE_F = {"0.5": 0.5, "-0.5": -0.5, "1.0": 1.0, "-1.0": -1.0, "2.0": 2.0,
            "-2.0": -2.0, "4.0": 4.0, "-4.0": -4.0,
            "0.15915494": 0.15915494309189532}


class Violation(Exception):
    pass


def mask_to_bool(m):
    return ((np.uint64(m) >> LANE) & np.uint64(1)).astype(bool)


def bool_to_mask(b):
    return int(np.sum(np.left_shift(np.uint64(1), LANE[b]), dtype=np.uint64)) \
        if np.any(b) else 0


def f64_from_u64(u):
    return np.asarray(u, dtype=np.uint64).view(np.float64)


def u64_from_f64(f):
    return np.asarray(f, dtype=np.float64).view(np.uint64)


def f32_from_u32(u):
    return np.asarray(u, dtype=np.uint32).view(np.float32)


def u32_from_f32(f):
    return np.asarray(f, dtype=np.float32).view(np.uint32)


def two_prod_fma(a, b, c):
    """fma(a, b, c) for float64 arrays, rounded once up to the last bit in
    all but pathological cases (Dekker product plus compensated sum)."""
    with np.errstate(all="ignore"):
        p = a * b
        split = 134217729.0
        t = split * a
        ah = t - (t - a)
        al = a - ah
        t = split * b
        bh = t - (t - b)
        bl = b - bh
        e = ((ah * bh - p) + ah * bl + al * bh) + al * bl
        s = p + c
        bb = s - p
        err = (p - (s - bb)) + (c - bb)
        r = s + (err + e)
        bad = ~np.isfinite(p) | ~np.isfinite(r)
        if np.any(bad):
            r = np.where(bad, a * b + c, r)
    return r


# ---------------------------------------------------------------------------
# code object parsing
# ---------------------------------------------------------------------------
class Kernel:
    def __init__(self, name):
        self.name = name
        self.insts = []      # (addr, op, operands, text)
        self.index = {}      # addr -> instruction index
        self.args = []       # (offset, size, kind)
        self.kernarg_size = 0
        self.lds_size = 0
        self.private_size = 0
        self.wg_id = (1, 0, 0)
        self.user_sgprs = 2
        self.dispatch_ptr = False


_INST = re.compile(r"^\s+([a-z_0-9]+)(.*?)(?:\s*//\s*([0-9A-F]+):.*)?$")
_TARGET = re.compile(r"<[^+>]+\+0x([0-9a-f]+)>")
_KSTART = re.compile(r"^([0-9a-f]+) <(\S+)>:")


def load_code_object(path, objdump="/opt/rocm/llvm/bin/llvm-objdump",
                     readelf="/opt/rocm/llvm/bin/llvm-readelf"):
    import subprocess
    dis = subprocess.run([objdump, "-d", "--mcpu=gfx950", path],
                         capture_output=True, text=True, check=True).stdout
    kernels = {}
    cur = None
    base = 0
    for line in dis.splitlines():
        m = _KSTART.match(line)
        if m:
            cur = Kernel(m.group(2))
            base = int(m.group(1), 16)
            kernels[cur.name] = cur
            continue
        if cur is None:
            continue
        m = _INST.match(line)
        if not m or not m.group(3):
            continue
        op = m.group(1)
        rest = m.group(2).strip()
        addr = int(m.group(3), 16)
        tgt = None
        t = _TARGET.search(line)
        if t and (op.startswith("s_cbranch") or op == "s_branch"):
            tgt = base + int(t.group(1), 16)
        cur.index[addr] = len(cur.insts)
        cur.insts.append((addr, op, rest, tgt))
    notes = subprocess.run([readelf, "--notes", path], capture_output=True,
                           text=True, check=True).stdout
    for block in notes.split("  - .agpr_count")[1:]:
        name = re.search(r"\.name:\s+(\S+)", block).group(1)
        if name not in kernels:
            continue
        k = kernels[name]
        k.args = [(int(o), int(s), kd) for o, s, kd in re.findall(
            r"\.offset:\s+(\d+)\s+\.size:\s+(\d+)\s+\.value_kind:\s+(\w+)", block)]
        k.kernarg_size = int(re.search(r"\.kernarg_segment_size:\s+(\d+)",
                                       block).group(1))
        k.lds_size = int(re.search(r"\.group_segment_fixed_size:\s+(\d+)",
                                   block).group(1))
        k.private_size = int(re.search(r"\.private_segment_fixed_size:\s+(\d+)",
                                       block).group(1))
    kd = subprocess.run([objdump, "-d", "--mcpu=gfx950", "--section=.rodata",
                         path], capture_output=True, text=True).stdout
    for chunk in kd.split(".amdhsa_kernel ")[1:]:
        name = chunk.split()[0]
        if name not in kernels:
            continue
        k = kernels[name]

        def field(f):
            m = re.search(r"\.amdhsa_%s (\d+)" % f, chunk)
            return int(m.group(1)) if m else 0
        k.wg_id = (field("system_sgpr_workgroup_id_x"),
                   field("system_sgpr_workgroup_id_y"),
                   field("system_sgpr_workgroup_id_z"))
        for f in ("user_sgpr_queue_ptr", "user_sgpr_dispatch_id",
                  "user_sgpr_private_segment_size"):
            if field(f):
                raise NotImplementedError("kernel %s uses %s" % (name, f))
        # user SGPRs in hardware order: dispatch packet pointer, then the
        # kernarg segment pointer
        k.dispatch_ptr = bool(field("user_sgpr_dispatch_ptr"))
        k.user_sgprs = (2 if k.dispatch_ptr else 0) + \
            (2 if field("user_sgpr_kernarg_segment_ptr") else 0)
    return kernels


# ---------------------------------------------------------------------------
# memory
# ---------------------------------------------------------------------------
class Memory:
    """Device allocations (and the kernarg segment) by base address."""

    def __init__(self):
        self.bases = []
        self.bufs = []
        self.violations = []

    def add(self, base, data):
        self.bases.append(base)
        self.bufs.append(data)
        order = np.argsort(self.bases)
        self.bases = [self.bases[i] for i in order]
        self.bufs = [self.bufs[i] for i in order]

    def find(self, addr, n):
        import bisect
        i = bisect.bisect_right(self.bases, addr) - 1
        if i < 0:
            return None, 0
        off = addr - self.bases[i]
        if off + n <= len(self.bufs[i]):
            return self.bufs[i], off
        return None, 0

    def read(self, addr, n, where):
        buf, off = self.find(addr, n)
        if buf is None:
            self.violations.append((where, "read", addr, n))
            return bytes(n)
        return bytes(buf[off:off + n])

    def write(self, addr, data, where):
        buf, off = self.find(addr, len(data))
        if buf is None:
            self.violations.append((where, "write", addr, len(data)))
            return
        buf[off:off + len(data)] = data


# ---------------------------------------------------------------------------
# wave state and operand access
# ---------------------------------------------------------------------------
_REG = re.compile(r"^([vsa])(?:\[(\d+):(\d+)\]|(\d+))$")


POISON = 0xBAD0BAD1
DISPATCH_BASE = 0x800


def _two_over_pi_bits(nbits):
    """floor(2/pi * 2**nbits) as an integer (pi by Machin's formula)."""
    guard = 64
    one = 1 << (nbits + guard)

    def arctan_inv(x):
        t = one // x
        s = t
        x2 = x * x
        n, sign = 3, -1
        while t:
            t //= x2
            s += sign * (t // n)
            sign, n = -sign, n + 2
        return s
    pi = 4 * (4 * arctan_inv(5) - arctan_inv(239))   # pi * one
    return (2 * one * (1 << nbits)) // pi


# v_trig_preop_f64 reads 53-bit segments of the first 1201 bits of 2/pi
TWO_OVER_PI_NBITS = 1201 + 53 + 128
TWO_OVER_PI = _two_over_pi_bits(TWO_OVER_PI_NBITS)


class Wave:
    def __init__(self, nsgpr=112, nvgpr=512):
        # registers the dispatch does not initialise hold a poison value, so
        # that an address formed from one falls outside every allocation;
        # `sdef` records which SGPRs have been written (dispatch or code)
        self.s = [POISON] * nsgpr
        self.sdef = [False] * nsgpr
        self.v = np.full((nvgpr, LANES), POISON, dtype=np.uint32)
        self.exec = 0
        self.vcc = 0
        self.scc = 0
        self.m0 = 0
        self.pc = 0
        self.done = False
        self.at_barrier = False
        self.priv = None


def split_ops(rest):
    """Operand list of an instruction: top-level comma separation, then the
    trailing modifiers (offset:N, bitop3:0x.., ...)."""
    mods = {}
    parts = []
    depth = 0
    cur = ""
    for ch in rest:
        if ch in "[(":
            depth += 1
        elif ch in "])":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    if cur.strip():
        parts.append(cur.strip())
    # modifiers follow the last operand separated by spaces
    if parts:
        tail = parts[-1].split()
        parts[-1] = tail[0] if tail else ""
        for t in tail[1:]:
            if ":" in t:
                k, v = t.split(":", 1)
                mods[k] = v
            else:
                mods[t] = True
    return [p for p in parts if p != ""], mods


class Exec:
    """Runs one kernel launch over a memory image."""

    def __init__(self, kernel, mem, kernarg_addr, grid, block, log=None):
        self.k = kernel
        self.mem = mem
        self.ka = kernarg_addr
        self.grid = grid
        self.block = block
        self.cache = {}
        self.log = log
        # instructions each finished wavefront issued (a latency proxy: the
        # longest is the launch's critical path in issue slots)
        self.wave_instructions = []

    # ---- operand helpers ----
    def sreg(self, w, name, n=1):
        """Scalar operand as a Python int of 32*n bits."""
        if name in ("vcc",):
            return w.vcc & (M64 if n == 2 else M32)
        if name == "vcc_lo":
            return w.vcc & M32
        if name == "vcc_hi":
            return (w.vcc >> 32) & M32
        if name == "exec":
            return w.exec & (M64 if n == 2 else M32)
        if name == "exec_lo":
            return w.exec & M32
        if name == "exec_hi":
            return (w.exec >> 32) & M32
        if name == "m0":
            return w.m0
        if name == "scc":
            return w.scc
        m = _REG.match(name)
        if m and m.group(1) == "s":
            lo = int(m.group(2) if m.group(2) else m.group(4))
            hi = int(m.group(3)) if m.group(3) else lo
            val = 0
            for i, r in enumerate(range(lo, hi + 1)):
                val |= (w.s[r] & M32) << (32 * i)
            return val
        return None

    def lit(self, tok, n=1, fp=False):
        """Inline constant or literal as raw bits (32*n)."""
        if tok in INLINE_F:
            f = INLINE_F[tok]
            if n == 2:
                return struct.unpack("<Q", struct.pack("<d", f))[0]
            return struct.unpack("<I", struct.pack("<f", f))[0]
        v = int(tok, 0)
        if n == 2:
            if tok.startswith("0x") and fp:
                return (v & M32) << 32
            if -16 <= v <= 64 and not tok.startswith("0x"):
                return v & M64
            return (v if v >= 0 else v + (1 << 64)) & M64 if not tok.startswith("0x") \
                else (v & M32)
        return v & M32

    def get(self, w, tok, n=1, fp=False):
        """Operand as per-lane numpy array: uint32 (n=1) or uint64 (n=2);
        float modifiers applied for fp operands (f64 when n=2, f32 when n=1)."""
        neg = absv = False
        t = tok
        if t.startswith("-") and (t[1:2] in "vs|a" or t[1:4] == "vcc") and t[1:2] != "":
            if not re.match(r"^-\d", t):
                neg = True
                t = t[1:]
        if t.startswith("|") and t.endswith("|"):
            absv = True
            t = t[1:-1]
        if t.startswith("abs(") and t.endswith(")"):
            absv = True
            t = t[4:-1]
        if t.startswith("neg(") and t.endswith(")"):
            neg = True
            t = t[4:-1]
        m = _REG.match(t)
        if m and m.group(1) == "v":
            lo = int(m.group(2) if m.group(2) else m.group(4))
            if n == 2:
                val = w.v[lo].astype(np.uint64) | (w.v[lo + 1].astype(np.uint64) << np.uint64(32))
            else:
                val = w.v[lo].copy()
        else:
            sv = self.sreg(w, t, n)
            if sv is None:
                sv = self.lit(t, n, fp)
            if n == 2:
                val = np.full(LANES, sv & M64, dtype=np.uint64)
            else:
                val = np.full(LANES, sv & M32, dtype=np.uint32)
        if fp and (neg or absv):
            if n == 2:
                sign = np.uint64(1 << 63)
                if absv:
                    val = val & ~sign
                if neg:
                    val = val ^ sign
            else:
                sign = np.uint32(1 << 31)
                if absv:
                    val = val & ~sign
                if neg:
                    val = val ^ sign
        elif neg:
            val = (-val.astype(np.int64)).astype(val.dtype) if n == 1 else \
                (np.uint64(0) - val)
        return val

    def fget(self, w, tok):
        return f64_from_u64(self.get(w, tok, 2, True))

    def f32get(self, w, tok):
        return f32_from_u32(self.get(w, tok, 1, True))

    def vset(self, w, tok, val, n=1, active=None):
        m = _REG.match(tok)
        if not m or m.group(1) != "v":
            raise NotImplementedError("vector destination " + tok)
        lo = int(m.group(2) if m.group(2) else m.group(4))
        act = mask_to_bool(w.exec) if active is None else active
        if n == 2:
            val = np.asarray(val, dtype=np.uint64)
            w.v[lo][act] = (val & np.uint64(M32)).astype(np.uint32)[act]
            w.v[lo + 1][act] = (val >> np.uint64(32)).astype(np.uint32)[act]
        else:
            val = np.asarray(val, dtype=np.uint32)
            w.v[lo][act] = val[act]

    def sset(self, w, tok, val, n=1):
        if tok == "vcc":
            w.vcc = val & M64 if n == 2 else (w.vcc & ~M32) | (val & M32)
            return
        if tok == "vcc_lo":
            w.vcc = (w.vcc & ~M32) | (val & M32)
            return
        if tok == "vcc_hi":
            w.vcc = (w.vcc & M32) | ((val & M32) << 32)
            return
        if tok == "exec":
            w.exec = val & M64 if n == 2 else (w.exec & ~M32) | (val & M32)
            return
        if tok == "exec_lo":
            w.exec = (w.exec & ~M32) | (val & M32)
            return
        if tok == "exec_hi":
            w.exec = (w.exec & M32) | ((val & M32) << 32)
            return
        if tok == "m0":
            w.m0 = val & M32
            return
        m = _REG.match(tok)
        if not m or m.group(1) != "s":
            raise NotImplementedError("scalar destination " + tok)
        lo = int(m.group(2) if m.group(2) else m.group(4))
        hi = int(m.group(3)) if m.group(3) else lo
        for i, r in enumerate(range(lo, hi + 1)):
            w.s[r] = (val >> (32 * i)) & M32
            w.sdef[r] = True

    def check_sdef(self, w, tok, where, what):
        """Record a violation when SGPR operand `tok` was never written."""
        m = _REG.match(tok)
        if not m or m.group(1) != "s":
            return
        lo = int(m.group(2) if m.group(2) else m.group(4))
        hi = int(m.group(3)) if m.group(3) else lo
        for r in range(lo, hi + 1):
            if not w.sdef[r]:
                self.mem.violations.append((where, "undefined-sgpr-" + what, r, 0))
                return

    def width(self, tok):
        if tok in ("vcc", "exec"):
            return 2
        m = _REG.match(tok.strip("-|"))
        if m and m.group(2):
            return int(m.group(3)) - int(m.group(2)) + 1
        return 1

    # ---- launch ----
    def run(self):
        k = self.k
        if k.dispatch_ptr and not self.mem.find(DISPATCH_BASE, 64)[0]:
            # an hsa_kernel_dispatch_packet_t with the launch shape
            pk = bytearray(64)
            bx, by, bz = self.block
            gx, gy, gz = self.grid
            pk[4:10] = struct.pack("<HHH", bx, by, bz)
            pk[12:24] = struct.pack("<III", gx * bx, gy * by, gz * bz)
            pk[28:32] = struct.pack("<I", k.lds_size)
            pk[40:48] = struct.pack("<Q", self.ka)
            self.mem.add(DISPATCH_BASE, pk)
        bx, by, bz = self.block
        nth = bx * by * bz
        nwaves = (nth + 63) // 64
        gx, gy, gz = self.grid
        for wz in range(gz):
            for wy in range(gy):
                for wx in range(gx):
                    self.run_group((wx, wy, wz), nth, nwaves)

    def run_group(self, gid, nth, nwaves):
        k = self.k
        bx, by, bz = self.block
        lds = bytearray(max(k.lds_size, 0) + 0)
        self.lds = lds
        waves = []
        for wi in range(nwaves):
            w = Wave()
            t = wi * 64 + np.arange(LANES)
            act = t < nth
            w.exec = bool_to_mask(act)
            tx = t % bx
            ty = (t // bx) % by
            tz = t // (bx * by)
            w.v[0] = (tx | (ty << 10) | (tz << 20)).astype(np.uint32)
            ka_reg = 0
            if k.dispatch_ptr:
                w.s[0] = DISPATCH_BASE & M32
                w.s[1] = (DISPATCH_BASE >> 32) & M32
                w.sdef[0] = w.sdef[1] = True
                ka_reg = 2
            w.s[ka_reg] = self.ka & M32
            w.s[ka_reg + 1] = (self.ka >> 32) & M32
            w.sdef[ka_reg] = w.sdef[ka_reg + 1] = True
            r = k.user_sgprs
            for d in range(3):
                if k.wg_id[d]:
                    w.s[r] = gid[d]
                    w.sdef[r] = True
                    r += 1
            w.pc = 0
            w.wid = wi
            waves.append(w)
        while True:
            progressed = False
            for w in waves:
                if w.done or w.at_barrier:
                    continue
                self.step_until_barrier(w, gid)
                progressed = True
            live = [w for w in waves if not w.done]
            if not live:
                break
            if all(w.at_barrier for w in live):
                for w in live:
                    w.at_barrier = False
                continue
            if not progressed:
                raise RuntimeError("%s: waves deadlocked at a barrier" % k.name)

    def step_until_barrier(self, w, gid):
        k = self.k
        steps = 0
        while not w.done:
            addr, op, rest, tgt = k.insts[w.pc]
            w.pc += 1
            steps += 1
            w.issued = getattr(w, "issued", 0) + 1
            if steps > 50_000_000:
                raise RuntimeError("%s: runaway loop" % k.name)
            if op == "s_barrier":
                w.at_barrier = True
                return
            self.execute(w, op, rest, tgt, (k.name, addr, gid))
            if w.done:
                self.wave_instructions.append(w.issued)

    def branch(self, w, tgt):
        w.pc = self.k.index[tgt]

    # ---- instruction semantics ----
    def execute(self, w, op, rest, tgt, where):
        if op in ("s_waitcnt", "s_nop", "s_setprio", "s_sleep",
                  # L2 write-back / invalidate of a device-scope fence: the
                  # replay has one memory and no caches
                  "buffer_wbl2", "buffer_inv"):
            return
        if op == "s_endpgm":
            w.done = True
            return
        ops, mods = self.cache.get(rest) or split_ops(rest)
        self.cache[rest] = (ops, mods)
        if op.startswith("s_"):
            return self.salu(w, op, ops, mods, tgt, where)
        if op.startswith("v_"):
            return self.valu(w, op, ops, mods, where)
        if op.startswith("global_"):
            return self.gmem(w, op, ops, mods, where)
        if op.startswith("ds_"):
            return self.dsmem(w, op, ops, mods, where)
        if op.startswith("scratch_"):
            return self.pmem(w, op, ops, mods, where)
        raise NotImplementedError(op)

    def salu(self, w, op, ops, mods, tgt, where):
        if op == "s_branch":
            return self.branch(w, tgt)
        if op.startswith("s_cbranch_"):
            c = op[10:]
            take = {"scc0": w.scc == 0, "scc1": w.scc == 1,
                    "vccz": w.vcc == 0, "vccnz": w.vcc != 0,
                    "execz": w.exec == 0, "execnz": w.exec != 0}[c]
            if take:
                self.branch(w, tgt)
            return
        if op.startswith("s_load_dword"):
            n = {"s_load_dword": 1, "s_load_dwordx2": 2, "s_load_dwordx4": 4,
                 "s_load_dwordx8": 8, "s_load_dwordx16": 16}[op]
            self.check_sdef(w, ops[1], where, "base")
            base = self.sreg(w, ops[1], 2)
            off = ops[2]
            o = self.sreg(w, off, 1)
            if o is None:
                o = int(off, 0)
                if o >= 1 << 20:
                    o -= 1 << 21
            data = self.mem.read((base + o) & M64, 4 * n, where)
            self.sset(w, ops[0], int.from_bytes(data, "little"), n)
            return
        if op in ("s_mov_b32", "s_mov_b64", "s_movk_i32"):
            n = 2 if op.endswith("b64") else 1
            if op == "s_movk_i32":
                v = int(ops[1], 0) & 0xFFFF
                v = v - 0x10000 if v & 0x8000 else v
                return self.sset(w, ops[0], v & M32)
            v = self.sreg(w, ops[1], n)
            if v is None:
                v = self.lit(ops[1], n)
            return self.sset(w, ops[0], v, n)
        if op in ("s_cselect_b32", "s_cselect_b64"):
            n = 2 if op.endswith("b64") else 1
            a = self.sv(w, ops[1], n)
            b = self.sv(w, ops[2], n)
            return self.sset(w, ops[0], a if w.scc else b, n)
        if op.endswith("_saveexec_b64"):
            src = self.sv(w, ops[1], 2)
            old = w.exec
            if op == "s_and_saveexec_b64":
                new = src & old
            elif op == "s_or_saveexec_b64":
                new = src | old
            elif op == "s_andn2_saveexec_b64":
                new = src & (~old & M64)
            else:
                raise NotImplementedError(op)
            self.sset(w, ops[0], old, 2)
            w.exec = new & M64
            w.scc = int(w.exec != 0)
            return
        if op in ("s_bitcmp0_b32", "s_bitcmp1_b32"):
            a = self.sv(w, ops[0])
            bit = (a >> (self.sv(w, ops[1]) & 31)) & 1
            w.scc = int(bit == (0 if op == "s_bitcmp0_b32" else 1))
            return
        if op.startswith("s_cmp"):
            return self.scmp(w, op, ops)
        return self.salu_arith(w, op, ops, mods)

    def sv(self, w, tok, n=1):
        v = self.sreg(w, tok, n)
        if v is None:
            v = self.lit(tok, n)
        return v

    def scmp(self, w, op, ops):
        m = re.match(r"s_cmpk?_(eq|lg|gt|ge|lt|le)_(i32|u32|u64)", op)
        if not m:
            raise NotImplementedError(op)
        c, t = m.groups()
        n = 2 if t == "u64" else 1
        a = self.sv(w, ops[0], n)
        b = self.sv(w, ops[1], n)
        if t == "i32":
            a = a - (1 << 32) if a & (1 << 31) else a
            b = b - (1 << 32) if b & (1 << 31) else b
        w.scc = int({"eq": a == b, "lg": a != b, "gt": a > b, "ge": a >= b,
                     "lt": a < b, "le": a <= b}[c])

    def salu_arith(self, w, op, ops, mods):
        def s32(x):
            x &= M32
            return x - (1 << 32) if x & (1 << 31) else x

        def s64(x):
            x &= M64
            return x - (1 << 64) if x & (1 << 63) else x
        if op in ("s_add_u32", "s_addc_u32"):
            a, b = self.sv(w, ops[1]), self.sv(w, ops[2])
            r = a + b + (w.scc if op == "s_addc_u32" else 0)
            w.scc = int(r > M32)
            return self.sset(w, ops[0], r & M32)
        if op in ("s_sub_u32", "s_subb_u32"):
            a, b = self.sv(w, ops[1]), self.sv(w, ops[2])
            r = a - b - (w.scc if op == "s_subb_u32" else 0)
            w.scc = int(r < 0)
            return self.sset(w, ops[0], r & M32)
        if op in ("s_add_i32", "s_sub_i32", "s_addk_i32"):
            a = s32(self.sv(w, ops[0] if op == "s_addk_i32" else ops[1]))
            if op == "s_addk_i32":
                b = int(ops[1], 0) & 0xFFFF
                b = b - 0x10000 if b & 0x8000 else b
            else:
                b = s32(self.sv(w, ops[2]))
            r = a + b if op != "s_sub_i32" else a - b
            w.scc = int(r > 0x7FFFFFFF or r < -0x80000000)
            return self.sset(w, ops[0], r & M32)
        if op == "s_mulk_i32":
            b = int(ops[1], 0) & 0xFFFF
            b = b - 0x10000 if b & 0x8000 else b
            r = s32(self.sv(w, ops[0])) * b
            return self.sset(w, ops[0], r & M32)
        if op == "s_mul_i32":
            r = s32(self.sv(w, ops[1])) * s32(self.sv(w, ops[2]))
            return self.sset(w, ops[0], r & M32)
        if op == "s_mul_hi_u32":
            r = self.sv(w, ops[1]) * self.sv(w, ops[2])
            return self.sset(w, ops[0], (r >> 32) & M32)
        if op == "s_mul_hi_i32":
            r = s32(self.sv(w, ops[1])) * s32(self.sv(w, ops[2]))
            return self.sset(w, ops[0], (r >> 32) & M32)
        m = re.match(r"s_(and|or|xor|andn2|orn2|nor|nand|xnor)_b(32|64)$", op)
        if m:
            n = 2 if m.group(2) == "64" else 1
            msk = M64 if n == 2 else M32
            a, b = self.sv(w, ops[1], n), self.sv(w, ops[2], n)
            f = m.group(1)
            r = {"and": a & b, "or": a | b, "xor": a ^ b, "andn2": a & ~b,
                 "orn2": a | ~b, "nor": ~(a | b), "nand": ~(a & b),
                 "xnor": ~(a ^ b)}[f] & msk
            w.scc = int(r != 0)
            return self.sset(w, ops[0], r, n)
        if op in ("s_not_b32", "s_not_b64"):
            n = 2 if op.endswith("64") else 1
            r = ~self.sv(w, ops[1], n) & (M64 if n == 2 else M32)
            w.scc = int(r != 0)
            return self.sset(w, ops[0], r, n)
        m = re.match(r"s_(lshl|lshr|ashr)_(b32|b64|i32|i64)$", op)
        if m:
            n = 2 if m.group(2).endswith("64") else 1
            bits = 32 * n
            a = self.sv(w, ops[1], n)
            sh = self.sv(w, ops[2]) & (bits - 1)
            if m.group(1) == "lshl":
                r = (a << sh) & ((1 << bits) - 1)
            elif m.group(1) == "lshr":
                r = a >> sh
            else:
                r = ((s64(a) if n == 2 else s32(a)) >> sh) & ((1 << bits) - 1)
            w.scc = int(r != 0)
            return self.sset(w, ops[0], r, n)
        if op == "s_abs_i32":
            r = abs(s32(self.sv(w, ops[1]))) & M32
            w.scc = int(r != 0)
            return self.sset(w, ops[0], r)
        if op in ("s_min_i32", "s_max_i32", "s_min_u32", "s_max_u32"):
            a, b = self.sv(w, ops[1]), self.sv(w, ops[2])
            if op.endswith("i32"):
                a, b = s32(a), s32(b)
            if op.startswith("s_min"):
                w.scc = int(a < b)
                r = a if a < b else b
            else:
                w.scc = int(a > b)
                r = a if a > b else b
            return self.sset(w, ops[0], r & M32)
        if op == "s_bfe_u32":
            a, b = self.sv(w, ops[1]), self.sv(w, ops[2])
            off = b & 31
            width = (b >> 16) & 0x7F
            r = (a >> off) & ((1 << width) - 1) if width else 0
            w.scc = int(r != 0)
            return self.sset(w, ops[0], r)
        if op == "s_brev_b32":
            a = self.sv(w, ops[1])
            r = int("{:032b}".format(a)[::-1], 2)
            return self.sset(w, ops[0], r)
        if op == "s_bcnt1_i32_b64":
            r = bin(self.sv(w, ops[1], 2)).count("1")
            w.scc = int(r != 0)
            return self.sset(w, ops[0], r)
        if op in ("s_ff1_i32_b64", "s_ff1_i32_b32"):
            n = 2 if op.endswith("64") else 1
            a = self.sv(w, ops[1], n)
            r = (a & -a).bit_length() - 1 if a else M32
            return self.sset(w, ops[0], r & M32)
        raise NotImplementedError(op)

    # ---- vector ALU ----
    def sdwa_sel(self, v, sel, sext):
        shift, bits = {"DWORD": (0, 32), "BYTE_0": (0, 8), "BYTE_1": (8, 8),
                       "BYTE_2": (16, 8), "BYTE_3": (24, 8), "WORD_0": (0, 16),
                       "WORD_1": (16, 16)}[sel]
        x = (v.astype(np.uint64) >> np.uint64(shift)) & np.uint64((1 << bits) - 1)
        if sext and bits < 32:
            x = np.where(x >> np.uint64(bits - 1) & np.uint64(1),
                         x | np.uint64(M32 & ~((1 << bits) - 1)), x)
        return (x & np.uint64(M32)).astype(np.uint32)

    def sdwa(self, w, op, ops, mods, act):
        """SDWA form of a two-source 32-bit integer op: select the source
        fields, run the op, place the result by dst_sel."""
        base = op[:-5]
        srcs = []
        for i, key in ((1, "src0_sel"), (2, "src1_sel")):
            tok = ops[i]
            sext = tok.startswith("sext(")
            if sext:
                tok = tok[5:-1]
            srcs.append(self.sdwa_sel(self._u(w, tok), mods.get(key, "DWORD"), sext))
        a, b = srcs
        fn = {"v_and_b32": lambda: a & b, "v_or_b32": lambda: a | b,
              "v_xor_b32": lambda: a ^ b, "v_add_u32": lambda: a + b,
              "v_sub_u32": lambda: a - b,
              "v_lshlrev_b32": lambda: b << (a & np.uint32(31)),
              "v_lshrrev_b32": lambda: b >> (a & np.uint32(31)),
              "v_mul_u32_u24": lambda: ((a & np.uint32(0xFFFFFF)).astype(np.uint64)
                                         * (b & np.uint32(0xFFFFFF)).astype(np.uint64)
                                         ).astype(np.uint32)}.get(base)
        if fn is None:
            raise NotImplementedError(op)
        r = fn().astype(np.uint32)
        dsel = mods.get("dst_sel", "DWORD")
        if dsel != "DWORD":
            shift, bits = {"BYTE_0": (0, 8), "BYTE_1": (8, 8), "BYTE_2": (16, 8),
                           "BYTE_3": (24, 8), "WORD_0": (0, 16), "WORD_1": (16, 16)}[dsel]
            fieldm = np.uint32(((1 << bits) - 1) << shift)
            unused = mods.get("dst_unused", "UNUSED_PAD")
            old = self._u(w, ops[0]) if unused == "UNUSED_PRESERVE" else np.zeros(LANES, np.uint32)
            r = (old & ~fieldm) | ((r << np.uint32(shift)) & fieldm)
        self.vset(w, ops[0], r)

    def valu(self, w, op, ops, mods, where):
        act = mask_to_bool(w.exec)
        if op.endswith("_sdwa"):
            return self.sdwa(w, op, ops, mods, act)
        base = re.sub(r"_e(32|64)$", "", op)
        if base.startswith("v_cmp_") or base.startswith("v_cmpx_"):
            return self.vcmp(w, base, ops, act)
        f = getattr(self, "op_" + base[2:], None)
        if f is None:
            raise NotImplementedError(op)
        return f(w, ops, mods, act)

    def fclass(self, x, mask_bits):
        """v_cmp_class: bit 0 sNaN, 1 qNaN, 2 -inf, 3 -normal, 4 -denormal,
        5 -0, 6 +0, 7 +denormal, 8 +normal, 9 +inf."""
        tiny = np.finfo(x.dtype).tiny
        neg = np.signbit(x)
        nan = np.isnan(x)
        inf = np.isinf(x)
        zero = x == 0
        den = ~nan & ~inf & ~zero & (np.abs(x) < tiny)
        nor = ~nan & ~inf & ~zero & ~den
        bits = u64_from_f64(x) if x.dtype == np.float64 else u32_from_f32(x).astype(np.uint64)
        qbit = np.uint64(1 << 51) if x.dtype == np.float64 else np.uint64(1 << 22)
        quiet = (bits & qbit) != 0
        cls = [nan & ~quiet, nan & quiet, inf & neg, nor & neg, den & neg, zero & neg,
               zero & ~neg, den & ~neg, nor & ~neg, inf & ~neg]
        r = np.zeros(LANES, dtype=bool)
        for i, c in enumerate(cls):
            r |= c & (((mask_bits >> np.uint32(i)) & np.uint32(1)) == 1)
        return r

    def vcmp(self, w, base, ops, act):
        m = re.match(r"v_cmpx?_class_(f64|f32)$", base)
        if m:
            x = self.fget(w, ops[1]) if m.group(1) == "f64" else self.f32get(w, ops[1])
            r = self.fclass(x, self._u(w, ops[2]))
            res = bool_to_mask(r & act)
            if base.startswith("v_cmpx"):
                w.exec = res
            if ops[0] == "vcc":
                w.vcc = res
            else:
                self.sset(w, ops[0], res, 2)
            return
        m = re.match(r"v_cmpx?_([a-z]+)_(i32|u32|i64|u64|f64|f32)$", base)
        if not m:
            raise NotImplementedError(base)
        c, t = m.groups()
        dst = ops[0]
        a_tok, b_tok = ops[1], ops[2]
        if t in ("f64", "f32"):
            if t == "f64":
                a, b = self.fget(w, a_tok), self.fget(w, b_tok)
            else:
                a, b = self.f32get(w, a_tok), self.f32get(w, b_tok)
            with np.errstate(invalid="ignore"):
                r = {"eq": a == b, "lt": a < b, "le": a <= b, "gt": a > b,
                     "ge": a >= b, "lg": (a < b) | (a > b),
                     "neq": ~(a == b), "nlt": ~(a < b), "nle": ~(a <= b),
                     "ngt": ~(a > b), "nge": ~(a >= b), "nlg": ~((a < b) | (a > b)),
                     "o": ~(np.isnan(a) | np.isnan(b)),
                     "u": np.isnan(a) | np.isnan(b)}[c]
        else:
            n = 2 if t.endswith("64") else 1
            a, b = self.get(w, a_tok, n), self.get(w, b_tok, n)
            if t.startswith("i"):
                a = a.astype(np.int64 if n == 2 else np.int32)
                b = b.astype(np.int64 if n == 2 else np.int32)
            r = {"eq": a == b, "ne": a != b, "lt": a < b, "le": a <= b,
                 "gt": a > b, "ge": a >= b}[c]
        res = bool_to_mask(r & act)
        if base.startswith("v_cmpx"):
            w.exec = res
        if dst == "vcc" or base.endswith("_e32"):
            w.vcc = res
        else:
            self.sset(w, dst, res, 2)

    def _u(self, w, t, n=1):
        return self.get(w, t, n)

    def op_mov_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]))

    def op_mov_b64(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1], 2), 2)

    def op_add_u32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) + self._u(w, o[2]))

    def op_sub_u32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) - self._u(w, o[2]))

    def op_subrev_u32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[2]) - self._u(w, o[1]))

    def _carry_ops(self, w, o, kind, act):
        # v_add_co_u32 d, sdst, a, b [, carry-in]
        d, sd = o[0], o[1]
        a = self._u(w, o[2]).astype(np.uint64)
        b = self._u(w, o[3]).astype(np.uint64)
        cin = np.zeros(LANES, dtype=np.uint64)
        if len(o) > 4:
            cin = mask_to_bool(self.sv(w, o[4], 2)).astype(np.uint64)
        if kind == "add":
            r = a + b + cin
            co = r > np.uint64(M32)
        elif kind == "sub":
            r = a - b - cin
            co = a < b + cin
        else:
            r = b - a - cin
            co = b < a + cin
        self.vset(w, d, (r & np.uint64(M32)).astype(np.uint32))
        self.sset(w, sd, bool_to_mask(co & act), 2)

    def op_add_co_u32(self, w, o, m, act):
        self._carry_ops(w, o, "add", act)

    def op_addc_co_u32(self, w, o, m, act):
        self._carry_ops(w, o, "add", act)

    def op_sub_co_u32(self, w, o, m, act):
        self._carry_ops(w, o, "sub", act)

    def op_subb_co_u32(self, w, o, m, act):
        self._carry_ops(w, o, "sub", act)

    def op_subrev_co_u32(self, w, o, m, act):
        self._carry_ops(w, o, "rev", act)

    def op_subbrev_co_u32(self, w, o, m, act):
        self._carry_ops(w, o, "rev", act)

    def op_add3_u32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) + self._u(w, o[2]) + self._u(w, o[3]))

    def op_or3_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) | self._u(w, o[2]) | self._u(w, o[3]))

    def op_and_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) & self._u(w, o[2]))

    def op_or_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) | self._u(w, o[2]))

    def op_xor_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[1]) ^ self._u(w, o[2]))

    def op_not_b32(self, w, o, m, act):
        self.vset(w, o[0], ~self._u(w, o[1]))

    def op_lshlrev_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[2]) << (self._u(w, o[1]) & np.uint32(31)))

    def op_lshrrev_b32(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[2]) >> (self._u(w, o[1]) & np.uint32(31)))

    def op_ashrrev_i32(self, w, o, m, act):
        r = self._u(w, o[2]).astype(np.int32) >> (self._u(w, o[1]) & np.uint32(31)).astype(np.int32)
        self.vset(w, o[0], r.astype(np.uint32))

    def op_lshlrev_b64(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[2], 2) << (self._u(w, o[1]).astype(np.uint64) & np.uint64(63)), 2)

    def op_lshrrev_b64(self, w, o, m, act):
        self.vset(w, o[0], self._u(w, o[2], 2) >> (self._u(w, o[1]).astype(np.uint64) & np.uint64(63)), 2)

    def op_ashrrev_i64(self, w, o, m, act):
        r = self._u(w, o[2], 2).astype(np.int64) >> (self._u(w, o[1]).astype(np.int64) & 63)
        self.vset(w, o[0], r.astype(np.uint64), 2)

    def op_lshl_add_u32(self, w, o, m, act):
        self.vset(w, o[0], (self._u(w, o[1]) << (self._u(w, o[2]) & np.uint32(31))) + self._u(w, o[3]))

    def op_add_lshl_u32(self, w, o, m, act):
        self.vset(w, o[0], (self._u(w, o[1]) + self._u(w, o[2])) << (self._u(w, o[3]) & np.uint32(31)))

    def op_lshl_or_b32(self, w, o, m, act):
        self.vset(w, o[0], (self._u(w, o[1]) << (self._u(w, o[2]) & np.uint32(31))) | self._u(w, o[3]))

    def op_and_or_b32(self, w, o, m, act):
        self.vset(w, o[0], (self._u(w, o[1]) & self._u(w, o[2])) | self._u(w, o[3]))

    def op_lshl_add_u64(self, w, o, m, act):
        a = self._u(w, o[1], 2)
        sh = self._u(w, o[2]).astype(np.uint64) & np.uint64(63)
        c = self._u(w, o[3], 2)
        self.vset(w, o[0], (a << sh) + c, 2)

    def op_bfe_u32(self, w, o, m, act):
        a = self._u(w, o[1]).astype(np.uint64)
        off = self._u(w, o[2]).astype(np.uint64) & np.uint64(31)
        wd = self._u(w, o[3]).astype(np.uint64) & np.uint64(31)
        r = (a >> off) & ((np.uint64(1) << wd) - np.uint64(1))
        self.vset(w, o[0], r.astype(np.uint32))

    def op_bfe_i32(self, w, o, m, act):
        a = self._u(w, o[1]).astype(np.int64)
        off = self._u(w, o[2]).astype(np.int64) & 31
        wd = self._u(w, o[3]).astype(np.int64) & 31
        r = (a >> off) & ((1 << wd) - 1)
        sign = (r >> np.maximum(wd - 1, 0)) & 1
        r = np.where((wd > 0) & (sign == 1), r - (1 << wd), r)
        r = np.where(wd == 0, 0, r)
        self.vset(w, o[0], r.astype(np.int64).astype(np.uint32))

    def op_bfi_b32(self, w, o, m, act):
        a, b, c = self._u(w, o[1]), self._u(w, o[2]), self._u(w, o[3])
        self.vset(w, o[0], (a & b) | (~a & c))

    def op_bitop3_b32(self, w, o, m, act):
        tt = int(m["bitop3"], 0)
        a, b, c = self._u(w, o[1]), self._u(w, o[2]), self._u(w, o[3])
        r = np.zeros(LANES, dtype=np.uint32)
        for idx in range(8):
            if tt >> idx & 1:
                ta = a if idx & 4 else ~a
                tb = b if idx & 2 else ~b
                tc = c if idx & 1 else ~c
                r |= ta & tb & tc
        self.vset(w, o[0], r)

    def op_bitop3_b16(self, w, o, m, act):
        tt = int(m["bitop3"], 0)
        a, b, c = self._u(w, o[1]), self._u(w, o[2]), self._u(w, o[3])
        r = np.zeros(LANES, dtype=np.uint32)
        for idx in range(8):
            if tt >> idx & 1:
                ta = a if idx & 4 else ~a
                tb = b if idx & 2 else ~b
                tc = c if idx & 1 else ~c
                r |= ta & tb & tc
        old = self._u(w, o[0])
        self.vset(w, o[0], (old & np.uint32(0xFFFF0000)) | (r & np.uint32(0xFFFF)))

    def op_cndmask_b32(self, w, o, m, act):
        msk = o[3] if len(o) > 3 else "vcc"
        sel = mask_to_bool(self.sv(w, msk, 2))
        a = self.get(w, o[1], 1, True)
        b = self.get(w, o[2], 1, True)
        self.vset(w, o[0], np.where(sel, b, a))

    def op_mul_lo_u32(self, w, o, m, act):
        r = self._u(w, o[1]).astype(np.uint64) * self._u(w, o[2]).astype(np.uint64)
        self.vset(w, o[0], (r & np.uint64(M32)).astype(np.uint32))

    def op_mul_hi_u32(self, w, o, m, act):
        r = self._u(w, o[1]).astype(np.uint64) * self._u(w, o[2]).astype(np.uint64)
        self.vset(w, o[0], (r >> np.uint64(32)).astype(np.uint32))

    def op_mul_hi_i32(self, w, o, m, act):
        r = self._u(w, o[1]).astype(np.int32).astype(np.int64) * \
            self._u(w, o[2]).astype(np.int32).astype(np.int64)
        self.vset(w, o[0], (r >> 32).astype(np.uint32))

    def op_mul_u32_u24(self, w, o, m, act):
        a = self._u(w, o[1]) & np.uint32(0xFFFFFF)
        b = self._u(w, o[2]) & np.uint32(0xFFFFFF)
        r = a.astype(np.uint64) * b.astype(np.uint64)
        self.vset(w, o[0], (r & np.uint64(M32)).astype(np.uint32))

    def op_mul_i32_i24(self, w, o, m, act):
        def s24(x):
            x = (x & np.uint32(0xFFFFFF)).astype(np.int64)
            return np.where(x & 0x800000, x - 0x1000000, x)
        r = s24(self._u(w, o[1])) * s24(self._u(w, o[2]))
        self.vset(w, o[0], (r & M32).astype(np.uint32))

    def op_mad_u32_u24(self, w, o, m, act):
        a = self._u(w, o[1]) & np.uint32(0xFFFFFF)
        b = self._u(w, o[2]) & np.uint32(0xFFFFFF)
        r = a.astype(np.uint64) * b.astype(np.uint64) + self._u(w, o[3]).astype(np.uint64)
        self.vset(w, o[0], (r & np.uint64(M32)).astype(np.uint32))

    def op_mad_u64_u32(self, w, o, m, act):
        a = [int(x) for x in self._u(w, o[2])]
        b = [int(x) for x in self._u(w, o[3])]
        c = [int(x) for x in self._u(w, o[4], 2)]
        r = [a[i] * b[i] + c[i] for i in range(LANES)]
        self.vset(w, o[0], np.array([x & M64 for x in r], dtype=np.uint64), 2)
        self.sset(w, o[1], bool_to_mask(np.array([x > M64 for x in r]) & act), 2)

    def op_mad_i64_i32(self, w, o, m, act):
        a = [int(x) for x in self._u(w, o[2]).astype(np.int32)]
        b = [int(x) for x in self._u(w, o[3]).astype(np.int32)]
        c = [int(x) for x in self._u(w, o[4], 2).astype(np.int64)]
        r = [a[i] * b[i] + c[i] for i in range(LANES)]
        self.vset(w, o[0], np.array([x & M64 for x in r], dtype=np.uint64), 2)
        self.sset(w, o[1], bool_to_mask(np.array(
            [x > 0x7FFFFFFFFFFFFFFF or x < -0x8000000000000000 for x in r]) & act), 2)

    def op_max_i32(self, w, o, m, act):
        self.vset(w, o[0], np.maximum(self._u(w, o[1]).astype(np.int32),
                                      self._u(w, o[2]).astype(np.int32)).astype(np.uint32))

    def op_min_i32(self, w, o, m, act):
        self.vset(w, o[0], np.minimum(self._u(w, o[1]).astype(np.int32),
                                      self._u(w, o[2]).astype(np.int32)).astype(np.uint32))

    def op_max_u32(self, w, o, m, act):
        self.vset(w, o[0], np.maximum(self._u(w, o[1]), self._u(w, o[2])))

    def op_min_u32(self, w, o, m, act):
        self.vset(w, o[0], np.minimum(self._u(w, o[1]), self._u(w, o[2])))

    def op_min3_i32(self, w, o, m, act):
        a, b, c = (self._u(w, o[i]).astype(np.int32) for i in (1, 2, 3))
        self.vset(w, o[0], np.minimum(np.minimum(a, b), c).astype(np.uint32))

    def op_max3_i32(self, w, o, m, act):
        a, b, c = (self._u(w, o[i]).astype(np.int32) for i in (1, 2, 3))
        self.vset(w, o[0], np.maximum(np.maximum(a, b), c).astype(np.uint32))

    def op_readlane_b32(self, w, o, m, act):
        lane = self.sv(w, o[2]) & 63
        self.sset(w, o[0], int(w.v[int(_REG.match(o[1]).group(4))][lane]))

    def op_writelane_b32(self, w, o, m, act):
        lane = self.sv(w, o[2]) & 63
        val = self.sv(w, o[1])
        w.v[int(_REG.match(o[0]).group(4))][lane] = val

    def op_readfirstlane_b32(self, w, o, m, act):
        idx = int(np.argmax(act)) if np.any(act) else 0
        self.sset(w, o[0], int(self._u(w, o[1])[idx]))

    def op_bcnt_u32_b32(self, w, o, m, act):
        # D = popcount(S0) + S1
        x = self._u(w, o[1]).astype(np.uint32)
        x = x - ((x >> np.uint32(1)) & np.uint32(0x55555555))
        x = (x & np.uint32(0x33333333)) + ((x >> np.uint32(2)) & np.uint32(0x33333333))
        x = (x + (x >> np.uint32(4))) & np.uint32(0x0F0F0F0F)
        cnt = ((x * np.uint32(0x01010101)) >> np.uint32(24)).astype(np.uint32)
        self.vset(w, o[0], cnt + self._u(w, o[2]))

    def op_mbcnt_lo_u32_b32(self, w, o, m, act):
        msk = self.sv(w, o[1]) & M32
        c = self._u(w, o[2])
        cnt = np.array([bin(msk & ((1 << min(l, 32)) - 1)).count("1") for l in range(LANES)],
                       dtype=np.uint32)
        self.vset(w, o[0], cnt + c)

    def op_mbcnt_hi_u32_b32(self, w, o, m, act):
        msk = self.sv(w, o[1]) & M32
        c = self._u(w, o[2])
        cnt = np.array([bin(msk & ((1 << max(l - 32, 0)) - 1)).count("1") if l >= 32 else 0
                        for l in range(LANES)], dtype=np.uint32)
        self.vset(w, o[0], cnt + c)

    # ---- floating point ----
    def op_add_f64(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(self.fget(w, o[1]) + self.fget(w, o[2])), 2)

    def op_mul_f64(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(self.fget(w, o[1]) * self.fget(w, o[2])), 2)

    def op_fma_f64(self, w, o, m, act):
        r = two_prod_fma(self.fget(w, o[1]), self.fget(w, o[2]), self.fget(w, o[3]))
        self.vset(w, o[0], u64_from_f64(r), 2)

    def op_fmac_f64(self, w, o, m, act):
        r = two_prod_fma(self.fget(w, o[1]), self.fget(w, o[2]), self.fget(w, o[0]))
        self.vset(w, o[0], u64_from_f64(r), 2)

    def op_max_f64(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(np.fmax(self.fget(w, o[1]), self.fget(w, o[2]))), 2)

    def op_min_f64(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(np.fmin(self.fget(w, o[1]), self.fget(w, o[2]))), 2)

    def op_rcp_f64(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(1.0 / self.fget(w, o[1])), 2)

    def op_rsq_f64(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(1.0 / np.sqrt(self.fget(w, o[1]))), 2)

    def op_sqrt_f64(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(np.sqrt(self.fget(w, o[1]))), 2)

    def op_div_scale_f64(self, w, o, m, act):
        # no scaling in range; VCC-style flag cleared
        self.vset(w, o[0], u64_from_f64(self.fget(w, o[2])), 2)
        self.sset(w, o[1], 0, 2)

    def op_div_fmas_f64(self, w, o, m, act):
        r = two_prod_fma(self.fget(w, o[1]), self.fget(w, o[2]), self.fget(w, o[3]))
        self.vset(w, o[0], u64_from_f64(r), 2)

    def op_div_fixup_f64(self, w, o, m, act):
        q = self.fget(w, o[1])
        d = self.fget(w, o[2])
        n = self.fget(w, o[3])
        with np.errstate(all="ignore"):
            exact = n / d
        special = ~np.isfinite(q) | ~np.isfinite(d) | ~np.isfinite(n) | (d == 0)
        self.vset(w, o[0], u64_from_f64(np.where(special, exact, q)), 2)

    def op_trig_preop_f64(self, w, o, m, act):
        """53-bit segment S1[4:0] of 2/pi, skipping exponent(S0) - 1077
        further bits for large S0, scaled to its place value."""
        a = self._u(w, o[1], 2)
        seg = self._u(w, o[2]) & np.uint32(31)
        out = np.zeros(LANES, dtype=np.uint64)
        for l in np.nonzero(act)[0]:
            e = int((int(a[l]) >> 52) & 0x7FF)
            shift = int(seg[l]) * 53
            if e > 1077:
                shift += e - 1077
            c = (TWO_OVER_PI >> (TWO_OVER_PI_NBITS - shift - 53)) & ((1 << 53) - 1)
            scale = -53 - shift
            if e >= 1968:
                scale += 128
            out[l] = u64_from_f64(np.ldexp(np.float64(c), scale))
        self.vset(w, o[0], out, 2)

    def op_mad_i32_i24(self, w, o, m, act):
        def s24(x):
            x = x.astype(np.int64) & 0xFFFFFF
            return np.where(x & 0x800000, x - 0x1000000, x)
        a = s24(self._u(w, o[1]))
        b = s24(self._u(w, o[2]))
        c = self._u(w, o[3]).astype(np.int64)
        self.vset(w, o[0], ((a * b + c) & M32).astype(np.uint32))

    def op_ldexp_f64(self, w, o, m, act):
        a = self.fget(w, o[1])
        e = self._u(w, o[2]).astype(np.int32)
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(np.ldexp(a, e)), 2)

    def op_frexp_mant_f64(self, w, o, m, act):
        mt, _ = np.frexp(self.fget(w, o[1]))
        self.vset(w, o[0], u64_from_f64(mt), 2)

    def op_frexp_exp_i32_f64(self, w, o, m, act):
        _, ex = np.frexp(self.fget(w, o[1]))
        self.vset(w, o[0], ex.astype(np.int32).astype(np.uint32))

    def op_floor_f64(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(np.floor(self.fget(w, o[1]))), 2)

    def op_ceil_f64(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(np.ceil(self.fget(w, o[1]))), 2)

    def op_trunc_f64(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(np.trunc(self.fget(w, o[1]))), 2)

    def op_rndne_f64(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(np.rint(self.fget(w, o[1]))), 2)

    def op_fract_f64(self, w, o, m, act):
        a = self.fget(w, o[1])
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u64_from_f64(np.minimum(a - np.floor(a), 0.9999999999999999)), 2)

    def op_cvt_f64_i32(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(self._u(w, o[1]).astype(np.int32).astype(np.float64)), 2)

    def op_cvt_f64_u32(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(self._u(w, o[1]).astype(np.float64)), 2)

    def op_cvt_i32_f64(self, w, o, m, act):
        a = self.fget(w, o[1])
        with np.errstate(all="ignore"):
            r = np.where(np.isnan(a), 0, np.clip(np.trunc(a), -2147483648.0, 2147483647.0))
        self.vset(w, o[0], r.astype(np.int64).astype(np.uint32))

    def op_cvt_u32_f64(self, w, o, m, act):
        a = self.fget(w, o[1])
        with np.errstate(all="ignore"):
            r = np.where(np.isnan(a), 0, np.clip(np.trunc(a), 0.0, 4294967295.0))
        self.vset(w, o[0], r.astype(np.uint64).astype(np.uint32))

    def op_cvt_f32_f64(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32(self.fget(w, o[1]).astype(np.float32)))

    def op_cvt_f64_f32(self, w, o, m, act):
        self.vset(w, o[0], u64_from_f64(self.f32get(w, o[1]).astype(np.float64)), 2)

    def op_cvt_f32_u32(self, w, o, m, act):
        self.vset(w, o[0], u32_from_f32(self._u(w, o[1]).astype(np.float32)))

    def op_cvt_f32_i32(self, w, o, m, act):
        self.vset(w, o[0], u32_from_f32(self._u(w, o[1]).astype(np.int32).astype(np.float32)))

    def op_cvt_u32_f32(self, w, o, m, act):
        a = self.f32get(w, o[1]).astype(np.float64)
        with np.errstate(all="ignore"):
            r = np.where(np.isnan(a), 0, np.clip(np.trunc(a), 0.0, 4294967295.0))
        self.vset(w, o[0], r.astype(np.uint64).astype(np.uint32))

    def op_cvt_i32_f32(self, w, o, m, act):
        a = self.f32get(w, o[1]).astype(np.float64)
        with np.errstate(all="ignore"):
            r = np.where(np.isnan(a), 0, np.clip(np.trunc(a), -2147483648.0, 2147483647.0))
        self.vset(w, o[0], r.astype(np.int64).astype(np.uint32))

    def op_trunc_f32(self, w, o, m, act):
        self.vset(w, o[0], u32_from_f32(np.trunc(self.f32get(w, o[1]))))

    def op_mul_f32(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32(self.f32get(w, o[1]) * self.f32get(w, o[2])))

    def op_add_f32(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32(self.f32get(w, o[1]) + self.f32get(w, o[2])))

    def op_sub_f32(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32(self.f32get(w, o[1]) - self.f32get(w, o[2])))

    def op_fmac_f32(self, w, o, m, act):
        a = self.f32get(w, o[1]).astype(np.float64)
        b = self.f32get(w, o[2]).astype(np.float64)
        c = self.f32get(w, o[0]).astype(np.float64)
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32((a * b + c).astype(np.float32)))

    def op_fma_f32(self, w, o, m, act):
        a = self.f32get(w, o[1]).astype(np.float64)
        b = self.f32get(w, o[2]).astype(np.float64)
        c = self.f32get(w, o[3]).astype(np.float64)
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32((a * b + c).astype(np.float32)))

    def op_rcp_f32(self, w, o, m, act):
        with np.errstate(all="ignore"):
            self.vset(w, o[0], u32_from_f32(np.float32(1.0) / self.f32get(w, o[1])))

    def op_rcp_iflag_f32(self, w, o, m, act):
        self.op_rcp_f32(w, o, m, act)

    # ---- memory ----
    def _gaddr(self, w, ops, mods, vaddr_i, saddr_i):
        off = int(mods.get("offset", "0"), 0)
        saddr = ops[saddr_i]
        if saddr == "off":
            a = self._u(w, ops[vaddr_i], 2).astype(np.uint64)
        else:
            base = self.sreg(w, saddr, 2)
            a = np.uint64(base) + self._u(w, ops[vaddr_i]).astype(np.uint64)
        return (a.astype(np.int64) + off).astype(np.uint64)

    def gmem(self, w, op, ops, mods, where):
        act = mask_to_bool(w.exec)
        sizes = {"dword": 4, "dwordx2": 8, "dwordx3": 12, "dwordx4": 16,
                 "ubyte": 1, "ushort": 2, "sbyte": 1, "sshort": 2, "byte": 1,
                 "short": 2}
        m = re.match(r"global_(load|store|atomic)_(\w+)$", op)
        if not m:
            raise NotImplementedError(op)
        kind, t = m.groups()
        if kind == "atomic":
            # memory-side atomic operations: one per active lane
            self.atomic_lanes = getattr(self, "atomic_lanes", 0) + int(act.sum())
        if kind == "load":
            n = sizes[t]
            addr = self._gaddr(w, ops, mods, 1, 2)
            d = _REG.match(ops[0])
            lo = int(d.group(2) if d.group(2) else d.group(4))
            for l in np.nonzero(act)[0]:
                data = self.mem.read(int(addr[l]), n, where + (int(l),))
                if t in ("ubyte", "ushort"):
                    w.v[lo][l] = int.from_bytes(data, "little")
                elif t in ("sbyte", "sshort"):
                    w.v[lo][l] = int.from_bytes(data, "little", signed=True) & M32
                else:
                    for i in range(n // 4):
                        w.v[lo + i][l] = int.from_bytes(data[4 * i:4 * i + 4], "little")
            return
        if kind == "store":
            n = sizes[t]
            addr = self._gaddr(w, ops, mods, 0, 2)
            d = _REG.match(ops[1])
            lo = int(d.group(2) if d.group(2) else d.group(4))
            for l in np.nonzero(act)[0]:
                if n < 4:
                    data = int(w.v[lo][l] & ((1 << (8 * n)) - 1)).to_bytes(n, "little")
                else:
                    data = b"".join(int(w.v[lo + i][l]).to_bytes(4, "little")
                                    for i in range(n // 4))
                self.mem.write(int(addr[l]), data, where + (int(l),))
            return
        # atomics: global_atomic_<op> [vdst,] vaddr, vdata, saddr [glc]
        aop = t
        ret = "glc" in mods or "sc0" in mods
        if ret:
            dst, vaddr_i, data_i, saddr_i = ops[0], 1, 2, 3
        else:
            dst, vaddr_i, data_i, saddr_i = None, 0, 1, 2
        addr = self._gaddr(w, ops, mods, vaddr_i, saddr_i)
        dv = self._u(w, ops[data_i])
        out = np.zeros(LANES, dtype=np.uint32)
        for l in np.nonzero(act)[0]:
            old = int.from_bytes(self.mem.read(int(addr[l]), 4, where + (int(l),)), "little")
            so = old - (1 << 32) if old & (1 << 31) else old
            v = int(dv[l])
            sv_ = v - (1 << 32) if v & (1 << 31) else v
            new = {"add": (old + v) & M32, "smax": max(so, sv_) & M32,
                   "smin": min(so, sv_) & M32, "umax": max(old, v),
                   "umin": min(old, v), "swap": v}[aop]
            self.mem.write(int(addr[l]), new.to_bytes(4, "little"), where + (int(l),))
            out[l] = old
        if dst:
            self.vset(w, dst, out)

    def pmem(self, w, op, ops, mods, where):
        """scratch_{load,store}_*: per-lane private memory of the kernel's
        private segment size; lane offset = VGPR (or 0) + SGPR (or 0) +
        instruction offset."""
        act = mask_to_bool(w.exec)
        sizes = {"dword": 4, "dwordx2": 8, "dwordx3": 12, "dwordx4": 16}
        m = re.match(r"scratch_(load|store)_(\w+)$", op)
        if not m or m.group(2) not in sizes:
            raise NotImplementedError(op)
        kind, t = m.groups()
        n = sizes[t]
        if w.priv is None:
            w.priv = [bytearray([0xBA]) * max(self.k.private_size, 0)
                      for _ in range(LANES)]
        imm = int(mods.get("offset", "0"), 0)
        if kind == "load":
            dst, va, sa = ops[0], ops[1], ops[2]
        else:
            va, dst, sa = ops[0], ops[1], ops[2]
        voff = self._u(w, va).astype(np.int64) if va != "off" \
            else np.zeros(LANES, dtype=np.int64)
        soff = 0
        if sa != "off":
            self.check_sdef(w, sa, where, "scratch-offset")
            soff = self.sreg(w, sa, 1)
            soff = soff - (1 << 32) if soff & (1 << 31) else soff
        d = _REG.match(dst)
        lo = int(d.group(2) if d.group(2) else d.group(4))
        for l in np.nonzero(act)[0]:
            a = int(voff[l]) + soff + imm
            mem = w.priv[l]
            if a < 0 or a + n > len(mem):
                self.mem.violations.append((where + (int(l),), "private-" + kind, a, n))
                continue
            if kind == "load":
                for i in range(n // 4):
                    w.v[lo + i][l] = int.from_bytes(mem[a + 4 * i:a + 4 * i + 4], "little")
            else:
                for i in range(n // 4):
                    mem[a + 4 * i:a + 4 * i + 4] = int(w.v[lo + i][l]).to_bytes(4, "little")

    def dsmem(self, w, op, ops, mods, where):
        act = mask_to_bool(w.exec)
        lds = self.lds

        def rd(a, n, l):
            if a < 0 or a + n > len(lds):
                self.mem.violations.append((where + (int(l),), "lds-read", a, n))
                return bytes(n)
            return bytes(lds[a:a + n])

        def wr(a, data, l):
            if a < 0 or a + len(data) > len(lds):
                self.mem.violations.append((where + (int(l),), "lds-write", a, len(data)))
                return
            lds[a:a + len(data)] = data

        def vlo(tok):
            d = _REG.match(tok)
            return int(d.group(2) if d.group(2) else d.group(4))
        if op == "ds_bpermute_b32":
            addr = self._u(w, ops[1])
            data = self._u(w, ops[2])
            src = (addr >> np.uint32(2)) & np.uint32(63)
            self.vset(w, ops[0], data[src.astype(np.int64)])
            return
        m = re.match(r"ds_(read|write)(2st64|2)?_b(32|64|96|128)$", op)
        if not m:
            raise NotImplementedError(op)
        kind, two, bits = m.group(1), m.group(2), int(m.group(3))
        n = bits // 8
        if kind == "read":
            addr = self._u(w, ops[1])
            dl = vlo(ops[0])
            if two:
                o0 = int(mods.get("offset0", "0"), 0)
                o1 = int(mods.get("offset1", "0"), 0)
                scale = n * (64 if two == "2st64" else 1)
                for l in np.nonzero(act)[0]:
                    for j, oo in enumerate((o0, o1)):
                        data = rd(int(addr[l]) + oo * scale, n, l)
                        for i in range(n // 4):
                            w.v[dl + j * (n // 4) + i][l] = int.from_bytes(
                                data[4 * i:4 * i + 4], "little")
            else:
                off = int(mods.get("offset", "0"), 0)
                for l in np.nonzero(act)[0]:
                    data = rd(int(addr[l]) + off, n, l)
                    for i in range(n // 4):
                        w.v[dl + i][l] = int.from_bytes(data[4 * i:4 * i + 4], "little")
            return
        addr = self._u(w, ops[0])
        if two:
            o0 = int(mods.get("offset0", "0"), 0)
            o1 = int(mods.get("offset1", "0"), 0)
            scale = n * (64 if two == "2st64" else 1)
            d0, d1 = vlo(ops[1]), vlo(ops[2])
            for l in np.nonzero(act)[0]:
                for oo, dl in ((o0, d0), (o1, d1)):
                    data = b"".join(int(w.v[dl + i][l]).to_bytes(4, "little")
                                    for i in range(n // 4))
                    wr(int(addr[l]) + oo * scale, data, l)
        else:
            off = int(mods.get("offset", "0"), 0)
            dl = vlo(ops[1])
            for l in np.nonzero(act)[0]:
                data = b"".join(int(w.v[dl + i][l]).to_bytes(4, "little")
                                for i in range(n // 4))
                wr(int(addr[l]) + off, data, l)
