"""Rewrite the kernel-launch syntax of a .hip source for the CPU emulation
build (test infrastructure): `kern<<<grid, block[, shmem, stream]>>>(args)`
becomes `bb_emu::launch("kern", kern, grid, block[, shmem, stream])(args)`.
Nothing else in the source changes."""
import re
import sys

LAUNCH = re.compile(
    r"([A-Za-z_][A-Za-z0-9_]*(?:<[^<>;()]*>)?)\s*<<<(.*?)>>>\s*\(", re.S)


def translate(text):
    def rep(m):
        callee, cfg = m.group(1), m.group(2)
        return 'bb_emu::launch("%s", %s, %s)(' % (callee, callee, cfg)
    out, n = LAUNCH.subn(rep, text)
    if "<<<" in out:
        raise SystemExit("translate.py: launch syntax left untranslated")
    return out, n


if __name__ == "__main__":
    src, dst = sys.argv[1], sys.argv[2]
    text = open(src).read()
    out, n = translate(text)
    with open(dst, "w") as f:
        f.write('#line 1 "%s"\n' % src)
        f.write(out)
