"""Build check (CPU): the committed Montgomery multiply assembly (mont_asm.h) is
exactly what gen_mont_asm.py emits, so its checked schedule holds for the
code that is compiled: every carry add at least two instructions after the
multiply that wrote its SGPR pair, and every multiply whose carry-out is
dropped either an m * N product (modulus limbs at most 0xF0000001, carried-in
value below 2^37: no overflow) or a function's first product (accumulator
zero)."""
import os
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "era-bellman-cuda_amd", "csrc")


def test_mont_asm_matches_generator():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_mont_asm.py")], check=True, capture_output=True, text=True).stdout
    assert out == open(os.path.join(CSRC, "mont_asm.h")).read()


def test_dropped_carries_are_provably_zero():
    """Re-checks the no-overflow argument from the field constants."""
    text = open(os.path.join(CSRC, "field.cuh")).read()
    for name in ("FqParams", "FrParams"):
        body = text[text.index("struct %s" % name):]
        limbs = [int(v, 16) for v in re.search(r"m\[8\] = \{([^}]*)\}", body).group(1).replace("0x", "").split(",")]
        assert max(limbs) <= 0xF0000001
        # a column sums at most 24 products below 2^64, so the carried-in value is below 24 * 2^32
        assert (2 ** 32 - 1) * max(limbs) + 24 * 2 ** 32 < 2 ** 64
    stmts = re.findall(r'asm\("([^"]*)"', open(os.path.join(CSRC, "mont_asm.h")).read())
    assert len(stmts) == 46
    for idx, stmt in enumerate(stmts):
        for pos, line in enumerate(stmt.split("\\n\\t")):
            ops = re.findall(r"%\[(\w+)\]", line)
            if line.startswith("v_mad_u64_u32") and ops[1] == "dd":
                assert (ops[2].startswith("m") and ops[3].startswith("n")) or (idx % 23 == 0 and pos == 0), line
