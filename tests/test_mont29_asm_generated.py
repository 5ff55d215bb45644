"""Build check (CPU): the committed 29-bit Montgomery product assembly
(mont29_asm.h) is exactly what gen_mont29_asm.py emits, and an emulation of
the committed statements (v_mad_u64_u32 and v_lshrrev_b64 on Python integers,
the C++ steps between them as written) gives x * y / 2^261 mod m, normalized,
for Fq and Fr on random and extreme inputs, without any column ever carrying
out of its 64-bit accumulator (the statements drop those carries)."""
import os
import random
import re
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "era-bellman-cuda_amd", "csrc")
HEADER = os.path.join(CSRC, "mont29_asm.h")
M29 = (1 << 29) - 1


def test_mont29_asm_matches_generator():
    out = subprocess.run([sys.executable, os.path.join(CSRC, "gen_mont29_asm.py")], check=True, capture_output=True, text=True).stdout
    assert out == open(HEADER).read()


def moduli():
    text = open(os.path.join(CSRC, "field.cuh")).read()
    out = {}
    for name in ("FqParams", "FrParams"):
        body = text[text.index("struct %s" % name):]
        words = [int(v, 16) for v in re.search(r"m\[8\] = \{([^}]*)\}", body).group(1).replace("0x", "").split(",")]
        out[name] = sum(w << (32 * i) for i, w in enumerate(words))
    return out


def functions():
    """name -> list of steps: ("asm", [instructions], {operand: source}) or ("cpp", line)."""
    text = open(HEADER).read()
    out = {}
    for m in re.finditer(r"(?:f29|void) (\w+)\(([^)]*)\) \{(.*?)\n\}", text, flags=re.S):
        steps = []
        for part in re.finditer(r'asm\("((?:[^"\\]|\\.)*)"\s*:(.*?):(.*?)\);|\n  ([^\n]*;)', m.group(3), flags=re.S):
            if part.group(1) is not None:
                code = [line.strip() for line in part.group(1).split("\\n\\t")]
                srcs = dict(re.findall(r'\[(\w+)\]\s*"[vs]"\(([^)]*\]?)\)', part.group(3)))
                steps.append(("asm", code, srcs))
            else:
                steps.append(("cpp", part.group(4)))
        out[m.group(1)] = steps
    return out


def emulate(steps, m, operands):
    """Runs the statements; returns the output limbs by name (t, or r and s)."""
    inv = (-pow(m, -1, 1 << 29)) % (1 << 29)
    N = [(m >> (29 * i)) & M29 if i < 8 else m >> 232 for i in range(9)]
    env = {k: list(v) for k, v in operands.items()}
    env["N"] = N
    for name in ("q0", "q1", "q2", "t", "r", "s"):
        env[name] = [0] * 9
    for x in ("a", "c"):
        if x in env:
            env[x + "2x"] = [v << 1 for v in env[x]]
    acc = {}

    def value(src):
        mm = re.fullmatch(r"(\w+?)(?:\.v)?\[(\d)\]", src)
        return env[mm.group(1)][int(mm.group(2))]

    for step in steps:
        if step[0] == "asm":
            _, code, srcs = step
            for line in code:
                op, _, rest = line.partition(" ")
                args = [a.strip() for a in rest.split(",")]
                dst = re.fullmatch(r"%\[(acc\d)\]", args[0]).group(1)
                if op == "v_mad_u64_u32":
                    x, y = (value(srcs[re.fullmatch(r"%\[(\w+)\]", a).group(1)]) for a in args[2:4])
                    if args[4] == "0":
                        addend = 0
                    else:
                        assert args[4] == args[0]
                        addend = acc[dst]
                    acc[dst] = x * y + addend
                    assert acc[dst] < 1 << 64, "a column carries out of its accumulator"
                elif op == "v_lshrrev_b64":
                    assert args[1] == "29" and args[2] == args[0]
                    acc[dst] >>= 29
                else:
                    raise AssertionError(line)
        else:
            line = step[1].strip()
            if mm := re.fullmatch(r"(q\d)\[(\d)\] = \(\(uint32_t\)(acc\d) \* P29<F>::inv\) & MASK29;", line):
                env[mm.group(1)][int(mm.group(2))] = ((acc[mm.group(3)] & 0xFFFFFFFF) * inv) & M29
            elif mm := re.fullmatch(r"(\w)\.v\[(\d)\] = \(uint32_t\)(acc\d) & MASK29;", line):
                env[mm.group(1)][int(mm.group(2))] = acc[mm.group(3)] & M29
            elif mm := re.fullmatch(r"(\w)\.v\[8\] = \(uint32_t\)\((acc\d) >> 29\);", line):
                assert acc[mm.group(2)] >> 29 < 1 << 32
                env[mm.group(1)][8] = acc[mm.group(2)] >> 29
            elif line.startswith(("constexpr", "uint32_t q", "f29 ", "uint64_t", "(void)", "return", "#pragma", "for (", "a2x.v[i]", "c2x.v[i]")):
                continue
            else:
                raise AssertionError(line)
    return env


def limbs(x):
    return [(x >> (29 * i)) & M29 for i in range(8)] + [x >> 232]


def value29(v):
    return sum(x << (29 * i) for i, x in enumerate(v))


def montgomery_ok(t, num, m):
    assert all(x <= M29 for x in t[:8]), t
    assert (value29(t) << 261) % m == num % m
    assert value29(t) < (num >> 261) + 2 * m


def check(name, steps, m, ops):
    env = emulate(steps, m, ops)
    vals = {k: value29(v) for k, v in ops.items()}
    if name == "mul29_device":
        montgomery_ok(env["t"], vals["a"] * vals["b"], m)
    elif name == "mul2_29_device":
        montgomery_ok(env["t"], vals["a"] * vals["b"] + vals["c"] * vals["e"], m)
    elif name == "sqr29_device":
        montgomery_ok(env["t"], vals["a"] * vals["a"], m)
    elif name == "mul29x2_device":
        montgomery_ok(env["r"], vals["a"] * vals["b"], m)
        montgomery_ok(env["s"], vals["c"] * vals["e"], m)
    elif name == "mul2_mul_mul29_device":
        montgomery_ok(env["r"], vals["a"] * vals["b"] + vals["c"] * vals["e"], m)
        montgomery_ok(env["s"], vals["f"] * vals["g"], m)
        montgomery_ok(env["t"], vals["h"] * vals["k"], m)
    else:
        assert name == "sqr29x2_device"
        montgomery_ok(env["r"], vals["a"] * vals["a"], m)
        montgomery_ok(env["s"], vals["c"] * vals["c"], m)


OPERANDS = {"mul29_device": "ab", "mul2_29_device": "abce", "sqr29_device": "a", "mul29x2_device": "abce", "sqr29x2_device": "ac",
            "mul2_mul_mul29_device": "abcefghk"}


def test_emulated_products_are_montgomery_products():
    fns = functions()
    assert set(fns) == set(OPERANDS)
    mads = {n: sum(len([c for c in s[1] if c.startswith("v_mad")]) for s in st if s[0] == "asm") for n, st in fns.items()}
    assert mads == {"mul29_device": 162, "mul2_29_device": 243, "sqr29_device": 126, "mul29x2_device": 324, "sqr29x2_device": 252,
                    "mul2_mul_mul29_device": 567}
    rng = random.Random(0xBC29)
    for m in moduli().values():
        top = (1 << 260) - 1  # inputs below 2^260: top limbs below 2^28
        cases = [[top] * 8, [0] * 8, [m - 1] * 8, [1, m - 1, 2 * m - 1, 3 * m] * 2]
        cases += [[rng.randrange(1 << 260) for _ in range(8)] for _ in range(200)]
        for xs in cases:
            vals = dict(zip("abcefghk", (limbs(x) for x in xs)))
            for name, steps in fns.items():
                check(name, steps, m, {k: vals[k] for k in OPERANDS[name]})
        # every limb at its bound (normalized lower limbs at 2^29 - 1), and
        # mul29's one wide operand with limbs up to 2^31 - 1
        wide = [(1 << 31) - 1] * 8 + [(1 << 28) - 1]
        full = [M29] * 8 + [(1 << 28) - 1]
        check("mul29_device", fns["mul29_device"], m, {"a": wide, "b": full})
        check("mul29x2_device", fns["mul29x2_device"], m, {"a": wide, "b": full, "c": full, "e": wide})
        # operands below 2^261, every limb at 2^29 - 1
        top29 = [M29] * 9
        check("mul29x2_device", fns["mul29x2_device"], m, {"a": top29, "b": full, "c": top29, "e": top29})
        check("mul2_29_device", fns["mul2_29_device"], m, {"a": full, "b": full, "c": full, "e": full})
        check("sqr29_device", fns["sqr29_device"], m, {"a": full})
        check("sqr29x2_device", fns["sqr29x2_device"], m, {"a": full, "c": full})
        check("mul2_mul_mul29_device", fns["mul2_mul_mul29_device"], m,
              {"a": full, "b": full, "c": full, "e": full, "f": wide, "g": full, "h": full, "k": wide})


def test_emulation_detects_a_wrong_statement():
    """A schedule that drops one product fails the congruence check."""
    steps = functions()["mul29_device"]
    broken = []
    dropped = False
    for s in steps:
        if s[0] == "asm" and not dropped and any("%[l0a1], %[l0b0]" in c for c in s[1]):
            broken.append(("asm", [c for c in s[1] if "%[l0a1], %[l0b0]" not in c], s[2]))
            dropped = True
        else:
            broken.append(s)
    m = moduli()["FqParams"]
    ops = {"a": limbs(12345678901234567890123), "b": limbs(98765432109876543210987)}
    try:
        check("mul29_device", broken, m, ops)
    except AssertionError:
        return
    raise AssertionError("the dropped product went unnoticed")
