"""Build checks (CPU) on the device assembly hipcc writes for every kernel.

1. No kernel keeps a private (scratch) segment.  Kernels that use the
   Montgomery multiply's inline assembly are only validated without register
   spills (a spilling kernel computed wrong values on gfx950), so every kernel
   must fit its registers.
2. No inline-assembly block of the field arithmetic has an output register
   that is also one of its inputs.  The multiply's statements declare their
   column words and carries early-clobber, but the register allocator has
   been seen to give the column's overflow word the register of an input
   limb still needed by the next statement (the top limb of a value loaded
   into a register tuple), which silently corrupts the product; the check
   reads the allocation back from the assembly."""
import os
import re
import shutil
import subprocess
import tempfile
from concurrent.futures import ThreadPoolExecutor

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "era-bellman-cuda_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "--cuda-device-only", "-S"]

pytestmark = pytest.mark.skipif(not os.path.exists(HIPCC) or shutil.which("ld.lld") is None and not os.path.exists("/opt/rocm/llvm/bin/ld.lld"),
                                reason="hipcc not available")


def compile_to_asm(src, out_dir):
    out = os.path.join(out_dir, os.path.basename(src) + ".s")
    subprocess.run([HIPCC] + FLAGS + ["-o", out, src], check=True, capture_output=True)
    return open(out).read()


@pytest.fixture(scope="module")
def device_asm():
    sources = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))
    with tempfile.TemporaryDirectory() as d, ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        texts = list(pool.map(lambda s: compile_to_asm(s, d), sources))
    return dict(zip((os.path.basename(s) for s in sources), texts))


def test_no_kernel_spills_to_scratch(device_asm):
    kernels = []
    for text in device_asm.values():
        for name, body in re.findall(r"\.amdhsa_kernel (\S+)(.*?)\.end_amdhsa_kernel", text, flags=re.S):
            kernels.append((name, int(re.search(r"\.amdhsa_private_segment_fixed_size (\d+)", body).group(1))))
    assert len(kernels) > 40
    spilling = [(name, size) for name, size in kernels if size > 0]
    assert not spilling, spilling


def registers(operand):
    m = re.fullmatch(r"([vs])\[(\d+):(\d+)\]", operand)
    if m:
        return {m.group(1) + str(i) for i in range(int(m.group(2)), int(m.group(3)) + 1)}
    return {operand} if re.fullmatch(r"[vs]\d+", operand) else set()


def instructions(text):
    out = []
    for line in text.split("\n"):
        line = line.strip()
        if line:
            op, _, rest = line.partition(" ")
            out.append((op, [a.strip() for a in rest.split(",")] if rest else []))
    return out


def asm_templates():
    """(instructions, output names, input names) of every inline-asm statement."""
    text = "".join(open(os.path.join(CSRC, h)).read() for h in ("mont_asm.h", "mont29_asm.h"))
    out = []
    for m in re.finditer(r'asm\("((?:[^"\\]|\\.)*)"\s*:(.*?):(.*?)\);', text, flags=re.S):
        body = m.group(1).replace("\\n", "\n").replace("\\t", " ")
        outs = re.findall(r'\[(\w+)\]\s*"[=+]', m.group(2))
        ins = re.findall(r'\[(\w+)\]\s*"', m.group(3))
        out.append((instructions(body), outs, ins))
    return out


def operand_clash(block, template):
    """None if `block` is not an instance of `template`; otherwise the output
    operands whose registers overlap another operand's."""
    code, outs, _ = template
    if len(code) != len(block):
        return None
    bound = {}
    for (op, args), (t_op, t_args) in zip(block, code):
        if op != t_op or len(args) != len(t_args):
            return None
        for a, t in zip(args, t_args):
            m = re.fullmatch(r"%\[(\w+)\]", t)
            if not m:
                if a != t:
                    return None
            elif bound.setdefault(m.group(1), a) != a:
                return None
    regs = {name: registers(a) for name, a in bound.items()}
    return sorted(o for o in outs if o in regs and any(regs[o] & r for name, r in regs.items() if name != o))


def test_operand_check_flags_shared_registers():
    template = (instructions("v_mad_u64_u32 %[acc], %[p0], %[a7], %[b6], %[acc]\nv_addc_co_u32_e64 %[c2], %[dd], 0, 0, %[p0]"),
                ["acc", "c2", "p0", "dd"], ["a7", "b6"])
    ok = instructions("v_mad_u64_u32 v[70:71], s[18:19], v33, v36, v[70:71]\nv_addc_co_u32_e64 v34, s[22:23], 0, 0, s[18:19]")
    assert operand_clash(ok, template) == []
    # the overflow word on the input limb a7, then on the accumulator's high word
    on_input = instructions("v_mad_u64_u32 v[70:71], s[18:19], v33, v36, v[70:71]\nv_addc_co_u32_e64 v33, s[22:23], 0, 0, s[18:19]")
    assert operand_clash(on_input, template) == ["c2"]
    on_acc = instructions("v_mad_u64_u32 v[70:71], s[18:19], v33, v36, v[70:71]\nv_addc_co_u32_e64 v71, s[22:23], 0, 0, s[18:19]")
    assert operand_clash(on_acc, template) == ["acc", "c2"]
    other = instructions("v_mad_u64_u32 v[70:71], s[18:19], v33, v36, v[70:71]")
    assert operand_clash(other, template) is None


def test_no_inline_asm_operand_overlaps(device_asm):
    templates = asm_templates()
    assert len(templates) >= 40
    blocks, unmatched, bad = 0, [], []
    for name, text in device_asm.items():
        lines = text.splitlines()
        fn, i = None, 0
        while i < len(lines):
            m = re.match(r"^(_Z\w+):", lines[i])
            if m:
                fn = m.group(1)
            if ";;#ASMSTART" in lines[i]:
                j = i + 1
                while ";;#ASMEND" not in lines[j]:
                    j += 1
                blocks += 1
                block = instructions("\n".join(lines[i + 1:j]))
                results = [r for r in (operand_clash(block, t) for t in templates) if r is not None]
                if not results:
                    unmatched.append((name, fn, i + 1))
                elif any(results):
                    bad.append((name, fn, i + 1, [r for r in results if r][0]))
                i = j
            i += 1
    assert blocks > 1000
    assert not unmatched, unmatched[:5]
    assert not bad, bad[:10]
