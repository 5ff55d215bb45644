"""Generates mont_asm.h: the 8 x 32-bit-limb Montgomery products used by
field.cuh on gfx950, as a few inline-asm statements per column.

mul_lazy_device(a, b) = a b / R and mul2_lazy_device(a, b, c, d) =
(a b + c d) / R (one reduction for the sum of two products).

Column k of the finely integrated product scan accumulates a_i * b_(k-i) (and
c_i * d_(k-i)) and m_i * N_(k-i) into a 64-bit register pair with
v_mad_u64_u32.  Every carry-out goes to an SGPR pair and is added into the
column's overflow word c2 by a v_addc_co_u32_e64 that reads the pair.  gfx950
needs wait states between a VALU write of an SGPR and a VALU read of it as a
carry (the compiler itself keeps two instructions between them), so the
carries rotate over three SGPR pairs and each add is scheduled at least two
instructions after the multiply that wrote its pair; the adds write their own
(always zero) carry-out to a fourth pair that nothing reads.  Only where a
column is too short to fill the gap does the scheduler emit s_nop.

The first product of each column keeps no carry.  From column 1 on it is an
m_i * N_j product: every modulus limb is at most 0xF0000001, so it is below
2^64 - 2^60, and the carried-in value is below 2^37 (a column holds at most 24
products), so their sum stays below 2^64.  Column 0 starts from zero.

Statements: per low column k < 8 one for the products and m_0..m_(k-1) * N,
then (after m_k = lo(acc) * inv in C++) one for m_k * N_0 that also drains
the carries the first left pending; per high column one statement.  Between
statements the compiler emits at least the v_mul_lo_u32 of m_k, so carries
pending across a statement boundary count one instruction of distance.

Run: python3 gen_mont_asm.py > mont_asm.h
"""

PAIRS = ["p0", "p1", "p2"]
GAP = 2  # instructions required between an SGPR write and its read as a carry


class Statement:
    def __init__(self, incoming, c2_started):
        self.code = []
        # (pair, position of its write); carries pending from the previous statement
        # were written at least one instruction before this statement starts
        self.pending = [(p, -2) for p in incoming]
        self.free = [p for p in PAIRS if p not in incoming]
        self.c2_started = c2_started
        self.operands = {}

    def use(self, name, text):
        self.operands[name] = text

    def ready(self, w):
        return len(self.code) >= w + GAP + 1

    def add_carry(self):
        pair, w = self.pending.pop(0)
        while not self.ready(w):
            self.code.append("s_nop 0")
        if self.c2_started:
            self.code.append("v_addc_co_u32_e64 %[c2], %[dd], 0, %[c2], %[" + pair + "]")
        else:
            self.code.append("v_addc_co_u32_e64 %[c2], %[dd], 0, 0, %[" + pair + "]")
            self.c2_started = True
        self.free.append(pair)

    def drain_ready(self):
        while self.pending and self.ready(self.pending[0][1]):
            self.add_carry()

    def mad(self, x, y, carry):
        if not carry:
            self.code.append("v_mad_u64_u32 %[acc], %[dd], %[" + x + "], %[" + y + "], %[acc]")
        else:
            if not self.free:
                self.add_carry()
            pair = self.free.pop(0)
            self.code.append("v_mad_u64_u32 %[acc], %[" + pair + "], %[" + x + "], %[" + y + "], %[acc]")
            self.pending.append((pair, len(self.code) - 1))
        self.drain_ready()

    def finish(self, keep):
        while len(self.pending) > keep:
            self.add_carry()
        return [p for p, _ in self.pending]


def operand(name, idx, src):
    return "%s%d" % (name, idx), '[%s%d] "v"(%s.v[%d])' % (name, idx, src, idx)


def operand_m(i):
    return "m%d" % i, '[m%d] "v"(m[%d])' % (i, i)


def operand_n(j):
    return "n%d" % j, '[n%d] "s"(F::m[%d])' % (j, j)


def render(st, incoming, c2_was_started):
    """asm statement text; carries pending across statements travel as "+s" operands."""
    text = "\\n\\t".join(st.code)
    outs = ['[acc] "+v"(acc)']
    if c2_was_started:
        outs.append('[c2] "+v"(c2)')
    elif st.c2_started:
        outs.append('[c2] "=&v"(c2)')
    for p in PAIRS:
        if p in incoming:
            outs.append('[%s] "+s"(%s)' % (p, p))
        elif any(("%[" + p + "]") in line for line in st.code):
            outs.append('[%s] "=&s"(%s)' % (p, p))
    outs.append('[dd] "=&s"(dd)')
    ins = list(st.operands.values())
    return '  asm("%s"\n      : %s\n      : %s);' % (text, ", ".join(outs), ",\n        ".join(ins))


def emit_function(name, streams):
    """streams: the products' operand names, e.g. [("a", "b")] or [("a", "b"), ("u", "w")]."""
    params = ", ".join("const fe &%s" % x for pair in streams for x in pair)
    out = []
    out.append("template <class F> __device__ __forceinline__ fe %s(%s) {" % (name, params))
    out.append("  uint32_t m[8];")
    out.append("  fe r;")
    out.append("  uint64_t acc = 0;")
    out.append("  uint32_t c2 = 0;")
    out.append("  uint64_t p0, p1, p2, dd;")
    for k in range(15):
        low = k < 8
        out.append("  // column %d" % k)
        rng = range(0, k + 1) if low else range(k - 7, 8)
        prods = [(x, i, y, k - i) for (x, y) in streams for i in rng]
        mp = [(i, k - i) for i in (range(0, k) if low else range(k - 7, 8))]
        st = Statement([], False)

        def mad_prod(pr, carry):
            x, i, y, j = pr
            xa, ta = operand(x, i, x)
            xb, tb = operand(y, j, y)
            st.use(xa, ta)
            st.use(xb, tb)
            st.mad(xa, xb, carry=carry)

        def mad_mn(i, j, carry):
            xm, tm = operand_m(i)
            xn, tn = operand_n(j)
            st.use(xm, tm)
            st.use(xn, tn)
            st.mad(xm, xn, carry=carry)

        if mp:
            mad_mn(mp[0][0], mp[0][1], carry=False)
            for pr in prods:
                mad_prod(pr, carry=True)
            for i, j in mp[1:]:
                mad_mn(i, j, carry=True)
        else:  # column 0: the accumulator starts at zero
            for n, pr in enumerate(prods):
                mad_prod(pr, carry=n > 0)
        pend = st.finish(keep=2 if low else 0)
        out.append(render(st, [], False))
        started = st.c2_started
        if low:
            out.append("  m[%d] = (uint32_t)acc * F::inv;" % k)
            st2 = Statement(pend, started)
            xm, tm = operand_m(k)
            st2.use(xm, tm)
            xn, tn = operand_n(0)
            st2.use(xn, tn)
            st2.mad(xm, xn, carry=True)
            st2.finish(keep=0)
            out.append(render(st2, pend, started))
        else:
            out.append("  r.v[%d] = (uint32_t)acc;" % (k - 8))
        out.append("  acc = (acc >> 32) | ((uint64_t)c2 << 32);")
    out.append("  r.v[7] = (uint32_t)acc;")
    out.append("  (void)p0;")
    out.append("  (void)p1;")
    out.append("  (void)p2;")
    out.append("  (void)dd;")
    out.append("  return r;")
    out.append("}")
    return out


def emit():
    out = []
    out.append("// Generated by gen_mont_asm.py: 8 x 32-bit-limb Montgomery products")
    out.append("// (finely integrated product scanning) for gfx950.")
    out.append("// Included by field.cuh inside namespace bc, device compilation only.")
    out.append("")
    out += emit_function("mul_lazy_device", [("a", "b")])
    out.append("")
    out += emit_function("mul2_lazy_device", [("a", "b"), ("u", "w")])
    return "\n".join(out) + "\n"


def check_schedule(text):
    """Re-derives, from the emitted text, every carry read's distance from its write."""
    import re

    stmts = re.findall(r'asm\("([^"]*)"', text)
    assert len(stmts) == 2 * 23
    for idx, stmt in enumerate(stmts):
        lines = stmt.split("\\n\\t")
        last_write = {}
        for pos, line in enumerate(lines):
            if line.startswith("v_mad_u64_u32"):
                ops = re.findall(r"%\[(\w+)\]", line)  # acc, carry-out, x, y, acc
                pair = ops[1]
                last_write[pair] = pos
                if pair == "dd":  # carry dropped: an m * N product, or the first product of a function
                    assert (ops[2].startswith("m") and ops[3].startswith("n")) or (idx % 23 == 0 and pos == 0), line
            elif line.startswith("v_addc"):
                pair = re.search(r"%\[(p\d)\]$", line).group(1)
                w = last_write.get(pair, -2)  # written in the previous statement
                assert pos - w - 1 >= GAP, (stmt, pos, w)
    return True


if __name__ == "__main__":
    text = emit()
    assert check_schedule(text)
    print(text, end="")
