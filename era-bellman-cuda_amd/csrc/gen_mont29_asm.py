"""Generates mont29_asm.h: the 9 x 29-bit-limb Montgomery products of
field29.cuh on gfx950 (mul29, mul2_29, sqr29, and pairs of independent
products), one inline-asm statement per column.

With 29-bit limbs a column of a*b (+ c*d) + q*N never leaves 64 bits, so each
column is a strict chain of v_mad_u64_u32 into one VGPR pair whose carry-outs
go to a scratch SGPR pair that nothing reads.  The compiler's own code for the
same C++ splits the chain into several accumulators and merges them with
64-bit adds and moves (932 modelled issue cycles per product against 784
here; 166 against 186 G products/s measured on MI355X).

Column k's statement first finishes column k - 1: for a reduction column it
adds q_(k-1) * N_0, then it shifts the accumulator right by 29 and adds column
k's products.  The reduction digit q_k = (lo(acc) * inv) mod 2^29 and the
result limbs are taken in C++ between statements.

A function may compute several independent products ("lanes"): each column's
statement then alternates between the lanes' chains instruction by
instruction, so a wave always has a second chain to issue from while a
multiply's result is in flight (the compiler never interleaves statements of
different products by itself).

Run: python3 gen_mont29_asm.py > mont29_asm.h
"""

# A lane is ("mul", [(x, y), ...]): the sum of the products x*y, or
# ("sqr", x): x*x with the cross products taken once against 2x.


def lane_code(li, lane, k):
    """(instructions, {operand: constraint}) of lane li's part of column k."""
    kind, arg = lane
    code, ins = [], {}
    acc = "acc%d" % li

    def use(key, text):
        ins[key] = text

    def mad(x, y, first=False):
        if first:
            code.append("v_mad_u64_u32 %%[%s], %%[dd], %%[%s], %%[%s], 0" % (acc, x, y))
        else:
            code.append("v_mad_u64_u32 %%[%s], %%[dd], %%[%s], %%[%s], %%[%s]" % (acc, x, y, acc))

    low = k < 9
    if k > 0:
        if k - 1 < 9:
            use("l%dq%d" % (li, k - 1), '[l%dq%d] "v"(q%d[%d])' % (li, k - 1, li, k - 1))
            use("n0", '[n0] "s"(N.v[0])')
            mad("l%dq%d" % (li, k - 1), "n0")
        code.append("v_lshrrev_b64 %%[%s], 29, %%[%s]" % (acc, acc))
    rng = range(0, k + 1) if low else range(k - 8, 9)
    terms = []
    if kind == "sqr":
        x = arg
        for i in rng:
            j = k - i
            if j > i and j < 9:
                terms.append(("%s2x" % x, i, x, j))
        if k % 2 == 0 and k // 2 < 9:
            terms.append((x, k // 2, x, k // 2))
    else:
        for (x, y) in arg:
            for i in rng:
                terms.append((x, i, y, k - i))
    for n, (x, i, y, j) in enumerate(terms):
        xa, xb = "l%d%s%d" % (li, x, i), "l%d%s%d" % (li, y, j)
        use(xa, '[%s] "v"(%s.v[%d])' % (xa, x, i))
        use(xb, '[%s] "v"(%s.v[%d])' % (xb, y, j))
        mad(xa, xb, first=(k == 0 and n == 0))
    qr = range(0, k) if low else range(k - 8, 9)
    for i in qr:
        j = k - i
        use("l%dq%d" % (li, i), '[l%dq%d] "v"(q%d[%d])' % (li, i, li, i))
        use("n%d" % j, '[n%d] "s"(N.v[%d])' % (j, j))
        mad("l%dq%d" % (li, i), "n%d" % j)
    return code, ins


def emit_function(name, params, lanes, outputs):
    """lanes[i] is written to outputs[i]; a single-lane function returns it."""
    single = len(lanes) == 1
    if single:
        head = "template <class F> __device__ __forceinline__ f29 %s(%s) {" % (name, params)
    else:
        head = "template <class F> __device__ __forceinline__ void %s(%s) {" % (name, params)
    out = [head, "  constexpr c29::limbs N = P29<F>::n;", "  uint64_t dd;"]
    for li, lane in enumerate(lanes):
        out.append("  uint32_t q%d[9];" % li)
        out.append("  uint64_t acc%d;" % li)
        if lane[0] == "sqr":
            x = lane[1]
            out.append("  f29 %s2x;" % x)
            out.append("#pragma unroll")
            out.append("  for (int i = 0; i < 9; i++)")
            out.append("    %s2x.v[i] = %s.v[i] << 1;" % (x, x))
    if single:
        out.append("  f29 %s;" % outputs[0])
    for k in range(17):
        parts = [lane_code(li, lane, k) for li, lane in enumerate(lanes)]
        code, ins = [], {}
        for p in parts:
            ins.update(p[1])
        longest = max(len(p[0]) for p in parts)
        for n in range(longest):
            for p in parts:
                if n < len(p[0]):
                    code.append(p[0][n])
        outs = ", ".join(('[acc%d] "=&v"(acc%d)' if k == 0 else '[acc%d] "+v"(acc%d)') % (li, li) for li in range(len(lanes)))
        text = "\\n\\t".join(code)
        out.append("  // column %d" % k)
        out.append('  asm("%s"\n      : %s, [dd] "=&s"(dd)\n      : %s);' % (text, outs, ",\n        ".join(ins.values())))
        for li in range(len(lanes)):
            if k < 9:
                out.append("  q%d[%d] = ((uint32_t)acc%d * P29<F>::inv) & MASK29;" % (li, k, li))
            else:
                out.append("  %s.v[%d] = (uint32_t)acc%d & MASK29;" % (outputs[li], k - 9, li))
    for li in range(len(lanes)):
        out.append("  %s.v[8] = (uint32_t)(acc%d >> 29);" % (outputs[li], li))
    out.append("  (void)dd;")
    if single:
        out.append("  return %s;" % outputs[0])
    out.append("}")
    return out


def emit():
    out = []
    out.append("// Generated by gen_mont29_asm.py: 9 x 29-bit-limb Montgomery products for gfx950")
    out.append("// (one strict v_mad_u64_u32 chain per column and product).  Included by")
    out.append("// field29.cuh inside namespace bc, device compilation only.")
    out.append("")
    out += emit_function("mul29_device", "const f29 &a, const f29 &b", [("mul", [("a", "b")])], ["t"])
    out.append("")
    out += emit_function("mul2_29_device", "const f29 &a, const f29 &b, const f29 &c, const f29 &e", [("mul", [("a", "b"), ("c", "e")])], ["t"])
    out.append("")
    out += emit_function("sqr29_device", "const f29 &a", [("sqr", "a")], ["t"])
    out.append("")
    out += emit_function("mul29x2_device", "const f29 &a, const f29 &b, const f29 &c, const f29 &e, f29 &r, f29 &s",
                         [("mul", [("a", "b")]), ("mul", [("c", "e")])], ["r", "s"])
    out.append("")
    out += emit_function("sqr29x2_device", "const f29 &a, const f29 &c, f29 &r, f29 &s", [("sqr", "a"), ("sqr", "c")], ["r", "s"])
    out.append("")
    out += emit_function("mul2_mul_mul29_device",
                         "const f29 &a, const f29 &b, const f29 &c, const f29 &e, const f29 &f, const f29 &g, const f29 &h, const f29 &k, "
                         "f29 &r, f29 &s, f29 &t",
                         [("mul", [("a", "b"), ("c", "e")]), ("mul", [("f", "g")]), ("mul", [("h", "k")])], ["r", "s", "t"])
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    print(emit(), end="")
