"""Generates fe8_asm.inc: GF(2^255-19) multiply and square in 8 x 32-bit limbs
as gfx950 inline assembly (v_mad_u64_u32 product scanning with carry chains).

    python gen_fe8.py > fe8_asm.inc

Scheme (both functions). Columns are scanned with a 64-bit accumulator held
in one of two fixed VGPR pairs (v[0:1], v[2:3]); each v_mad_u64_u32 adds one
32 x 32 product and sends its carry-out (an SGPR lane mask, three in rotation)
to a counter in the high half of the other pair, which becomes the next
column's accumulator after one move. A VALU write of an SGPR is read by a VALU
instruction no sooner than two instructions later (the generator schedules
that, padding with s_nop when nothing else is ready).

Reduction modulo 2^256 - 38 is folded into the scan: the high columns 8..14
are scanned first, without carry-in, into H' (the true high half H minus the
carry out of the low half, which the low scan keeps); then the low columns
0..7 are scanned with 38 * H'[k] as one more product each, so the low scan
computes V = L + 38 H' (L: the low columns with their own carries). Its
carry-out T < 2^36 + 38 (two words in the accumulator pair) is folded once
more as 38 T < 2^42, and a final wrap past 2^256 adds 38 (the value is then
below 2^42, so that addition stops within two words). These two chains carry
through VCC in the VOP2 forms, whose carry-in read of VCC right after its
write needs no wait states; only the v_cndmask that reads VCC as a lane mask
waits two.

The square forms each cross product once against the words of 2A
(a_i * 2 A_{>i}), so it needs 35 + 8 products instead of 64.
"""

import sys

RING = [(0, 1), (2, 3)]


class Gen:
    def __init__(self):
        self.ins = []
        self.pending = []   # (counter, sgpr, index of the mad)
        self.sg = 0
        self.cur = 0
        self.counter_used = False

    def emit(self, s):
        self.ins.append(s)

    def pair(self, k):
        lo, hi = RING[k]
        return "v[%d:%d]" % (lo, hi)

    def next_sgpr(self):
        s = "%%[c%d]" % self.sg
        self.sg = (self.sg + 1) % 3
        return s

    def counter_reg(self):
        return "v%d" % RING[1 - self.cur][1]

    def flush(self, force=False):
        keep = []
        for cnt, sg, at in self.pending:
            gap = len(self.ins) - 1 - at
            if force or gap >= 2:
                if gap < 2:
                    self.emit("s_nop %d" % (1 - gap))
                if self.counter_used:
                    self.emit("v_addc_co_u32_e64 %s, %%[cj], %s, 0, %s" % (cnt, cnt, sg))
                else:
                    self.emit("v_addc_co_u32_e64 %s, %%[cj], 0, 0, %s" % (cnt, sg))
                    self.counter_used = True
            else:
                keep.append((cnt, sg, at))
        self.pending = keep

    def column(self, products, carry_in):
        """products: list of (x, y) operand strings; carry_in: whether the
        current pair already holds the previous column's carry."""
        self.counter_used = False
        cnt = self.counter_reg()
        for idx, (x, y) in enumerate(products):
            acc = self.pair(self.cur)
            src2 = acc if (idx > 0 or carry_in) else "0"
            # flush carries ready before issuing the next mad, so the
            # SGPR ring slot is free again
            self.flush()
            if src2 == "0":
                self.emit("v_mad_u64_u32 %s, %%[cj], %s, %s, 0" % (acc, x, y))
            else:
                sg = self.next_sgpr()
                # never reuse an SGPR still pending
                while any(p[1] == sg for p in self.pending):
                    self.flush(force=True)
                self.emit("v_mad_u64_u32 %s, %s, %s, %s, %s" % (acc, sg, x, y, src2))
                self.pending.append((cnt, sg, len(self.ins) - 1))
        return cnt

    def transition(self, out):
        """Column done: its low word to `out`, next pair = (high word, counter)."""
        lo, hi = RING[self.cur]
        nlo, nhi = RING[1 - self.cur]
        self.emit("v_mov_b32 %s, v%d" % (out, lo))
        self.emit("v_mov_b32 v%d, v%d" % (nlo, hi))
        self.flush(force=True)
        if not self.counter_used:
            self.emit("v_mov_b32 v%d, 0" % nhi)
        self.cur = 1 - self.cur


def mul_products(k, a, b):
    return [(a(i), b(k - i)) for i in range(8) if 0 <= k - i < 8]


def gen_body(kind):
    g = Gen()
    if kind == "mul":
        A = lambda i: "%%[a%d]" % i
        B = lambda j: "%%[b%d]" % j
        prods = lambda k: mul_products(k, A, B)
    else:
        A = lambda i: "%%[a%d]" % i
        # words of 2A: e_j = a_j << 1 (lowest word of 2 A_{>i}), d_j = (a_j << 1) | (a_{j-1} >> 31)
        for j in range(1, 8):
            g.emit("v_lshlrev_b32 %%[e%d], 1, %%[a%d]" % (j, j))
        for j in range(2, 8):
            g.emit("v_alignbit_b32 %%[d%d], %%[a%d], %%[a%d], 31" % (j, j, j - 1))
        g.emit("v_lshrrev_b32 %[d8], 31, %[a7]")

        def prods(k):
            out = []
            for i in range(0, 7):
                j = k - i
                if j < i + 1 or j > 8:
                    continue
                out.append((A(i), "%%[e%d]" % j if j == i + 1 else "%%[d%d]" % j))
            if k % 2 == 0 and k // 2 < 8:
                out.append((A(k // 2), A(k // 2)))
            return out
    # high columns 8..14 (no carry-in), into h8..h15
    first = True
    for k in range(8, 15):
        g.column(prods(k), carry_in=not first)
        first = False
        g.transition("%%[h%d]" % k)
    # h15 = the accumulator's low word now (its high word is provably zero)
    g.flush(force=True)
    g.emit("v_mov_b32 %%[h15], v%d" % RING[g.cur][0])
    # low columns 0..7 with 38 * h[k+8] folded in
    g.cur = 0
    first = True
    for k in range(0, 8):
        ps = prods(k) + [("%%[h%d]" % (k + 8), "38")]
        g.column(ps, carry_in=not first)
        first = False
        g.transition("%%[r%d]" % k)
    g.flush(force=True)
    # T = (t_lo, t_hi) = carry-out of the low scan: r += 38 T, then a wrap adds 38.
    t_lo, t_hi = RING[g.cur]
    g.emit("v_mul_u32_u24 v4, v%d, 38" % t_hi)
    g.emit("v_mad_u64_u32 v[%d:%d], %%[cj], v%d, 38, 0" % (t_lo, t_hi, t_lo))
    g.emit("v_add_u32 v%d, v%d, v4" % (t_hi, t_hi))
    # the chain carries through VCC (VOP2 forms: a carry-in read of VCC
    # needs no wait states); v_cndmask reads it as a lane mask two later
    g.emit("v_add_co_u32_e32 %%[r0], vcc, v%d, %%[r0]" % t_lo)
    g.emit("v_addc_co_u32_e32 %%[r1], vcc, v%d, %%[r1], vcc" % t_hi)
    for i in range(2, 8):
        g.emit("v_addc_co_u32_e32 %%[r%d], vcc, 0, %%[r%d], vcc" % (i, i))
    g.emit("s_nop 1")
    g.emit("v_cndmask_b32_e64 v4, 0, 38, vcc")
    g.emit("v_add_co_u32_e32 %[r0], vcc, v4, %[r0]")
    g.emit("v_addc_co_u32_e32 %[r1], vcc, 0, %[r1], vcc")
    return g.ins


def emit_function(name, kind):
    body = gen_body(kind)
    outs = ["[r%d] \"=&v\"(r[%d])" % (i, i) for i in range(8)]
    outs += ["[h%d] \"=&v\"(h[%d])" % (k, k - 8) for k in range(8, 16)]
    if kind == "sq":
        outs += ["[e%d] \"=&v\"(e[%d])" % (j, j) for j in range(1, 8)]
        outs += ["[d%d] \"=&v\"(d[%d])" % (j, j) for j in range(2, 9)]
    outs += ["[c0] \"=&s\"(c0)", "[c1] \"=&s\"(c1)", "[c2] \"=&s\"(c2)", "[cj] \"=&s\"(cj)"]
    ins = ["[a%d] \"v\"(a[%d])" % (i, i) for i in range(8)]
    if kind == "mul":
        ins += ["[b%d] \"v\"(b[%d])" % (i, i) for i in range(8)]
    sig = ("LXV_D void %s(uint32_t r[8], const uint32_t a[8], const uint32_t b[8])" % name if kind == "mul"
           else "LXV_D void %s(uint32_t r[8], const uint32_t a[8])" % name)
    lines = [sig + " {", "    uint32_t h[8];"]
    if kind == "sq":
        lines.append("    uint32_t e[8], d[9];")
    lines.append("    uint64_t c0, c1, c2, cj;")
    lines.append("    asm volatile(")
    for s in body:
        lines.append("        \"%s\\n\"" % s)
    lines.append("        : " + ",\n          ".join(outs))
    lines.append("        : " + ",\n          ".join(ins))
    lines.append("        : \"v0\", \"v1\", \"v2\", \"v3\", \"v4\", \"vcc\");")
    lines.append("}")
    return "\n".join(lines), body


def main():
    out = ["// Generated by gen_fe8.py; do not edit. See that script for the scheme.", ""]
    stats = []
    for name, kind in (("fe8_mul_asm", "mul"), ("fe8_sq_asm", "sq")):
        text, body = emit_function(name, kind)
        out.append(text)
        out.append("")
        n_mad = sum(1 for s in body if s.startswith("v_mad_u64_u32"))
        n_nop = sum(1 for s in body if s.startswith("s_nop"))
        stats.append("// %s: %d instructions, %d v_mad_u64_u32, %d s_nop" % (name, len(body), n_mad, n_nop))
    out = out[:1] + stats + out[1:]
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main()
