"""Generates pair_table.inc: the affine points T_i = i Q, Q = 8B, of the pair
walk (lxv_pairwalk.h) on Curve25519 in Montgomery form v^2 = u^3 + A u^2 + u,
A = 486662, as canonical little-endian 32-bit words.

    python gen_pair_table.py > pair_table.inc

Q = 8B is the walk step (the reference's walk steps the private key by 8,
src/main.rs:124-137): u(Q) = X25519(8, 9), and v(Q) is the square root of
u^3 + A u^2 + u whose low bit is 0 (either root works; the walk only needs
every point it uses to be a multiple of the same Q, and the start-point
recovery uses this v). T_1 = Q, T_2 by the doubling formula, T_{i+1} =
T_i + Q by the affine addition formula. Entries 0..127 hold T_1..T_128 and
entry 128 holds T_256 (the largest batch's centre step).

Everything is Python big-integer arithmetic from the curve equation; the
table is checked against an independent x-only ladder by the CPU tests.
"""

import sys

P = 2**255 - 19
A = 486662


def inv(v):
    return pow(v % P, P - 2, P)


def sqrt(a):
    # p = 5 (mod 8): a^((p+3)/8) is a root of a or of -a; fix with sqrt(-1)
    r = pow(a, (P + 3) // 8, P)
    if r * r % P != a % P:
        r = r * pow(2, (P - 1) // 4, P) % P
    if r * r % P != a % P:
        raise ValueError("not a square")
    return r


def ladder_u(k, u):
    """u(kP) for u(P) = u by the RFC 7748 ladder over all bits of k."""
    x1, x2, z2, x3, z3 = u, 1, 0, u, 1
    for t in reversed(range(k.bit_length())):
        if (k >> t) & 1:
            x2, x3, z2, z3 = x3, x2, z3, z2
        a, b = (x2 + z2) % P, (x2 - z2) % P
        aa, bb = a * a % P, b * b % P
        e = (aa - bb) % P
        c, d = (x3 + z3) % P, (x3 - z3) % P
        da, cb = d * a % P, c * b % P
        x3, z3 = (da + cb) ** 2 % P, u * (da - cb) ** 2 % P
        x2, z2 = aa * bb % P, e * (aa + 121665 * e) % P
        if (k >> t) & 1:
            x2, x3, z2, z3 = x3, x2, z3, z2
    return x2 * inv(z2) % P


def affine_add(p1, p2):
    (x1, y1), (x2, y2) = p1, p2
    if x1 == x2:
        lam = (3 * x1 * x1 + 2 * A * x1 + 1) * inv(2 * y1) % P
    else:
        lam = (y2 - y1) * inv(x2 - x1) % P
    x3 = (lam * lam - A - x1 - x2) % P
    return x3, (lam * (x1 - x3) - y1) % P


def q_point():
    xq = ladder_u(8, 9)
    yq = sqrt((xq ** 3 + A * xq * xq + xq) % P)
    if yq & 1:
        yq = P - yq
    return xq, yq


def table():
    q = q_point()
    pts = [q]
    while len(pts) < 256:
        pts.append(affine_add(pts[-1], q))
    return pts[:128] + [pts[255]]


def words(v):
    return ["0x%08xu" % ((v >> (32 * i)) & 0xffffffff) for i in range(8)]


def main(out=sys.stdout):
    pts = table()
    out.write("// Generated by gen_pair_table.py (do not edit): T_i = i Q, Q = 8B, affine\n")
    out.write("// (u, v) on v^2 = u^3 + 486662 u^2 + u as canonical little-endian words;\n")
    out.write("// entry i - 1 holds T_i for i <= 128, entry 128 holds T_256.\n")
    out.write("#define LXV_PAIR_TABLE_ENTRIES %d\n" % len(pts))
    out.write("#define LXV_PAIR_TABLE_INIT { \\\n")
    for i, (x, y) in enumerate(pts):
        sep = "," if i + 1 < len(pts) else ""
        out.write("    %s, \\\n" % ", ".join(words(x)))
        out.write("    %s%s \\\n" % (", ".join(words(y)), sep))
    out.write("}\n")


if __name__ == "__main__":
    main()
