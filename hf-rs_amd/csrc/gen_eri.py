"""Generate straight-line Head-Gordon-Pople ERI code for each shell-quartet class.

For a class (la lb | lc ld) with la >= lb, lc >= ld the generated device
function evaluates, per primitive quartet, the Obara-Saika vertical recurrence

  [a+1_i|c]^m = PA_i [a|c]^m + WP_i [a|c]^{m+1}
              + a_i/(2p) ([a-1_i|c]^m - q/(p+q) [a-1_i|c]^{m+1})          (c = 0)
  [a|c+1_i]^m = QC_i [a|c]^m + WQ_i [a|c]^{m+1}
              + c_i/(2q) ([a|c-1_i]^m - p/(p+q) [a|c-1_i]^{m+1})
              + a_i/(2(p+q)) [a-1_i|c]^{m+1}

from [0|0]^m = 2 pi^{5/2} K_ab K_cd / (p q sqrt(p+q)) F_m(T), accumulates the
contracted [e0|f0] for la<=|e|<=la+lb and lc<=|f|<=lc+ld in registers, and after
the primitive loop applies the horizontal recurrences

  (a, b+1_i| = (a+1_i, b| + AB_i (a, b|        |c, d+1_i) = |c+1_i, d) + CD_i |c, d)

to obtain the Cartesian block (ab|cd) in the hf-rs component order
(src/basis.rs:170-188, 430-436).  The algorithm differs from the reference's
McMurchie-Davidson evaluation (src/integrals.rs:259-346, 905-971) on purpose:
primitive work is limited to the vertical step and the horizontal step runs
once per contracted quartet.

Usage: python gen_eri.py OUTDIR   (writes eri_classes.inc)
"""
import itertools
import os
import sys

LMAX = 2


def cart(L):
    """Cartesian components in hf-rs order."""
    if L == 2:
        return [(2, 0, 0), (0, 2, 0), (0, 0, 2), (1, 1, 0), (1, 0, 1), (0, 1, 1)]
    return [(lx, ly, L - lx - ly) for lx in range(L, -1, -1) for ly in range(L - lx, -1, -1)]


def ncart(L):
    return (L + 1) * (L + 2) // 2


def sub(t, i, k=1):
    u = list(t)
    u[i] -= k
    return tuple(u)


def add(t, i, k=1):
    u = list(t)
    u[i] += k
    return tuple(u)


def tag(t):
    return "%d%d%d" % t


class Emitter:
    def __init__(self):
        self.lines = []
        self.nflops = 0

    def emit(self, s):
        self.lines.append(s)


def choose_axis(t, other):
    """Axis to lower in t: a component that is 1 drops the second-order term;
    ties broken towards the axis where `other` is zero (drops the coupling term)."""
    best, score = None, None
    for i in range(3):
        if t[i] == 0:
            continue
        s = (t[i] - 1 > 0) * 2 + (other[i] > 0)
        if score is None or s < score:
            best, score = i, s
    return best


class VRR:
    """Memoised generation of [a|c]^m nodes inside the primitive loop."""

    def __init__(self, em):
        self.em = em
        self.memo = {}

    def node(self, a, c, m):
        key = (a, c, m)
        if key in self.memo:
            return self.memo[key]
        if any(x < 0 for x in a) or any(x < 0 for x in c):
            return None
        em = self.em
        name = "v%s_%s_%d" % (tag(a), tag(c), m)
        if sum(a) == 0 and sum(c) == 0:
            name = "F%d" % m
            self.memo[key] = name
            return name
        terms = []
        if sum(c) > 0:
            i = choose_axis(c, a)
            c1 = sub(c, i)
            t0 = self.node(a, c1, m)
            t1 = self.node(a, c1, m + 1)
            ax = "xyz"[i]
            terms.append("QC%s*%s" % (ax, t0))
            terms.append("WQ%s*%s" % (ax, t1))
            if c1[i] > 0:
                c2 = sub(c1, i)
                u0 = self.node(a, c2, m)
                u1 = self.node(a, c2, m + 1)
                terms.append("%d.0*oo2q*(%s - rq*%s)" % (c1[i], u0, u1))
                em.nflops += 4
            if a[i] > 0:
                w = self.node(sub(a, i), c1, m + 1)
                terms.append("%d.0*oo2s*%s" % (a[i], w))
                em.nflops += 3
        else:
            i = choose_axis(a, c)
            a1 = sub(a, i)
            t0 = self.node(a1, c, m)
            t1 = self.node(a1, c, m + 1)
            ax = "xyz"[i]
            terms.append("PA%s*%s" % (ax, t0))
            terms.append("WP%s*%s" % (ax, t1))
            if a1[i] > 0:
                a2 = sub(a1, i)
                u0 = self.node(a2, c, m)
                u1 = self.node(a2, c, m + 1)
                terms.append("%d.0*oo2p*(%s - rp*%s)" % (a1[i], u0, u1))
                em.nflops += 4
        em.nflops += 3
        em.emit("      const double %s = %s;" % (name, " + ".join(terms)))
        self.memo[key] = name
        return name


def levels(lo, hi):
    out = []
    for L in range(lo, hi + 1):
        out += cart(L)
    return out


def hrr_lines(la, lb, lc, ld, acc, pre, base, keep=None, used=None):
    """Horizontal recurrences on contracted [e0|f0] values `acc` (names keyed
    by (e, f)), writing the Cartesian block to I[base ...]; `pre` keeps the
    temporaries of different contraction combinations apart.  With `keep`,
    only the outputs whose a and c component indices pass keep(ia, ic) are
    formed; `used` collects the (e, f) keys they reach."""
    hl = []
    memo = {}

    def bra(a, b, f):
        # (a b | f0)
        key = ("b", a, b, f)
        if key in memo:
            return memo[key]
        if sum(b) == 0:
            if used is not None:
                used.add((a, f))
            memo[key] = acc[(a, f)]
            return memo[key]
        i = choose_axis(b, (0, 0, 0))
        b1 = sub(b, i)
        x = bra(add(a, i), b1, f)
        y = bra(a, b1, f)
        name = "%shb%s_%s_%s" % (pre, tag(a), tag(b), tag(f))
        hl.append("  const double %s = %s + AB[%d] * %s;" % (name, x, i, y))
        memo[key] = name
        return name

    def ket(a, b, c, d):
        key = ("k", a, b, c, d)
        if key in memo:
            return memo[key]
        if sum(d) == 0:
            return bra(a, b, c)
        i = choose_axis(d, (0, 0, 0))
        d1 = sub(d, i)
        x = ket(a, b, add(c, i), d1)
        y = ket(a, b, c, d1)
        name = "%shk%s%s_%s_%s" % (pre, tag(a), tag(b), tag(c), tag(d))
        hl.append("  const double %s = %s + CD[%d] * %s;" % (name, x, i, y))
        memo[key] = name
        return name

    idx = base
    for ia, a in enumerate(cart(la)):
        for b in cart(lb):
            for ic, c in enumerate(cart(lc)):
                for d in cart(ld):
                    if keep is None or keep(ia, ic):
                        v = ket(a, b, c, d)
                        hl.append("  I[%d] = %s;" % (idx, v))
                    idx += 1
    return hl


def ket_prim_lines(la, lb, lc, ld):
    """Per primitive quartet inside the ket loop: geometry, Boys values scaled
    by the prefactor (F0..FM) and the recurrence coefficients."""
    M = la + lb + lc + ld
    o = []
    o.append("      const int ik = k0 + jk * ks;")
    o.append("      const double q = kp.p[ik];")
    o.append("      const double PQx = Px - kp.Px[ik], PQy = Py - kp.Py[ik], PQz = Pz - kp.Pz[ik];")
    o.append("      const double rs = rsqrt_nr(p + q);  // 1/sqrt(p+q)")
    o.append("      const double inv_s = rs * rs;")
    o.append("      const double rho = p * q * inv_s;")
    o.append("      const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);")
    o.append("      const double pref = TWO_PI_2_5 * Kp * kp.K[ik] * rs;")
    o.append("      double Fm[%d];" % (M + 1))
    o.append("      boys_eval<%d>(T, boys, Fm);" % M)
    for m in range(M + 1):
        o.append("      const double F%d = pref * Fm[%d];" % (m, m))
    if M > 0:
        if la + lb > 0:
            o.append("      const double rp = q * inv_s;")
            o.append("      const double WPx = -rp * PQx, WPy = -rp * PQy, WPz = -rp * PQz;")
        if lc + ld > 0:
            o.append("      const double QCx = kp.PAx[ik], QCy = kp.PAy[ik], QCz = kp.PAz[ik], oo2q = kp.H[ik];")
            o.append("      const double rq = p * inv_s, oo2s = 0.5 * inv_s;")
            o.append("      const double WQx = rq * PQx, WQy = rq * PQy, WQz = rq * PQz;")
    return o


def gen_class(la, lb, lc, ld):
    """Return C++ source of eri_<la><lb><lc><ld>(...)."""
    M = la + lb + lc + ld
    em = Emitter()
    E = levels(la, la + lb)
    Fk = levels(lc, lc + ld)
    fname = "eri_%d%d%d%d" % (la, lb, lc, ld)
    nab, ncd = ncart(la) * ncart(lb), ncart(lc) * ncart(ld)
    out = []
    out.append("// (%s%s|%s%s): %d contracted [e0|f0] accumulators" % ("spdfg"[la], "spdfg"[lb], "spdfg"[lc],
                                                                        "spdfg"[ld], len(E) * len(Fk)))
    out.append("template <> HFG_DEV void eri_block<%d,%d,%d,%d>(" % (la, lb, lc, ld))
    out.append("    const PrimPairs& bp, int b0, int b1, const PrimPairs& kp, int k0, int nk, int ks,")
    out.append("    const double AB[3], const double CD[3], const double* __restrict__ boys, double* __restrict__ I) {")
    acc = {}
    for e in E:
        for f in Fk:
            acc[(e, f)] = "c%s_%s" % (tag(e), tag(f))
    out.append("  double " + ", ".join("%s = 0.0" % v for v in acc.values()) + ";")
    # Bra primitives outer (uniform across the wave in the J/K kernel), ket
    # primitives inner (one ket pair per lane).  The opposite nesting -- a
    # wave-uniform loop inside the per-lane loop -- produced wrong J/K on gfx950
    # in the digestion kernel while the same source was exact on the CPU.
    out.append("  for (int ib = b0; ib < b1; ++ib) {")
    out.append("    const double p = bp.p[ib], Px = bp.Px[ib], Py = bp.Py[ib], Pz = bp.Pz[ib], Kp = bp.K[ib];")
    if M > 0 and la + lb > 0:
        out.append("    const double PAx = bp.PAx[ib], PAy = bp.PAy[ib], PAz = bp.PAz[ib], oo2p = bp.H[ib];")
    out.append("    for (int jk = 0; jk < nk; ++jk) {")
    out += ket_prim_lines(la, lb, lc, ld)
    vrr = VRR(em)
    for e in E:
        for f in Fk:
            n = vrr.node(e, f, 0)
            em.emit("      %s += %s;" % (acc[(e, f)], n))
    out += em.lines
    out.append("    }")
    out.append("  }")
    # Horizontal recurrences on the contracted values.
    out += hrr_lines(la, lb, lc, ld, acc, "", 0)
    out.append("}")
    return "\n".join(out) + "\n", em.nflops


# Contraction members evaluated together by the J/K kernel, per quartet
# class: (RA, RB | RC, RD).  A pair of generally contracted shells (several
# contractions over one exponent set) is split into views of at most RA
# members of its first shell and RB of its second; the CA = RA * RB bra and
# CC = RC * RD ket combinations of a view share every primitive integral.  The
# kernel holds CA x CC contracted blocks in registers, so the caps keep
# CA * CC * (accumulators per block) within about 96 doubles.  Classes not
# listed evaluate one combination.
GC_CAPS = {
    (0, 0, 0, 0): (3, 3, 3, 3),
    (1, 0, 0, 0): (1, 3, 3, 3),
    (1, 0, 1, 0): (1, 3, 1, 3),
    (1, 1, 0, 0): (1, 1, 3, 3),
    (1, 1, 1, 0): (1, 1, 1, 3),
    (2, 0, 0, 0): (1, 1, 3, 3),
    (2, 0, 1, 0): (1, 1, 1, 3),
    (2, 1, 0, 0): (1, 1, 1, 3),
}


def gc_caps(c, V=0):
    """Caps of class c in kernel variant V: 0 the caps above, 1 one member per
    shell (segmented bases, where no shells share exponents: the variant-0
    kernels would carry every combination's arithmetic for one nonzero)."""
    return (1, 1, 1, 1) if V else GC_CAPS.get(tuple(c), (1, 1, 1, 1))


def has_unit_variant(c):
    return gc_caps(c) != (1, 1, 1, 1)


# Output slices of the high-L classes.  A slice is the part of a contracted
# quartet's Cartesian block with the a component (mode 1), the c component
# (mode 2) or both (mode 3) fixed; each slice runs in its own wave with only
# the [e0|f0] accumulators and recurrence nodes its outputs reach.  Slices
# repeat the per-primitive setup (Boys values, prefactors) and some of the
# vertical recurrences (x2-x10 for these classes) but hold a fraction of the
# registers: unsliced, (dp|dp) keeps 256 contracted accumulators, (dd|dd) 961,
# far beyond the 512 registers of a lane, and the spills bound those classes
# by memory latency at one wave per SIMD.  Classes with many primitive
# quartets per contracted quartet and little spilling ((pp|pp), (dp|pp),
# (dp|ds), (dd|ps)) were measured slower sliced and stay whole.  Among the
# modes, the one that keeps the costly digestion sums whole wins: per c
# component for (dp|dp) (its ket Coulomb sums, global atomics, stay one per
# output; 0.78 -> 0.42 ms on (H2O)16/cc-pVDZ against per a), per a component
# for (dd|pp) (6 slices instead of 18 (a, c) pairs; 0.43 -> 0.28 ms), per c
# for (dd|dp) (6 instead of 36; 0.46 -> 0.28 ms), per a for (dd|ds).
SLICE_MODES = {
    (2, 1, 2, 1): 2,
    (2, 2, 1, 1): 1,
    (2, 2, 2, 0): 1,
    (2, 2, 2, 1): 2,
    (2, 2, 2, 2): 3,
}


# Classes whose bra-primitive loop is unrolled by two: two independent Boys
# evaluations per iteration let the scheduler overlap the table reads of one
# primitive with the arithmetic of the other (at one or two waves per SIMD
# the low-L classes otherwise wait on those reads most of the time).
UNROLL_BRA = {(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (1, 1, 1, 0)}


def slice_mode(c):
    return SLICE_MODES.get(tuple(c), 0)


def slice_count(c):
    m = slice_mode(c)
    na, nc = ncart(c[0]), ncart(c[2])
    return {0: 1, 1: na, 2: nc, 3: na * nc}[m]


def slice_keep(c, S):
    """keep(ia, ic) of slice S of class c (None: the whole block)."""
    m = slice_mode(c)
    nc = ncart(c[2])
    if m == 0:
        return None
    if m == 1:
        return lambda ia, ic: ia == S
    if m == 2:
        return lambda ia, ic: ic == S
    return lambda ia, ic: ia * nc + ic == S


# Bra primitive record of the J/K kernel: 16 doubles, read with wide scalar
# loads from one pointer: p, K, Px, Py, Pz, then for pair classes beyond ss
# PAx, PAy, PAz, 1/(2p); the contraction coefficients of the view's
# combinations follow (from BRA_COEF_SS or BRA_COEF).
BRA_REC = 16
BRA_COEF_SS = 5
BRA_COEF = 9


def gen_class_gc(la, lb, lc, ld, S=0, V=0):
    """eri_gc<la,lb,lc,ld,S,V>: slice S of the blocks of CA x CC contraction combinations of one
    bra pair and one ket pair that share their primitive pairs.

    Bra primitives are records of BRA_REC doubles at br (wave-uniform, read
    through scalar loads), ket primitives the transposed arrays kp (one pair
    per lane).  Primitive data carry no contraction coefficients when a side
    has more than one combination (cap > 1): the bra coefficients are in the
    record, the ket ones at kc[(jk * CC + g) * ks].  With cap 1 the single
    combination's coefficient is folded into K as in eri_block.  Ket
    primitives are the outer loop, bra primitives the inner one: per ket
    primitive the bra sums t[gab] = sum_b c_b[gab] [e0|f0]_b are formed, then
    a[gab][gcd] += c_k[gcd] t[gab].  Output: I[(gab * CC + gcd) * NAB * NCD + ...],
    only the entries of slice S (slice_keep) written."""
    RA, RB, RC, RD = gc_caps((la, lb, lc, ld), V)
    CA, CC = RA * RB, RC * RD
    coff = BRA_COEF_SS if la + lb == 0 else BRA_COEF
    assert CA == 1 or coff + CA <= BRA_REC
    M = la + lb + lc + ld
    em = Emitter()
    E = levels(la, la + lb)
    Fk = levels(lc, lc + ld)
    nab, ncd = ncart(la) * ncart(lb), ncart(lc) * ncart(ld)
    keep = slice_keep((la, lb, lc, ld), S)
    used = set()
    hrr_lines(la, lb, lc, ld, {(e, f): "x" for e in E for f in Fk}, "", 0, keep, used)
    E_f = [(e, f) for e in E for f in Fk if (e, f) in used]  # accumulators the slice reaches
    out = []
    out.append("// (%s%s|%s%s) x %d x %d combinations, slice %d of %d, variant %d: %d accumulators per combination" % (
        "spdfg"[la], "spdfg"[lb], "spdfg"[lc], "spdfg"[ld], CA, CC, S, slice_count((la, lb, lc, ld)), V, len(E_f)))
    out.append("template <> HFG_DEV void eri_gc<%d,%d,%d,%d,%d,%d>(" % (la, lb, lc, ld, S, V))
    out.append("    const double* __restrict__ br, int b0, int b1, const PrimPairs& kp,")
    out.append("    const double* __restrict__ kc, int k0, int nk, int ks, const double AB[3], const double CD[3],")
    out.append("    const double* __restrict__ boys, double* __restrict__ I) {")
    acc = {}
    for ga in range(CA):
        for gc in range(CC):
            for e, f in E_f:
                acc[(ga, gc, e, f)] = "a%d_%d_%s_%s" % (ga, gc, tag(e), tag(f))
    out.append("  double " + ", ".join("%s = 0.0" % v for v in acc.values()) + ";")
    # Ket primitives outer (one ket pair per lane; every lane runs the same
    # count, idle lanes on a valid pair, masked at digestion), bra primitives
    # inner (wave-uniform: scalar loads), so each ket primitive and its
    # coefficients are loaded once per quartet.  The next bra record is loaded
    # after the Boys table reads of the current one (software prefetch: one
    # wave per SIMD hides no latency, and scalar and LDS loads share a counter).
    bra_fields = [("p", 0), ("Kp", 1), ("Bx", 2), ("By", 3), ("Bz", 4)]
    if M > 0 and la + lb > 0:
        bra_fields += [("PAx", 5), ("PAy", 6), ("PAz", 7), ("oo2p", 8)]
    if CA > 1:
        bra_fields += [("cb%d" % g, coff + g) for g in range(CA)]
    def bra_loads(rec, pre):
        return ["%s%s = %s[%d]" % (pre, v, rec, k) for v, k in bra_fields]
    out.append("  const cdptr rec0 = (cdptr)br;")
    out.append("  for (int jk = 0; jk < nk; ++jk) {")
    out.append("    const int ik = k0 + jk * ks;")
    out.append("    const double q = kp.p[ik], Qx = kp.Px[ik], Qy = kp.Py[ik], Qz = kp.Pz[ik], Kq = kp.K[ik];")
    if M > 0 and lc + ld > 0:
        out.append("    const double QCx = kp.PAx[ik], QCy = kp.PAy[ik], QCz = kp.PAz[ik], oo2q = kp.H[ik];")
    if CC > 1:
        out.append("    const double " + ", ".join("ck%d = kc[(jk * %d + %d) * ks]" % (g, CC, g) for g in range(CC)) + ";")
        tmp = {}
        for ga in range(CA):
            for e, f in E_f:
                tmp[(ga, e, f)] = "t%d_%s_%s" % (ga, tag(e), tag(f))
        out.append("    double " + ", ".join("%s = 0.0" % v for v in tmp.values()) + ";")
        tgt = lambda ga, e, f: tmp[(ga, e, f)]
    else:
        tgt = lambda ga, e, f: acc[(ga, 0, e, f)]
    out.append("    const cdptr rb = rec0 + (size_t)b0 * %d;" % BRA_REC)
    out.append("    double " + ", ".join(bra_loads("rb", "n_")) + ";")
    if (la, lb, lc, ld) in UNROLL_BRA:
        out.append("#pragma unroll 2")
    out.append("    for (int ib = b0; ib < b1; ++ib) {")
    names = [v for v, k in bra_fields]
    out.append("      const double " + ", ".join("%s = n_%s" % (v, v) for v in names) + ";")
    out.append("      const cdptr rn = rec0 + (size_t)(ib + 1 < b1 ? ib + 1 : ib) * %d;" % BRA_REC)
    out.append("      const double PQx = Bx - Qx, PQy = By - Qy, PQz = Bz - Qz;")
    out.append("      const double rs = rsqrt_nr(p + q);  // 1/sqrt(p+q)")
    out.append("      const double inv_s = rs * rs;")
    out.append("      const double rho = p * q * inv_s;")
    out.append("      const double T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);")
    out.append("      const double pref = TWO_PI_2_5 * Kp * Kq * rs;")
    out.append("      double Fm[%d];" % (M + 1))
    out.append("      boys_eval<%d, %s, %s>(T, boys, Fm);" % (M, "true" if (la, lb, lc, ld) in UNROLL_BRA else "false",
                                                      "true" if V == 1 and M > 0 else "false"))  # boys_compact(V, M)
    for m in range(M + 1):
        out.append("      const double F%d = pref * Fm[%d];" % (m, m))
    out.append("      " + "; ".join(bra_loads("rn", "n_")) + ";")
    if M > 0:
        if la + lb > 0:
            out.append("      const double rp = q * inv_s;")
            out.append("      const double WPx = -rp * PQx, WPy = -rp * PQy, WPz = -rp * PQz;")
        if lc + ld > 0:
            out.append("      const double rq = p * inv_s, oo2s = 0.5 * inv_s;")
            out.append("      const double WQx = rq * PQx, WQy = rq * PQy, WQz = rq * PQz;")
    vrr = VRR(em)
    for e, f in E_f:
        n = vrr.node(e, f, 0)
        for ga in range(CA):
            if CA > 1:
                em.emit("      %s += cb%d * %s;" % (tgt(ga, e, f), ga, n))
            else:
                em.emit("      %s += %s;" % (tgt(ga, e, f), n))
    out += em.lines
    out.append("    }")
    if CC > 1:
        for ga in range(CA):
            for gc in range(CC):
                for e, f in E_f:
                    out.append("    %s += ck%d * %s;" % (acc[(ga, gc, e, f)], gc, tmp[(ga, e, f)]))
    out.append("  }")
    for ga in range(CA):
        for gc in range(CC):
            sub_acc = {(e, f): acc[(ga, gc, e, f)] for e, f in E_f}
            out += hrr_lines(la, lb, lc, ld, sub_acc, "g%d_%d_" % (ga, gc), (ga * CC + gc) * nab * ncd, keep)
    out.append("}")
    return "\n".join(out) + "\n"


# Pair classes: every (la, lb) with lb <= la <= 2 (orbital shell pairs up to
# d), then (3,0) and (4,0): an auxiliary f or g shell paired with the
# exponent-0 unit function of the density-fitting integrals (src/df.rs:30-40).
PAIR_CLASSES = [(a, b) for a in range(LMAX + 1) for b in range(a + 1)] + [(3, 0), (4, 0)]


def classes():
    out = []
    for i, (la, lb) in enumerate(PAIR_CLASSES):
        for (lc, ld) in PAIR_CLASSES[:i + 1]:
            out.append((la, lb, lc, ld))
    return out


def grad_classes():
    """Block classes the nuclear gradient needs beyond classes(): a d shell
    raised to f in one pair (fp, fd, in kernel orientation) against every
    orbital pair; the other raised and lowered pairs fall in classes()."""
    orbital = [c for c in PAIR_CLASSES if c[0] <= LMAX]
    return [(3, b, lc, ld) for b in (1, 2) for (lc, ld) in orbital]


def df_f_classes():
    """Block classes of the density-fitting 3-index integrals with an orbital
    f shell paired with p, d or f (fp, fd, ff, kernel orientation) against an
    auxiliary shell with the unit function (s..g); the f-s orbital pairs fall
    in classes()."""
    return [(3, b, lc, 0) for b in (1, 2, 3) for lc in range(5)]


def main(outdir):
    os.makedirs(outdir, exist_ok=True)
    parts = ["// Generated by gen_eri.py -- do not edit.\n"]
    report = []
    for c in classes():
        src, fl = gen_class(*c)
        parts.append(src)
        report.append("%s %d" % ("".join(map(str, c)), fl))
    with open(os.path.join(outdir, "eri_classes.inc"), "w") as f:
        f.write("\n".join(parts))
    jk = classes()[:21]  # orbital classes (bra and ket up to dd): the J/K kernel's
    with open(os.path.join(outdir, "eri_gc.inc"), "w") as f:
        f.write("// Generated by gen_eri.py -- do not edit.\n")
        for c in jk:
            for S in range(slice_count(c)):
                f.write(gen_class_gc(*c, S=S))
            if has_unit_variant(c):
                assert slice_count(c) == 1
                f.write(gen_class_gc(*c, S=0, V=1))
    with open(os.path.join(outdir, "eri_grad.inc"), "w") as f:
        f.write("// Generated by gen_eri.py -- do not edit.  Block classes of the gradient (grad.hip).\n")
        for c in grad_classes():
            src, fl = gen_class(*c)
            f.write(src)
            report.append("%s %d" % ("".join(map(str, c)), fl))
        f.write("#define HFG_FOR_EACH_GRAD_CLASS(X) " + " ".join("X(%d, %d, %d, %d)" % c for c in grad_classes()) + "\n")
    with open(os.path.join(outdir, "eri_df.inc"), "w") as f:
        f.write("// Generated by gen_eri.py -- do not edit.  3-index classes with orbital f pairs (df_f.hip).\n")
        for c in df_f_classes():
            src, fl = gen_class(*c)
            f.write(src)
            report.append("%s %d" % ("".join(map(str, c)), fl))
        f.write("#define HFG_FOR_EACH_DF_F_CLASS(X) " + " ".join("X(%d, %d, %d, %d)" % c for c in df_f_classes()) + "\n")
    with open(os.path.join(outdir, "eri_classes.txt"), "w") as f:
        f.write("\n".join(report) + "\n")
    # class list for the C++ side, in quartet_class() order
    with open(os.path.join(outdir, "class_list.inc"), "w") as f:
        f.write("// Generated by gen_eri.py: X(la,lb,lc,ld) for every class, quartet_class() order.\n")
        f.write("#define HFG_FOR_EACH_CLASS_ALL(X) " + " ".join("X(%d, %d, %d, %d)" % c for c in classes()) + "\n")
        f.write("#define HFG_NCLASS_ALL %d\n" % len(classes()))
        f.write("// Contraction-member caps (RA, RB, RC, RD) of the J/K classes, quartet_class() order.\n")
        f.write("#define HFG_GC_CAPS " + ", ".join("{%d, %d, %d, %d}" % gc_caps(c) for c in jk) + "\n")
        f.write("// Classes with a one-member (segmented-basis) kernel variant, bit per quartet_class() index.\n")
        f.write("#define HFG_UNIT_VARIANT_MASK %#xu\n" % sum(1 << i for i, c in enumerate(jk) if has_unit_variant(c)))
        f.write("// Output-slice mode of the J/K classes (0 whole block, 1 per a component, 2 per c, 3 per (a, c)).\n")
        f.write("#define HFG_SLICE_MODES " + ", ".join("%d" % slice_mode(c) for c in jk) + "\n")
        f.write("// Bra primitive record of the J/K kernels: doubles per record, first coefficient (ss / other pairs).\n")
        f.write("#define HFG_BRA_REC %d\n#define HFG_BRA_COEF_SS %d\n#define HFG_BRA_COEF %d\n" % (BRA_REC, BRA_COEF_SS, BRA_COEF))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "gen"))
