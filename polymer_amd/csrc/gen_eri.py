"""Generate straight-line Head-Gordon-Pople ERI code for every shell-quartet
class (la lb|lc ld) with la >= lb, lc >= ld, (la,lb) >= (lc,ld), l <= LMAX.

For each class two device functions are emitted into eri_generated.h:

  vrr_<la><lb><lc><ld>(...)   one primitive quartet: Obara-Saika vertical
                              recursion from F_m(T) to [e0|f0]^(0) for
                              e in [la, la+lb], f in [lc, lc+ld], summed into
                              the contracted accumulator acc[].
  hrr_<la><lb><lc><ld>(...)   after the primitive loop: horizontal transfer
                              (e0|f0) -> (ab|cd) on the ket, then the bra.

The same recursion as src/pfourindexdirect.f90:554-690 (Obara-Saika with
the W-centre terms), reorganised as VRR to [e0|f0] plus HRR, so the
per-primitive work only builds the e+f ladder and the transfer to the
final components runs once per contracted quartet.

Cartesian component order inside a shell of angular momentum l is
x^i y^j z^k with i descending, then j descending (xx, xy, xz, yy, yz, zz).
The code is templated on a scalar type `real` so the same recursion can be
instantiated in double and in float.
"""
import os
import sys

LMAX = 2
# classes with more integrals per quartet than this run in the cooperative
# LDS kernel (POLYERI_COOP_MIN_INTEGRALS in eri_kernels.h)
COOP_MIN_INTEGRALS = 108


def comps(l):
    out = []
    for i in range(l, -1, -1):
        for k in range(0, l - i + 1):
            out.append((i, l - i - k, k))
    return out


def ncart(l):
    return (l + 1) * (l + 2) // 2


def name(v):
    return "".join(str(x) for x in v)


def sub(v, i, n=1):
    w = list(v)
    w[i] -= n
    return tuple(w)


def add(v, i, n=1):
    w = list(v)
    w[i] += n
    return tuple(w)


def pick_dir(v):
    """Reduce along the direction with the smallest non-zero exponent: it
    avoids the (n-1) two-step term whenever possible."""
    best, bi = None, -1
    for i in range(3):
        if v[i] > 0 and (best is None or v[i] < best):
            best, bi = v[i], i
    return bi


XYZ = "xyz"


class Vrr:
    def __init__(self):
        self.lines = []
        self.done = {}

    def node(self, e, f, m):
        key = (e, f, m)
        if key in self.done:
            return self.done[key]
        if sum(e) == 0 and sum(f) == 0:
            nm = f"F[{m}]"
            self.done[key] = nm
            return nm
        nm = f"v_{name(e)}_{name(f)}_{m}"
        terms = []
        if sum(f) == 0:
            i = pick_dir(e)
            e1 = sub(e, i)
            a = self.node(e1, f, m)
            b = self.node(e1, f, m + 1)
            terms.append(f"PA{XYZ[i]} * {a}")
            terms.append(f"WP{XYZ[i]} * {b}")
            if e[i] >= 2:
                e2 = sub(e, i, 2)
                c = self.node(e2, f, m)
                d = self.node(e2, f, m + 1)
                terms.append(f"real({e[i] - 1}) * oo2z * ({c} - roz * {d})")
        else:
            i = pick_dir(f)
            f1 = sub(f, i)
            a = self.node(e, f1, m)
            b = self.node(e, f1, m + 1)
            terms.append(f"QC{XYZ[i]} * {a}")
            terms.append(f"WQ{XYZ[i]} * {b}")
            if f[i] >= 2:
                f2 = sub(f, i, 2)
                c = self.node(e, f2, m)
                d = self.node(e, f2, m + 1)
                terms.append(f"real({f[i] - 1}) * oo2e * ({c} - roe * {d})")
            if e[i] >= 1:
                g = self.node(sub(e, i), f1, m + 1)
                terms.append(f"real({e[i]}) * oo2ze * {g}")
        self.lines.append(f"    const real {nm} = " + " + ".join(terms) + ";")
        self.done[key] = nm
        return nm


def gen_class(la, lb, lc, ld):
    L = la + lb + lc + ld
    ecomps = [c for l in range(la, la + lb + 1) for c in comps(l)]
    fcomps = [c for l in range(lc, lc + ld + 1) for c in comps(l)]
    tag = f"{la}{lb}{lc}{ld}"
    v = Vrr()
    acc_lines = []
    for ie, e in enumerate(ecomps):
        for jf, f in enumerate(fcomps):
            nm = v.node(e, f, 0)
            acc_lines.append(f"    acc[{ie * len(fcomps) + jf}] += {nm};")
    nacc = len(ecomps) * len(fcomps)
    out = []
    out.append(f"// class ({'spdf'[la]}{'spdf'[lb]}|{'spdf'[lc]}{'spdf'[ld]}), L = {L}, "
               f"{nacc} contracted [e0|f0] accumulators")
    out.append("template <typename real>")
    out.append(f"POLY_HD void vrr_{tag}(")
    out.append("    const real PAx, const real PAy, const real PAz, const real WPx, const real WPy, const real WPz,")
    out.append("    const real QCx, const real QCy, const real QCz, const real WQx, const real WQy, const real WQz,")
    out.append("    const real oo2z, const real oo2e, const real oo2ze, const real roz, const real roe,")
    out.append("    const real* __restrict__ F, real* __restrict__ acc) {")
    used = set()
    body = "\n".join(v.lines + acc_lines)
    for s in ["PAx", "PAy", "PAz", "WPx", "WPy", "WPz", "QCx", "QCy", "QCz", "WQx", "WQy", "WQz",
              "oo2z", "oo2e", "oo2ze", "roz", "roe"]:
        if s not in body:
            out.append(f"    (void){s};")
    out.append(body)
    out.append("}")
    out.append("")

    # HRR: ket first on every e component, then bra
    lines = []
    done = {}
    fidx = {f: k for k, f in enumerate(fcomps)}
    eidx = {e: k for k, e in enumerate(ecomps)}

    def ket(e, c, d):
        key = ("k", e, c, d)
        if key in done:
            return done[key]
        if sum(d) == 0:
            nm = f"acc[{eidx[e] * len(fcomps) + fidx[c]}]"
            done[key] = nm
            return nm
        i = pick_dir(d)
        a = ket(e, add(c, i), sub(d, i))
        b = ket(e, c, sub(d, i))
        nm = f"k_{name(e)}_{name(c)}_{name(d)}"
        lines.append(f"    const real {nm} = {a} + CD{XYZ[i]} * {b};")
        done[key] = nm
        return nm

    def bra(a_, b_, c, d):
        key = ("b", a_, b_, c, d)
        if key in done:
            return done[key]
        if sum(b_) == 0:
            return ket(a_, c, d)
        i = pick_dir(b_)
        x = bra(add(a_, i), sub(b_, i), c, d)
        y = bra(a_, sub(b_, i), c, d)
        nm = f"h_{name(a_)}_{name(b_)}_{name(c)}_{name(d)}"
        lines.append(f"    const real {nm} = {x} + AB{XYZ[i]} * {y};")
        done[key] = nm
        return nm

    na, nb, nc, nd = ncart(la), ncart(lb), ncart(lc), ncart(ld)
    # grouped per ket component (c,d): the ket transfer for every e, then the
    # bra transfer, each final integral handed to the callback at once -- no
    # (ab|cd) block is ever materialised
    stores = []
    for ic, c in enumerate(comps(lc)):
        for id_, d in enumerate(comps(ld)):
            for ia, a_ in enumerate(comps(la)):
                for ib, b_ in enumerate(comps(lb)):
                    nm = bra(a_, b_, c, d)
                    lines.append(f"    cb({ia}, {ib}, {ic}, {id_}, {nm});")
    out.append("template <typename real, typename Cb>")
    out.append(f"POLY_HD void hrr_{tag}(const real ABx, const real ABy, const real ABz,")
    out.append("    const real CDx, const real CDy, const real CDz, const real* __restrict__ acc, Cb& cb) {")
    body = "\n".join(lines)
    for s_ in ["ABx", "ABy", "ABz", "CDx", "CDy", "CDz"]:
        if s_ not in body:
            out.append(f"    (void){s_};")
    out.append(body)
    out.append("}")
    out.append("")
    return "\n".join(out), nacc


def classes(lmax=LMAX):
    pairs = [(a, b) for a in range(lmax + 1) for b in range(a + 1)]
    # pair-class order used everywhere: by (la, lb) lexicographic
    res = []
    for i, (la, lb) in enumerate(pairs):
        for j, (lc, ld) in enumerate(pairs):
            if j <= i:
                res.append((la, lb, lc, ld))
    return pairs, res


def main(out_path):
    pairs, cls = classes()
    parts = ["// Generated by gen_eri.py -- do not edit.",
             "#pragma once", ""]
    nacc = {}
    for c in cls:
        code, n = gen_class(*c)
        parts.append(code)
        nacc[c] = n
    parts.append(f"#define POLYERI_LMAX {LMAX}")
    parts.append(f"#define POLYERI_NPAIRCLASS {len(pairs)}")
    parts.append("")
    with open(out_path, "w") as f:
        f.write("\n".join(parts))
    return cls, nacc


def write_kernel_units(out_dir):
    """One translation unit per class (the five k_eri modes, and the segment
    kernel as getter modes 5 and 6) plus the getter, so the class kernels
    compile in parallel."""
    _, cls = classes()
    os.makedirs(out_dir, exist_ok=True)
    names = []
    for (la, lb, lc, ld) in cls:
        tag = f"{la}{lb}{lc}{ld}"
        path = os.path.join(out_dir, f"kgen_{tag}.hip")
        body = ["// Generated by gen_eri.py -- do not edit.",
                '#include "../eri_kernels.h"', "",
                f"EriKernel polyeri_kernel_{tag}(int mode) {{"]
        if ncart(la) * ncart(lb) * ncart(lc) * ncart(ld) > COOP_MIN_INTEGRALS:
            # runs only through the cooperative kernel (kcoop.hip)
            body += ["    (void)mode;", "    return nullptr;", "}", ""]
        else:
            body.append("    switch (mode) {")
            for m in range(5):
                body.append(f"    case {m}: return k_eri<{la}, {lb}, {lc}, {ld}, {m}>;")
            # getter modes 5 and 6: the segment kernel in FP64 and mixed precision
            body.append(f"    case 5: return k_eri_seg<{la}, {lb}, {lc}, {ld}, 0>;")
            body.append(f"    case 6: return k_eri_seg<{la}, {lb}, {lc}, {ld}, 3>;")
            body += ["    }", "    return nullptr;", "}", ""]
        text = "\n".join(body)
        if not os.path.exists(path) or open(path).read() != text:
            with open(path, "w") as f:
                f.write(text)
        names.append(tag)
    lines = ["// Generated by gen_eri.py -- do not edit.", '#include "../eri_kernels.h"', ""]
    for t in names:
        lines.append(f"EriKernel polyeri_kernel_{t}(int mode);")
    lines += ["", "EriKernel polyeri_class_kernel(int la, int lb, int lc, int ld, int mode) {",
              "    switch (la * 1000 + lb * 100 + lc * 10 + ld) {"]
    for t in names:
        lines.append(f"    case {int(t)}: return polyeri_kernel_{t}(mode);")
    lines += ["    }", "    return nullptr;", "}", ""]
    text = "\n".join(lines)
    path = os.path.join(out_dir, "kgen_table.hip")
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    return names


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    out = os.path.join(here, "eri_generated.h")
    text_before = open(out).read() if os.path.exists(out) else None
    tmp = out + ".new"
    cls, nacc = main(tmp)
    if text_before == open(tmp).read():
        os.remove(tmp)
    else:
        os.replace(tmp, out)
    names = write_kernel_units(os.path.join(here, "kgen"))
    print(f"{len(cls)} classes -> {out}, {len(names)} kernel units")


# ---------------------------------------------------------------------------
# Level-scheduled programs for the cooperative (one wavefront per quartet)
# kernel: node values live in LDS; every op of one level is independent.
# ---------------------------------------------------------------------------
def class_program(la, lb, lc, ld):
    """Returns dict with op lists.  Slot 0 holds 0.0; slots 1..nacc the
    contracted accumulators; the VRR region (F_0..F_L then VRR nodes) and,
    after the primitive loop, the HRR region both start at nacc+1."""
    L = la + lb + lc + ld
    ec = [c for l in range(la, la + lb + 1) for c in comps(l)]
    fc = [c for l in range(lc, lc + ld + 1) for c in comps(l)]
    nacc = len(ec) * len(fc)
    base = nacc + 1
    slot, level, ops = {}, {}, []   # VRR
    for m in range(L + 1):
        slot[((0, 0, 0), (0, 0, 0), m)] = base + m
        level[((0, 0, 0), (0, 0, 0), m)] = 0
    nxt = [base + L + 1]

    def node(e, f, m):
        key = (e, f, m)
        if key in slot:
            return key
        if sum(f) == 0:
            i = pick_dir(e)
            e1 = sub(e, i)
            s1, s2 = node(e1, f, m), node(e1, f, m + 1)
            s3 = s4 = None
            n3 = 0
            if e[i] >= 2:
                s3, s4 = node(sub(e, i, 2), f, m), node(sub(e, i, 2), f, m + 1)
                n3 = e[i] - 1
            srcs, typ, n5, s5 = [s1, s2, s3, s4], 0, 0, None
        else:
            i = pick_dir(f)
            f1 = sub(f, i)
            s1, s2 = node(e, f1, m), node(e, f1, m + 1)
            s3 = s4 = s5 = None
            n3 = n5 = 0
            if f[i] >= 2:
                s3, s4 = node(e, sub(f, i, 2), m), node(e, sub(f, i, 2), m + 1)
                n3 = f[i] - 1
            if e[i] >= 1:
                s5 = node(sub(e, i), f1, m + 1)
                n5 = e[i]
            srcs, typ = [s1, s2, s3, s4, s5], 1
        lv = 1 + max(level[s] for s in srcs if s is not None)
        slot[key] = nxt[0]
        nxt[0] += 1
        level[key] = lv
        get = lambda k: slot[k] if k is not None else 0
        ops.append((lv, slot[key], get(s1), get(s2), get(s3), get(s4), get(s5 if typ else None), typ, i, n3, n5))
        return key

    accmap = []
    for ie, e in enumerate(ec):
        for jf, f in enumerate(fc):
            k = node(e, f, 0)
            accmap.append((1 + ie * len(fc) + jf, slot[k]))
    nvrr_slots = nxt[0]
    # HRR
    hslot, hlevel, hops = {}, {}, []
    hn = [base]
    fidx = {f: k for k, f in enumerate(fc)}
    eidx = {e: k for k, e in enumerate(ec)}

    def ket(e, c, d):
        key = ("k", e, c, d)
        if key in hslot:
            return key
        if sum(d) == 0:
            hslot[key] = 1 + eidx[e] * len(fc) + fidx[c]
            hlevel[key] = 0
            return key
        i = pick_dir(d)
        a = ket(e, add(c, i), sub(d, i))
        b = ket(e, c, sub(d, i))
        hslot[key] = hn[0]; hn[0] += 1
        hlevel[key] = 1 + max(hlevel[a], hlevel[b])
        hops.append((hlevel[key], hslot[key], hslot[a], hslot[b], 1, i))
        return key

    def bra(a_, b_, c, d):
        key = ("b", a_, b_, c, d)
        if key in hslot:
            return key
        if sum(b_) == 0:
            return ket(a_, c, d)
        i = pick_dir(b_)
        x = bra(add(a_, i), sub(b_, i), c, d)
        y = bra(a_, sub(b_, i), c, d)
        hslot[key] = hn[0]; hn[0] += 1
        hlevel[key] = 1 + max(hlevel[x], hlevel[y])
        hops.append((hlevel[key], hslot[key], hslot[x], hslot[y], 0, i))
        return key

    finals = []
    for ia, a_ in enumerate(comps(la)):
        for ib, b_ in enumerate(comps(lb)):
            for ic, c in enumerate(comps(lc)):
                for id_, d in enumerate(comps(ld)):
                    k = bra(a_, b_, c, d)
                    finals.append((hslot[k], ia, ib, ic, id_))
    nslot = max(nvrr_slots, hn[0])
    return dict(L=L, nacc=nacc, base=base, ops=sorted(ops), hops=sorted(hops), accmap=accmap,
                finals=finals, nslot=nslot)


def write_programs(path):
    _, cls = classes()
    vops, hops, aux, descs = [], [], [], []
    for c in cls:
        p = class_program(*c)
        d = {}
        # VRR ops grouped by level
        # VRR op words carry LDS byte offsets of 8-byte slots (slot 0 = the
        # zero slot for absent terms); w: typ | i << 1 | n3 << 3 | n5 << 6,
        # bit 9 set when the op's node is a contracted target, whose
        # accumulator's byte offset is then in w >> 16
        dsts = {op[1] for op in p["ops"]}
        acc_of = {s: a for a, s in p["accmap"] if s in dsts}
        assert 8 * p["nslot"] < 65536
        lv_off = []
        cur = None
        for op in p["ops"]:
            if op[0] != cur:
                cur = op[0]
                lv_off.append(len(vops))
            _, dst, s1, s2, s3, s4, s5, typ, i, n3, n5 = op
            w = typ | (i << 1) | (n3 << 3) | (n5 << 6)
            if dst in acc_of:
                w |= (1 << 9) | ((8 * acc_of[dst]) << 16)
            vops.append((8 * dst | (8 * s1 << 16), 8 * s2 | (8 * s3 << 16), 8 * s4 | (8 * s5 << 16), w))
        lv_off.append(len(vops))
        h_off = []
        cur = None
        for op in p["hops"]:
            if op[0] != cur:
                cur = op[0]
                h_off.append(len(hops))
            _, dst, s1, s2, typ, i = op
            hops.append((dst | (s1 << 16), s2, 0, typ | (i << 1)))
        h_off.append(len(hops))
        # the accumulation map keeps only the targets no VRR op produces
        # ((00|00)'s F_0, a Boys value): the others accumulate in their op
        acc_off = len(aux)
        accrest = [(a, s) for a, s in p["accmap"] if s not in dsts]
        for a, s in accrest:
            aux.append(a | (s << 16))
        fin_off = len(aux)
        for s, a, b, cc, dd in p["finals"]:
            aux.append(s | (a << 16) | (b << 20) | (cc << 24) | (dd << 28))
        descs.append((c, lv_off, h_off, acc_off, len(accrest), fin_off, len(p["finals"]), p["nslot"], p["L"], p["nacc"], p["base"]))
    maxlv = max(max(len(d[1]), len(d[2])) for d in descs)
    out = ["// Generated by gen_eri.py -- do not edit.", "#pragma once", "#include <stdint.h>", "",
           "// The tables have external linkage (no const at namespace scope): the HIP",
           "// runtime resolves every registered device variable through the code",
           "// object's global symbols.",
           f"#define PROG_MAXLV {maxlv}",
           "struct ProgDesc {",
           "    int vrr_lv[PROG_MAXLV];   // level starts into g_prog_vrr, last = end",
           "    int nvrr_lv;",
           "    int hrr_lv[PROG_MAXLV];",
           "    int nhrr_lv;",
           "    int acc_off, nacc_map, fin_off, nfin, nslot, L, nacc, base;",
           "};", ""]

    def arr(name, rows):
        s = [f"__device__ uint4 {name}[] = {{"]
        for r in rows:
            s.append("    {%du, %du, %du, %du}," % r)
        s.append("};")
        return s
    out += arr("g_prog_vrr", vops) + [""] + arr("g_prog_hrr", hops) + [""]
    out.append("__device__ uint32_t g_prog_aux[] = {")
    for k in range(0, len(aux), 12):
        out.append("    " + ", ".join(f"{x}u" for x in aux[k:k + 12]) + ",")
    out.append("};")
    out.append("")
    out.append(f"__device__ ProgDesc g_prog_desc[{len(descs)}] = {{")
    for (c, lv, hv, ao, na, fo, nf, ns, L, nacc, base) in descs:
        lvs = ", ".join(str(x) for x in lv + [0] * (maxlv - len(lv)))
        hvs = ", ".join(str(x) for x in hv + [0] * (maxlv - len(hv)))
        out.append(f"    // ({c[0]}{c[1]}|{c[2]}{c[3]})")
        out.append(f"    {{{{{lvs}}}, {len(lv) - 1}, {{{hvs}}}, {len(hv) - 1}, {ao}, {na}, {fo}, {nf}, {ns}, {L}, {nacc}, {base}}},")
    out.append("};")
    out.append("")
    out.append("// LDS slots per class for the host (no device-symbol lookup needed)")
    out.append(f"static const int g_prog_nslot[{len(descs)}] = {{" + ", ".join(str(d[7]) for d in descs) + "};")
    out.append("")
    out.append("// class index of (la lb|lc ld) in g_prog_desc")
    out.append("static inline int prog_index(int la, int lb, int lc, int ld) {")
    out.append("    switch (la * 1000 + lb * 100 + lc * 10 + ld) {")
    for k, (c, *_r) in enumerate(descs):
        out.append(f"    case {c[0] * 1000 + c[1] * 100 + c[2] * 10 + c[3]}: return {k};")
    out.append("    }")
    out.append("    return -1;")
    out.append("}")
    text = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != text:
        with open(path, "w") as f:
            f.write(text)
    return max(d[7] for d in descs)


if __name__ == "__main__" and os.environ.get("GEN_ERI_PROGRAMS", "1") == "1":
    here = os.path.dirname(os.path.abspath(__file__))
    mx = write_programs(os.path.join(here, "eri_programs.h"))
    print(f"programs -> eri_programs.h (max {mx} LDS slots)")
