"""Generator of the gradient kernel's per-class tables (grad_tables.inc).

The nuclear-gradient kernel (qk_grad.hip) is one table-driven wavefront-
cooperative kernel for every quartet class.  For class (la lb|lc ld) it needs:

* a VRR node program (the Obara-Saika DAG, levelled; the same node encoding as
  the Fock team kernel) whose targets are the [e0|f0] of four contracted sets,
  the VRR raised by one on each side (shell.f90:1906-1912, 1968):
      S0: |e| in [la-1, la+lb],   |f| in [lc-1, lc+ld]   weight 1
      SA: |e| in [la+1, la+lb+1], |f| in [lc, lc+ld]     weight 2 alpha
      SB: |e| in [la, la+lb+1],   |f| in [lc, lc+ld]     weight 2 beta
      SC: |e| in [la, la+lb],     |f| in [lc+1, lc+ld+1] weight 2 gamma
  (alpha, beta, gamma: exponents of the primitives on A, B, C; classopt,
  shell.f90:2056-2114);
* per derivative term ("stage") a bra HRR table, a ket HRR table and a
  contraction table.  d/dA_x (ab|cd) = 2 alpha (a+1x b|cd) - a_x (a-1x b|cd)
  and likewise for B and C (hrrwholeopt, subs/hrr.f90:862-1014); the stages
  are A+ (from SA), A- (S0), B+ (SB), B- (S0), C+ (SC), C- (S0).

Everything is emitted as one int32 blob per class with a header of offsets
(layout in qk_internal.h, GradHdr) that the host uploads once.
"""
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from gen_eri import QUICK_ORDER, carts, erange, first_dir, sub, tag, quartet_classes, binom_terms  # noqa: E402

MAX_TERMS = 8       # HRR terms per output row: prod (b_i + 1) <= 8 for |b| <= 3
N_STAGE_FIELDS = 8  # dir, src, nR1, nCol, nR2, off_bra, off_ket, off_con
HDR = 24            # fixed header fields before the stage records


def order(L):
    """Cartesian order of shell L: QUICK's function order for s/p/d, carts(3) for f."""
    return QUICK_ORDER[L] if L in QUICK_ORDER else carts(L)


def grad_program(sets):
    """Levelled VRR node program for the union of the sets' targets."""
    targets = []
    for E, F in sets:
        targets += [(e, f, 0) for e in E for f in F]
    deps_of, order_, seen = {}, [], set()

    def deps(node):
        e, f, m = node
        d = {}
        if sum(f) > 0:
            i = first_dir(f)
            f1 = sub(f, i)
            d["a"], d["b"] = (e, f1, m), (e, f1, m + 1)
            if f1[i] > 0:
                f2 = sub(f1, i)
                d["c"], d["d"] = (e, f2, m), (e, f2, m + 1)
                d["n1"] = f1[i]
            if e[i] > 0:
                d["x"] = (sub(e, i), f1, m + 1)
                d["ne"] = e[i]
            d["kind"], d["dir"] = 2, i
        elif sum(e) > 0:
            i = first_dir(e)
            e1 = sub(e, i)
            d["a"], d["b"] = (e1, f, m), (e1, f, m + 1)
            if e1[i] > 0:
                e2 = sub(e1, i)
                d["c"], d["d"] = (e2, f, m), (e2, f, m + 1)
                d["n1"] = e1[i]
            d["kind"], d["dir"] = 1, i
        else:
            d["kind"] = 0
        return d

    def visit(node):
        if node in seen:
            return
        d = deps(node)
        deps_of[node] = d
        for k in ("a", "b", "c", "d", "x"):
            if k in d:
                visit(d[k])
        seen.add(node)
        order_.append(node)

    for t in targets:
        visit(t)
    level = {}
    for node in order_:
        d = deps_of[node]
        lv = 0
        for k in ("a", "b", "c", "d", "x"):
            if k in d:
                lv = max(lv, level[d[k]] + 1)
        level[node] = lv
    nlev = max(level.values()) + 1
    by_level = [[n for n in order_ if level[n] == L] for L in range(nlev)]
    slot, s = {}, 1
    for L in range(nlev):
        for n in by_level[L]:
            slot[n] = s
            s += 1
    nodes, lev_start = [], [0]
    for L in range(nlev):
        for n in by_level[L]:
            d = deps_of[n]
            g = lambda k: slot[d[k]] if k in d else 0
            if d["kind"] == 0:
                flags = n[2] << 8
            else:
                flags = d["kind"] | (d["dir"] << 2) | (d.get("n1", 0) << 4) | (d.get("ne", 0) << 8)
            nodes.append((slot[n] | (g("a") << 16), g("b") | (g("c") << 16), g("d") | (g("x") << 16), flags))
        lev_start.append(len(nodes))
    tslots = [[slot[(e, f, 0)] for e in E for f in F] for E, F in sets]
    return nodes, lev_start, tslots, s


def hrr_rows(l1, l2, src_list):
    """(a b| for a in order(l1), b in order(l2) from rows of src_list: MAX_TERMS
    terms (row, powcode, coef), padded with coef 0."""
    idx = {e: k for k, e in enumerate(src_list)}
    rows = []
    for a in order(l1):
        for b in order(l2):
            t = [(idx[e], px | (py << 2) | (pz << 4), c) for e, px, py, pz, c in binom_terms(a, b)]
            assert len(t) <= MAX_TERMS
            t += [(0, 0, 0)] * (MAX_TERMS - len(t))
            rows.append(t)
    return rows


def pack_term(row, pc, coef):
    """one HRR term in 32 bits: row (bits 0-11), (A-B)/(C-D) power code (12-17),
    binomial coefficient (18-21); coefficient 0 marks padding."""
    assert 0 <= row < 4096 and 0 <= pc < 64 and 0 <= coef < 16
    return row | (pc << 12) | (coef << 18)


def pack_con(ix, fac):
    """contraction entry: output index (bits 0-15, 0xffff = none), factor (signed, bits 16-31)."""
    if ix < 0:
        return 0xffff
    assert ix < 0xffff and -32768 <= fac < 32768
    v = ix | ((fac & 0xffff) << 16)
    return v - (1 << 32) if v >= (1 << 31) else v     # as a signed int32


def class_blob(cls):
    la, lb, lc, ld = cls
    r = lambda lo, hi: erange(max(lo, 0), hi)
    E0, F0 = r(la - 1, la + lb), r(lc - 1, lc + ld)
    EA, FA = r(la + 1, la + lb + 1), r(lc, lc + ld)
    EB, FB = r(la, la + lb + 1), r(lc, lc + ld)
    EC, FC = r(la, la + lb), r(lc + 1, lc + ld + 1)
    sets = [(E0, F0), (EA, FA), (EB, FB), (EC, FC)]
    nodes, lev, tslots, nslots = grad_program(sets)
    NA, NB, NC, ND = (len(order(l)) for l in cls)
    nout = NA * NB * NC * ND
    # stages: (dir, src set, bra (l1,l2), ket (l3,l4), shifted shell index, sign)
    stages = [(0, 1, (la + 1, lb), (lc, ld), 0, +1)]
    if la >= 1:
        stages.append((0, 0, (la - 1, lb), (lc, ld), 0, -1))
    stages.append((1, 2, (la, lb + 1), (lc, ld), 1, +1))
    if lb >= 1:
        stages.append((1, 0, (la, lb - 1), (lc, ld), 1, -1))
    stages.append((2, 3, (la, lb), (lc + 1, ld), 2, +1))
    if lc >= 1:
        stages.append((2, 0, (la, lb), (lc - 1, ld), 2, -1))

    blob = [0] * (HDR + N_STAGE_FIELDS * 6)
    body = []

    def put(vals):
        # every table starts on a 16-byte boundary (read as int4 / uint4)
        while (len(blob) + len(body)) % 4:
            body.append(0)
        off = len(blob) + len(body)
        body.extend(vals)
        return off

    off_nodes = put([v for nd in nodes for v in nd])
    off_lev = put(lev)
    off_tgt = put([s for ts in tslots for s in ts])
    stage_recs = []
    base_l = [la, lb, lc, ld]
    for dirn, src, (l1, l2), (l3, l4), which, sign in stages:
        Es, Fs = sets[src]
        bra = hrr_rows(l1, l2, Es)
        ket = hrr_rows(l3, l4, Fs)
        nR1, nR2, ncol = len(bra), len(ket), len(Fs)
        off_bra = put([pack_term(*term) for row in bra for term in row])
        off_ket = put([pack_term(*term) for row in ket for term in row])
        # contraction: for each function quartet of the class and axis, the
        # element of this stage's output and its factor (+1 or -n_x)
        shp = list(base_l)
        shp[which] += sign
        o1 = {a: k for k, a in enumerate(order(shp[0]))}
        o2 = {b: k for k, b in enumerate(order(shp[1]))}
        o3 = {c: k for k, c in enumerate(order(shp[2]))}
        o4 = {d: k for k, d in enumerate(order(shp[3]))}
        n2, n3, n4 = len(o2), len(o3), len(o4)
        con = []
        for a in order(la):
            for b in order(lb):
                for c in order(lc):
                    for d in order(ld):
                        for x in range(3):
                            v = [a, b, c, d]
                            nx = v[which][x]
                            if sign < 0 and nx == 0:
                                con.append(pack_con(-1, 0))
                                continue
                            w = list(v[which])
                            w[x] += sign
                            v[which] = tuple(w)
                            i1, i2, i3, i4 = o1[v[0]], o2[v[1]], o3[v[2]], o4[v[3]]
                            con.append(pack_con((i1 * n2 + i2) * (n3 * n4) + i3 * n4 + i4, 1 if sign > 0 else -nx))
                        con.append(0xffff)   # pad each quartet's record to four entries
        off_con = put(con)
        assert nR1 == len(o1) * n2 and nR2 == n3 * n4
        # the most terms any row of the stage's bra / ket table has (the rest of
        # a row's MAX_TERMS slots are padding), stored above the direction
        nb = max(sum(1 for t in row if t[2] != 0) for row in bra)
        nk = max(sum(1 for t in row if t[2] != 0) for row in ket)
        stage_recs.append([dirn | (nb << 8) | (nk << 16), src, nR1, ncol, nR2, off_bra, off_ket, off_con])
    n_sets = [len(E) * len(F) for E, F in sets]
    max_stage = max(rec[2] * rec[3] + rec[2] * rec[4] for rec in stage_recs)
    hdr = [len(nodes), len(lev) - 1, nslots, la + lb + lc + ld,
           n_sets[0], n_sets[1], n_sets[2], n_sets[3],
           len(F0), len(FA), len(FB), len(FC), len(stage_recs), nout, NA, NB, NC, ND,
           off_nodes, off_lev, off_tgt, max_stage, 0, 0]
    assert len(hdr) == HDR
    blob[:HDR] = hdr
    for k, rec in enumerate(stage_recs):
        blob[HDR + N_STAGE_FIELDS * k: HDR + N_STAGE_FIELDS * (k + 1)] = rec
    return blob + body, dict(nodes=len(nodes), levels=len(lev) - 1, slots=nslots, sets=n_sets,
                             max_stage=max_stage, nout=nout)


def write(path):
    out = ["// Generated by quick_amd/csrc/gen_grad.py -- do not edit.",
           "// Per-class tables of the nuclear-gradient kernel (layout: GradHdr in qk_internal.h).",
           "#pragma once", "namespace qk {"]
    names = []
    for cls in quartet_classes():
        t = tag(cls)
        blob, info = class_blob(cls)
        out.append(f"static const int kGradBlob_{t}[{len(blob)}] = {{")
        for i in range(0, len(blob), 24):
            out.append("    " + ", ".join(str(v) for v in blob[i:i + 24]) + ",")
        out.append("};")
        names.append((t, len(blob)))
    out.append("static const int* const kGradBlobs[21] = {" + ", ".join(f"kGradBlob_{t}" for t, _ in names) + "};")
    out.append("static const int kGradBlobLen[21] = {" + ", ".join(str(n) for _, n in names) + "};")
    out.append("}  // namespace qk")
    new = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != new:
        open(path, "w").write(new)


# ---------------------------------------------------------------------------
# Straight-line gradient code for the light classes (one quartet per lane):
#   vrg_<tag>: the raised VRR, accumulating the four contracted sets
#              (acc0, accA, accB, accC with weights 1, wA, wB, wC);
#   gcon_<tag>: the six derivative terms' bra/ket HRR and their contraction
#              with the density factors K[o] into g[9] (A, B, C x xyz).

LIGHT_CLASSES = [(0, 0, 0, 0), (1, 0, 0, 0), (1, 0, 1, 0), (1, 1, 0, 0), (2, 0, 0, 0), (1, 1, 1, 0), (2, 0, 1, 0),
                 (2, 1, 0, 0), (1, 1, 1, 1), (2, 0, 1, 1), (2, 0, 2, 0), (2, 1, 1, 0), (2, 2, 0, 0)]


def grad_sets(cls):
    la, lb, lc, ld = cls
    r = lambda lo, hi: erange(max(lo, 0), hi)
    return [(r(la - 1, la + lb), r(lc - 1, lc + ld)), (r(la + 1, la + lb + 1), r(lc, lc + ld)),
            (r(la, la + lb + 1), r(lc, lc + ld)), (r(la, la + lb), r(lc + 1, lc + ld + 1))]


def gen_vrr_grad(cls, mask=15):
    """vrg_<tag>_<mask>: the raised VRR accumulating the sets in mask (bit 0: S0,
    1: SA, 2: SB, 3: SC)."""
    from gen_eri import name_of
    sets = grad_sets(cls)
    targets = []
    for s_, (E, F) in enumerate(sets):
        if mask >> s_ & 1:
            targets += [(e, f, 0) for e in E for f in F]
    order_, seen = [], set()

    def visit(node):
        if node in seen:
            return
        e, f, m = node
        deps = []
        if sum(f) > 0:
            i = first_dir(f)
            f1 = sub(f, i)
            deps += [(e, f1, m), (e, f1, m + 1)]
            if f1[i] > 0:
                f2 = sub(f1, i)
                deps += [(e, f2, m), (e, f2, m + 1)]
            if e[i] > 0:
                deps.append((sub(e, i), f1, m + 1))
        elif sum(e) > 0:
            i = first_dir(e)
            e1 = sub(e, i)
            deps += [(e1, f, m), (e1, f, m + 1)]
            if e1[i] > 0:
                e2 = sub(e1, i)
                deps += [(e2, f, m), (e2, f, m + 1)]
        for d in deps:
            visit(d)
        seen.add(node)
        order_.append(node)

    for t in targets:
        visit(t)
    X = "xyz"
    # where each target goes: node -> [(set, index)]
    where = {}
    for s_, (E, F) in enumerate(sets):
        if not mask >> s_ & 1:
            continue
        k = 0
        for e in E:
            for f in F:
                where.setdefault((e, f, 0), []).append((s_, k))
                k += 1
    lines = []
    for node in order_:
        e, f, m = node
        nm = name_of(e, f, m)
        if sum(e) == 0 and sum(f) == 0:
            lines.append(f"    const R {nm} = F[{m}];")
        else:
            if sum(f) > 0:
                i = first_dir(f)
                f1 = sub(f, i)
                expr = f"QC{X[i]} * {name_of(e, f1, m)} + WQ{X[i]} * {name_of(e, f1, m + 1)}"
                if f1[i] > 0:
                    f2 = sub(f1, i)
                    expr += f" + R({f1[i]}) * oo2e * ({name_of(e, f2, m)} - roe * {name_of(e, f2, m + 1)})"
                if e[i] > 0:
                    expr += f" + R({e[i]}) * oo2ze * {name_of(sub(e, i), f1, m + 1)}"
            else:
                i = first_dir(e)
                e1 = sub(e, i)
                expr = f"PA{X[i]} * {name_of(e1, f, m)} + WP{X[i]} * {name_of(e1, f, m + 1)}"
                if e1[i] > 0:
                    e2 = sub(e1, i)
                    expr += f" + R({e1[i]}) * oo2z * ({name_of(e2, f, m)} - roz * {name_of(e2, f, m + 1)})"
            lines.append(f"    const R {nm} = {expr};")
        # accumulate a target as soon as it exists, so it need not stay live
        for s_, k in where.get(node, []):
            w = ["", "wA * ", "wB * ", "wC * "][s_]
            lines.append(f"    acc{'0ABC'[s_]}[{k}] += {w}{nm};")
    acc = []
    sig = (f"template <typename R>\n__device__ __forceinline__ void vrg_{tag(cls)}_{mask}(const R* __restrict__ F, "
           "R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, "
           "R oo2z, R oo2e, R oo2ze, R roz, R roe, R wA, R wB, R wC, "
           "R* __restrict__ acc0, R* __restrict__ accA, R* __restrict__ accB, R* __restrict__ accC)")
    return f"{sig}\n{{\n" + "\n".join(lines + acc) + "\n}\n", [len(E) * len(F) for E, F in sets]


def gen_grad_contract(cls, mask=15):
    """gcon_<tag>_<mask>: the derivative terms whose source set is in mask."""
    la, lb, lc, ld = cls
    sets = grad_sets(cls)
    X = "xyz"
    lines = []
    counter = [0]

    def fresh(p):
        counter[0] += 1
        return f"{p}{counter[0]}"

    def stage_values(src, l1, l2, l3, l4):
        """names of (a b|c d) for a in order(l1), b in order(l2), c in order(l3), d in order(l4)
        built from set src by bra then ket HRR."""
        E, F = sets[src]
        eidx = {e: k for k, e in enumerate(E)}
        fidx = {f: k for k, f in enumerate(F)}
        nF = len(F)
        arr = f"acc{'0ABC'[src]}"
        bra = {}

        def bra_expr(a, b, fk):
            key = (a, b, fk)
            if key in bra:
                return bra[key]
            if sum(b) == 0:
                bra[key] = f"{arr}[{eidx[a] * nF + fk}]"
                return bra[key]
            i = first_dir(b)
            b1 = sub(b, i)
            a1 = list(a)
            a1[i] += 1
            x1 = bra_expr(tuple(a1), b1, fk)
            x2 = bra_expr(a, b1, fk)
            nm = fresh("hb")
            lines.append(f"    const R {nm} = {x1} + AB{X[i]} * {x2};")
            bra[key] = nm
            return nm

        ket = {}

        def ket_expr(a, b, c, d):
            key = (a, b, c, d)
            if key in ket:
                return ket[key]
            if sum(d) == 0:
                ket[key] = bra_expr(a, b, fidx[c])
                return ket[key]
            i = first_dir(d)
            d1 = sub(d, i)
            c1 = list(c)
            c1[i] += 1
            x1 = ket_expr(a, b, tuple(c1), d1)
            x2 = ket_expr(a, b, c, d1)
            nm = fresh("hk")
            lines.append(f"    const R {nm} = {x1} + CD{X[i]} * {x2};")
            ket[key] = nm
            return nm

        return ket_expr

    stages = [(0, 1, +1), (0, 0, -1), (1, 2, +1), (1, 0, -1), (2, 3, +1), (2, 0, -1)]
    base = [la, lb, lc, ld]
    for which, src, sign in stages:
        if sign < 0 and base[which] == 0:
            continue
        if not mask >> src & 1:
            continue
        shp = list(base)
        shp[which] += sign
        val = stage_values(src, *shp)
        sums = {x: [] for x in range(3)}
        o = 0
        for a in order(la):
            for b in order(lb):
                for c in order(lc):
                    for d in order(ld):
                        v = [a, b, c, d]
                        for x in range(3):
                            nx = v[which][x]
                            if sign < 0 and nx == 0:
                                continue
                            w = list(v[which])
                            w[x] += sign
                            vv = list(v)
                            vv[which] = tuple(w)
                            term = val(*vv)
                            fac = "" if sign > 0 else (f"R({nx}) * " if nx > 1 else "")
                            op = "+=" if sign > 0 else "-="
                            # summed as soon as the value exists, so it need not stay live
                            lines.append(f"    g[{3 * which + x}] {op} {fac}K[{o}] * {term};")
                        o += 1
    sig = (f"template <typename R>\n__device__ __forceinline__ void gcon_{tag(cls)}_{mask}(const R* __restrict__ acc0, "
           "const R* __restrict__ accA, const R* __restrict__ accB, const R* __restrict__ accC, "
           "R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, const R* __restrict__ K, R* __restrict__ g)")
    return f"{sig}\n{{\n" + "\n".join(lines) + "\n}\n"


PASS_BUDGET = 104   # accumulators + density factors (FP64 values) per lane and pass


def grad_passes(cls):
    """Partition of the four sets into passes over the primitive quartets (the
    VRR is recomputed per pass): the fewest passes whose accumulators plus the
    NOUT density factors stay within PASS_BUDGET; four single-set passes at most."""
    sizes = [len(E) * len(F) for E, F in grad_sets(cls)]
    nout = 1
    for l in cls:
        nout *= (l + 1) * (l + 2) // 2
    for parts in ([[0, 1, 2, 3]], [[0, 1], [2, 3]], [[1], [2], [3], [0]]):
        if max(sum(sizes[i] for i in p) for p in parts) + nout <= PASS_BUDGET:
            return parts
    return [[1], [2], [3], [0]]


def write_light(path):
    out = ["// Generated by quick_amd/csrc/gen_grad.py -- do not edit.",
           "// Straight-line gradient code of the light classes: vrg_<class>_<mask> (raised VRR",
           "// into the contracted sets in mask), gcon_<class>_<mask> (derivative HRR and density",
           "// contraction of the terms those sets feed), and the passes a kernel makes.", "#pragma once", ""]
    for cls in LIGHT_CLASSES:
        t = tag(cls)
        passes = grad_passes(cls)
        masks = [sum(1 << i for i in p) for p in passes]
        sizes = [len(E) * len(F) for E, F in grad_sets(cls)]
        for m in masks:
            v, _ = gen_vrr_grad(cls, m)
            out += [v, gen_grad_contract(cls, m)]
        width = max(sum(sizes[i] for i in p) for p in passes)
        out.append(f"struct GradSets_{t} {{")
        out.append(f"    static constexpr int n0 = {sizes[0]}, nA = {sizes[1]}, nB = {sizes[2]}, nC = {sizes[3]};")
        out.append(f"    static constexpr int kPasses = {len(masks)}, kWidth = {width};")
        out.append(f"    static constexpr int kMask[4] = {{{', '.join(str(m) for m in masks + [0] * (4 - len(masks)))}}};")
        out.append("};")
        # dispatch on the mask at compile time
        out.append("template <int M, typename R>")
        out.append(f"__device__ __forceinline__ void vrg_{t}(const R* F, R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, "
                   "R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, R oo2z, R oo2e, R oo2ze, R roz, R roe, R wA, R wB, "
                   "R wC, R* a0, R* aA, R* aB, R* aC) {")
        for k, m in enumerate(masks):
            kw = "if" if k == 0 else "else if"
            out.append(f"    {kw} constexpr (M == {m}) vrg_{t}_{m}<R>(F, PAx, PAy, PAz, WPx, WPy, WPz, QCx, QCy, "
                       "QCz, WQx, WQy, WQz, oo2z, oo2e, oo2ze, roz, roe, wA, wB, wC, a0, aA, aB, aC);")
        out.append("}")
        out.append("template <int M, typename R>")
        out.append(f"__device__ __forceinline__ void gcon_{t}(const R* a0, const R* aA, const R* aB, const R* aC, "
                   "R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, const R* K, R* g) {")
        for k, m in enumerate(masks):
            kw = "if" if k == 0 else "else if"
            out.append(f"    {kw} constexpr (M == {m}) gcon_{t}_{m}<R>(a0, aA, aB, aC, ABx, ABy, ABz, CDx, CDy, CDz, K, g);")
        out.append("}")
    new = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != new:
        open(path, "w").write(new)


def write_light_units(outdir):
    """One translation unit per light gradient class."""
    os.makedirs(outdir, exist_ok=True)
    for cls in LIGHT_CLASSES:
        t = tag(cls)
        body = ["// Generated by quick_amd/csrc/gen_grad.py -- do not edit.", '#include "../qk_grad_light.h"',
                "namespace qk {",
                f"void launch_grad_light_{t}(const GradArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s) {{",
                f"    static const long long cap = scratch_grid_cap(kernel_scratch(reinterpret_cast<const void*>(&grad_light_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, double>)), 4);",
                f"    hipLaunchKernelGGL((grad_light_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, double>), dim3((int)std::min<long long>(grid, cap)), dim3(256), lds, s, a, w);",
                "}",
                f"void launch_grad_light_f32_{t}(const GradArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s) {{",
                f"    static const long long cap = scratch_grid_cap(kernel_scratch(reinterpret_cast<const void*>(&grad_light_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, float>)), 4);",
                f"    hipLaunchKernelGGL((grad_light_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, float>), dim3((int)std::min<long long>(grid, cap)), dim3(256), lds, s, a, w);",
                "}", "}  // namespace qk"]
        fn = os.path.join(outdir, f"grad_{t}.hip")
        new = "\n".join(body) + "\n"
        if not os.path.exists(fn) or open(fn).read() != new:
            open(fn, "w").write(new)


if __name__ == "__main__":
    if len(sys.argv) > 1:
        write(sys.argv[1])
        write_light(os.path.join(os.path.dirname(sys.argv[1]), "grad_light.inc"))
        write_light_units(os.path.join(os.path.dirname(sys.argv[1]), "gen"))
    else:
        for cls in quartet_classes():
            _, info = class_blob(cls)
            print(tag(cls), info)
