"""Generates the per-class ERI device code (VRR + HRR) for the HIP kernels.

For every quartet class (la lb|lc ld) with la >= lb, lc >= ld and bra pair class
>= ket pair class (s,p,d: 21 classes) this emits straight-line code for

  * ``vrr_<tag>``: the Obara-Saika / Head-Gordon-Pople vertical recurrence
    [e0|f0]^(m) from F_m(T) (already multiplied by the primitive prefactor), for
    |e| in [la, la+lb] and |f| in [lc, lc+ld], accumulated into ``acc`` -- only
    the intermediates the targets need are emitted, in dependency order;
  * ``hrr_<tag>``: the horizontal recurrences (a,b+1_i) = (a+1_i,b) + AB_i (a,b)
    on the bra and (c,d+1_i) = (c+1_i,d) + CD_i (c,d) on the ket, producing
    (ab|cd) in QUICK's component order (p: x,y,z; d: xx,xy,yy,xz,yz,zz).

The recurrence variables are the ones QUICK's `vertical` takes
(subs/hrr.f90:1018; gpu_get2e_subs.h:679-682): P-A, W-P, Q-C, W-Q, 1/2zeta,
1/2eta, 1/2(zeta+eta), eta/(zeta+eta), zeta/(zeta+eta).  The direction chosen
at every step (first non-zero Cartesian index) is the one the CPU oracle uses,
so both evaluate the same arithmetic expressions.

usage: python gen_eri.py <out.inc>
"""
from __future__ import annotations

import sys
from functools import lru_cache

LMAX = 8


def carts(L):
    """Cartesian exponent triples of total L in the oracle's index order."""
    out = []
    for i in range(L, -1, -1):
        for j in range(L - i, -1, -1):
            out.append((i, j, L - i - j))
    return out


# QUICK's component order within a shell (basis.f90:415-574)
QUICK_ORDER = {0: [(0, 0, 0)], 1: [(1, 0, 0), (0, 1, 0), (0, 0, 1)],
               2: [(2, 0, 0), (1, 1, 0), (0, 2, 0), (1, 0, 1), (0, 1, 1), (0, 0, 2)]}

PAIR_CLASSES = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]


def quartet_classes():
    out = []
    for bi, bra in enumerate(PAIR_CLASSES):
        for ki, ket in enumerate(PAIR_CLASSES):
            if ki <= bi:
                out.append(bra + ket)
    return out


def tag(c):
    return "".join("spdf"[x] for x in c)


def erange(lo, hi):
    """Cartesians with lo <= |e| <= hi, in order of increasing |e|."""
    out = []
    for L in range(lo, hi + 1):
        out += carts(L)
    return out


def first_dir(v):
    return 0 if v[0] > 0 else (1 if v[1] > 0 else 2)


def sub(v, i, n=1):
    w = list(v)
    w[i] -= n
    return tuple(w)


def name_of(e, f, m):
    return "v_%d%d%d_%d%d%d_%d" % (e + f + (m,))


def gen_vrr(cls):
    la, lb, lc, ld = cls
    E, F = la + lb, lc + ld
    targets = [(e, f, 0) for e in erange(la, E) for f in erange(lc, F)]
    order, seen = [], set()

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
        order.append(node)

    for t in targets:
        visit(t)

    X = "xyz"
    lines = []
    nflop = 0
    for node in order:
        e, f, m = node
        nm = name_of(e, f, m)
        if sum(e) == 0 and sum(f) == 0:
            lines.append(f"    const R {nm} = F[{m}];")
            continue
        if sum(f) > 0:
            i = first_dir(f)
            f1 = sub(f, i)
            expr = f"QC{X[i]} * {name_of(e, f1, m)} + WQ{X[i]} * {name_of(e, f1, m + 1)}"
            nflop += 3
            if f1[i] > 0:
                f2 = sub(f1, i)
                expr += f" + R({f1[i]}) * oo2e * ({name_of(e, f2, m)} - roe * {name_of(e, f2, m + 1)})"
                nflop += 5
            if e[i] > 0:
                expr += f" + R({e[i]}) * oo2ze * {name_of(sub(e, i), f1, m + 1)}"
                nflop += 3
        else:
            i = first_dir(e)
            e1 = sub(e, i)
            expr = f"PA{X[i]} * {name_of(e1, f, m)} + WP{X[i]} * {name_of(e1, f, m + 1)}"
            nflop += 3
            if e1[i] > 0:
                e2 = sub(e1, i)
                expr += f" + R({e1[i]}) * oo2z * ({name_of(e2, f, m)} - roz * {name_of(e2, f, m + 1)})"
                nflop += 5
        lines.append(f"    const R {nm} = {expr};")
    ne, nf = len(erange(la, E)), len(erange(lc, F))
    acc = []
    for a, e in enumerate(erange(la, E)):
        for b, f in enumerate(erange(lc, F)):
            acc.append(f"    acc[{a * nf + b}] += {name_of(e, f, 0)};")
            nflop += 1
    body = "\n".join(lines + acc)
    sig = (f"template <typename R>\n__device__ __forceinline__ void vrr_{tag(cls)}(const R* __restrict__ F, "
           "R PAx, R PAy, R PAz, R WPx, R WPy, R WPz, R QCx, R QCy, R QCz, R WQx, R WQy, R WQz, "
           "R oo2z, R oo2e, R oo2ze, R roz, R roe, R* __restrict__ acc)")
    return f"{sig}\n{{\n{body}\n}}\n", ne * nf, nflop, len(order)


def gen_hrr(cls):
    """acc[e][f] -> out[a][b][c][d] in QUICK's component order."""
    la, lb, lc, ld = cls
    E, F = la + lb, lc + ld
    elist, flist = erange(la, E), erange(lc, F)
    eidx = {e: k for k, e in enumerate(elist)}
    fidx = {f: k for k, f in enumerate(flist)}
    nf = len(flist)
    X = "xyz"
    lines = []
    nflop = 0

    # bra HRR for every f: names hb_<a>_<b>_<f>
    def hb(a, b, fk):
        return "hb_%d%d%d_%d%d%d_%d" % (a + b + (fk,))

    @lru_cache(maxsize=None)
    def need_bra(a, b):
        return True

    bra_terms = {}

    def bra_expr(a, b, fk):
        key = (a, b, fk)
        if key in bra_terms:
            return bra_terms[key]
        if sum(b) == 0:
            nm = f"acc[{eidx[a] * nf + fk}]"
            bra_terms[key] = nm
            return nm
        i = first_dir(b)
        b1 = sub(b, i)
        a1 = list(a)
        a1[i] += 1
        a1 = tuple(a1)
        x1 = bra_expr(a1, b1, fk)
        x2 = bra_expr(a, b1, fk)
        nm = hb(a, b, fk)
        lines.append(f"    const R {nm} = {x1} + AB{X[i]} * {x2};")
        nonlocal_flops[0] += 2
        bra_terms[key] = nm
        return nm

    nonlocal_flops = [0]
    A_list, B_list = QUICK_ORDER[la], QUICK_ORDER[lb]
    C_list, D_list = QUICK_ORDER[lc], QUICK_ORDER[ld]
    # (ab|f0) for the needed f: all f in flist
    for a in A_list:
        for b in B_list:
            for fk in range(nf):
                bra_expr(a, b, fk)

    ket_terms = {}

    def ket_expr(abkey, c, d):
        key = (abkey, c, d)
        if key in ket_terms:
            return ket_terms[key]
        a, b = abkey
        if sum(d) == 0:
            nm = bra_terms[(a, b, fidx[c])]
            ket_terms[key] = nm
            return nm
        i = first_dir(d)
        d1 = sub(d, i)
        c1 = list(c)
        c1[i] += 1
        c1 = tuple(c1)
        x1 = ket_expr(abkey, c1, d1)
        x2 = ket_expr(abkey, c, d1)
        nm = "hk_%d%d%d_%d%d%d_%d%d%d_%d%d%d" % (a + b + c + d)
        lines.append(f"    const R {nm} = {x1} + CD{X[i]} * {x2};")
        nonlocal_flops[0] += 2
        ket_terms[key] = nm
        return nm

    outs = []
    k = 0
    for a in A_list:
        for b in B_list:
            for c in C_list:
                for d in D_list:
                    outs.append(f"    out[{k}] = {ket_expr((a, b), c, d)};")
                    k += 1
    sig = (f"template <typename R>\n__device__ __forceinline__ void hrr_{tag(cls)}(const R* __restrict__ acc, "
           "R ABx, R ABy, R ABz, R CDx, R CDy, R CDz, R* __restrict__ out)")
    return f"{sig}\n{{\n" + "\n".join(lines + outs) + "\n}\n", nonlocal_flops[0], k


def write_units(outdir):
    """One translation unit per quartet class (and per Schwarz diagonal class),
    so hipcc compiles the heavy classes in parallel."""
    import os
    os.makedirs(outdir, exist_ok=True)
    names = []
    for cls in quartet_classes():
        t = tag(cls)
        body = ['// Generated by quick_amd/csrc/gen_eri.py -- do not edit.', '#include "../qk_kernels.h"',
                'namespace qk {',
                f'void launch_fock_{t}(const FockArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s) {{',
            ] + ([
                '    if (a.stage_d) {',
                f'        const void* k = reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, true>);',
                '        if (lds > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);',
                f'        hipLaunchKernelGGL((fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, true>), dim3(grid), dim3(256), lds, s, a, w);',
                '        return;',
                '    }'] if cls not in TEAM_CLASSES else []) + [
                f'    const void* k = reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>);',
                '    if (lds > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);',
                f'    hipLaunchKernelGGL((fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>), dim3(grid), dim3(256), lds, s, a, w);',
                '}',
                f'size_t fock_scratch_{t}() {{',
                '    hipFuncAttributes at{};',
                f'    if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>)) != hipSuccess) return 0;',
            ] + ([
                '    hipFuncAttributes st{};   // the staged instantiation: the larger of the two',
                f'    if (hipFuncGetAttributes(&st, reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, true>)) != hipSuccess) return 0;',
                '    return at.localSizeBytes > st.localSizeBytes ? at.localSizeBytes : st.localSizeBytes;',
            ] if cls not in TEAM_CLASSES else ['    return at.localSizeBytes;']) + [
                '}',
                f'int fock_blocks_{t}(size_t lds, bool staged) {{',
                '    int nb = 0;',
                (f'    const void* k = staged ? reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, true>) : reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>);'
                 if cls not in TEAM_CLASSES else
                 f'    const void* k = reinterpret_cast<const void*>(&fock_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>); (void)staged;'),
                '    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, 256, lds) != hipSuccess) return 0;',
                '    return nb;',
                '}']
        if cls not in TEAM_CLASSES:
            body += [f'void launch_fock_f32_{t}(const FockArgs& a, const ClassWork& w, int grid, size_t lds, hipStream_t s) {{',
                     f'    const void* k = reinterpret_cast<const void*>(&fock_kernel_f32<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>);',
                     '    if (lds > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);',
                     f'    hipLaunchKernelGGL((fock_kernel_f32<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>), dim3(grid), dim3(256), lds, s, a, w);',
                     '}',
                     f'size_t fock_scratch_f32_{t}() {{',
                     '    hipFuncAttributes at{};',
                     f'    if (hipFuncGetAttributes(&at, reinterpret_cast<const void*>(&fock_kernel_f32<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}>)) != hipSuccess) return 0;',
                     '    return at.localSizeBytes;',
                     '}']
        if cls in TEAM_CLASSES:
            body = body[:2] + ['#include "../team_tables.inc"', '#include "../qk_team.h"'] + body[2:]
            for R, sfx in (("double", ""), ("float", "_f32")):
                body += [f'void launch_team{sfx}_{t}(const FockArgs& a, const ClassWork& w, int grid, hipStream_t s) {{',
                         f'    constexpr unsigned lds = team_lds_bytes<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, Team_{t}, {R}>();',
                         f'    const void* k = reinterpret_cast<const void*>(&team_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, Team_{t}, {R}>);',
                         '    if (lds > 64 * 1024) (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);',
                         f'    hipLaunchKernelGGL((team_kernel<{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, Team_{t}, {R}>), dim3(grid), dim3(64), lds, s, a, w);',
                         '}']
        if cls[:2] == cls[2:]:
            pt = t[:2]
            body += [f'void launch_schwarz_{pt}(const int* pairs, int np, const int* qa, const int* qb, const int* prim0,',
                     '                        const int* nprim, const DevQshells& qs, const DevPrims& pp, const double* xyz,',
                     '                        const double* cons, double* ymax, double* pmax, hipStream_t s) {',
                     f'    static const long long cap = scratch_grid_cap(kernel_scratch(reinterpret_cast<const void*>(&schwarz_kernel<{cls[0]}, {cls[1]}>)), 4);',
                     '    // one thread per pair, in launches of at most cap blocks (outputs are indexed by pair)',
                     '    for (long long off = 0; off < np; off += cap * 256) {',
                     '        const int m = (int)std::min<long long>(np - off, cap * 256);',
                     f'        hipLaunchKernelGGL((schwarz_kernel<{cls[0]}, {cls[1]}>), dim3((m + 255) / 256), dim3(256), 0, s,',
                     '                           pairs + off, m, qa, qb, prim0, nprim, qs, pp, xyz, cons, ymax, pmax);',
                     '    }',
                     '}']
        body.append('}  // namespace qk')
        fn = os.path.join(outdir, f'fock_{t}.hip')
        new = "\n".join(body) + "\n"
        if not os.path.exists(fn) or open(fn).read() != new:
            open(fn, 'w').write(new)
        names.append(fn)
    return names


def main(path):
    out = ["// Generated by quick_amd/csrc/gen_eri.py -- do not edit.",
           "// Per-class VRR/HRR for (la lb|lc ld), la>=lb, lc>=ld, bra pair class >= ket pair class.",
           "#pragma once", ""]
    table = []
    for cls in quartet_classes():
        v, nacc, vflop, nterms = gen_vrr(cls)
        h, hflop, nout = gen_hrr(cls)
        out += [v, h]
        table.append((cls, nacc, vflop, hflop, nout))
    hdr = ["// Generated by quick_amd/csrc/gen_eri.py -- do not edit.",
           "// Arithmetic operations of the generated code, per class (FMA = 2):",
           "// la, lb, lc, ld, accumulator size, VRR ops per primitive quartet, HRR ops per quartet, outputs",
           "#pragma once",
           "static const int kEriFlopTable[][8] = {"]
    for cls, nacc, vflop, hflop, nout in table:
        hdr.append(f"    {{{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, {nacc}, {vflop}, {hflop}, {nout}}},")
    hdr.append("};")
    import os
    fpath = os.path.join(os.path.dirname(path), "eri_flops.h")
    newh = "\n".join(hdr) + "\n"
    if not os.path.exists(fpath) or open(fpath).read() != newh:
        open(fpath, "w").write(newh)
    out.append("// class, accumulator size, VRR flops per primitive quartet, HRR flops, outputs")
    out.append("struct EriClassInfo { int la, lb, lc, ld, nacc, vrr_flops, hrr_flops, nout; };")
    out.append("static const EriClassInfo kEriClassInfo[] = {")
    for cls, nacc, vflop, hflop, nout in table:
        out.append(f"    {{{cls[0]}, {cls[1]}, {cls[2]}, {cls[3]}, {nacc}, {vflop}, {hflop}, {nout}}},")
    out.append("};")
    new = "\n".join(out) + "\n"
    import os
    if not os.path.exists(path) or open(path).read() != new:
        open(path, "w").write(new)
    for cls, nacc, vflop, hflop, nout in table:
        print(tag(cls), nacc, vflop, hflop, nout)


# ---------------------------------------------------------------------------
# Wavefront-cooperative ("team") form of the heavy classes: one wavefront
# evaluates one quartet, with every intermediate in LDS.  The VRR becomes a
# node program executed level by level (all nodes of a level are independent),
# the HRR and the digestion become index tables the lanes stride through.

TEAM_CLASSES = [(2, 1, 2, 1), (2, 2, 1, 1), (2, 2, 2, 0), (2, 2, 2, 1), (2, 2, 2, 2)]


def team_program(cls):
    """Nodes of the VRR DAG for class cls, levelled.  Slot 0 holds 0.0."""
    la, lb, lc, ld = cls
    E, F = la + lb, lc + ld
    targets = [(e, f, 0) for e in erange(la, E) for f in erange(lc, F)]
    deps_of = {}
    order = []
    seen = set()

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
        order.append(node)

    for t in targets:
        visit(t)
    level = {}
    for node in order:
        d = deps_of[node]
        lv = 0
        for k in ("a", "b", "c", "d", "x"):
            if k in d:
                lv = max(lv, level[d[k]] + 1)
        level[node] = lv
    nlev = max(level.values()) + 1
    # slots: 0 = zero, then nodes ordered by level
    by_level = [[n for n in order if level[n] == L] for L in range(nlev)]
    slot = {}
    s = 1
    for L in range(nlev):
        for n in by_level[L]:
            slot[n] = s
            s += 1
    nslots = s
    nodes = []      # uint4 per node
    lev_start = [0]
    for L in range(nlev):
        for n in by_level[L]:
            d = deps_of[n]
            g = lambda k: slot[d[k]] if k in d else 0
            if d["kind"] == 0:
                flags = (n[2] << 8)            # F_m: m in bits 8..15
            else:
                flags = d["kind"] | (d["dir"] << 2) | (d.get("n1", 0) << 4) | (d.get("ne", 0) << 8)
            nodes.append((slot[n] | (g("a") << 16), g("b") | (g("c") << 16), g("d") | (g("x") << 16), flags))
        lev_start.append(len(nodes))
    tslots = [slot[t] for t in targets]
    return nodes, lev_start, tslots, nslots


def binom_terms(a, b):
    """(a b| = sum_k prod_i C(b_i,k_i) AB_i^(b_i-k_i) (a+k 0|, as (e, px, py, pz, coef)."""
    from math import comb
    out = []
    for kx in range(b[0] + 1):
        for ky in range(b[1] + 1):
            for kz in range(b[2] + 1):
                e = (a[0] + kx, a[1] + ky, a[2] + kz)
                out.append((e, b[0] - kx, b[1] - ky, b[2] - kz, comb(b[0], kx) * comb(b[1], ky) * comb(b[2], kz)))
    return out


def team_tables(cls):
    la, lb, lc, ld = cls
    E, F = la + lb, lc + ld
    nodes, lev, tslots, nslots = team_program(cls)
    elist, flist = erange(la, E), erange(lc, F)
    eidx = {e: k for k, e in enumerate(elist)}
    fidx = {f: k for k, f in enumerate(flist)}
    A, B, C, D = QUICK_ORDER[la], QUICK_ORDER[lb], QUICK_ORDER[lc], QUICK_ORDER[ld]
    bra = []   # per (a,b): 4 terms (eidx, px|py<<2|pz<<4, coef) padded with coef 0
    for a in A:
        for b in B:
            t = binom_terms(a, b)
            t = [(eidx[e], px | (py << 2) | (pz << 4), c) for e, px, py, pz, c in t]
            t += [(0, 0, 0)] * (4 - len(t))
            bra.append(t)
    ket = []
    for c in C:
        for d in D:
            t = binom_terms(c, d)
            t = [(fidx[e], px | (py << 2) | (pz << 4), cc) for e, px, py, pz, cc in t]
            t += [(0, 0, 0)] * (4 - len(t))
            ket.append(t)
    return nodes, lev, tslots, nslots, bra, ket, len(elist), len(flist)


def write_team(path):
    import os
    out = ["// Generated by quick_amd/csrc/gen_eri.py -- do not edit.",
           "// Node programs and HRR tables of the wavefront-cooperative heavy classes.",
           "#pragma once", "#include <cstdint>", "namespace qk {"]
    for cls in TEAM_CLASSES:
        t = tag(cls)
        nodes, lev, tslots, nslots, bra, ket, ne, nf = team_tables(cls)
        out.append(f"__device__ const uint4 kTeamNodes_{t}[{len(nodes)}] = {{")
        out.append(",\n".join("    {%du, %du, %du, %du}" % nd for nd in nodes))
        out.append("};")
        out.append(f"__device__ const int kTeamLevels_{t}[{len(lev)}] = {{{', '.join(map(str, lev))}}};")
        out.append(f"__device__ const unsigned short kTeamTargets_{t}[{len(tslots)}] = {{{', '.join(map(str, tslots))}}};")
        # one int per term: index | power code << 12 | binomial << 18 (0 = padding)
        flat = lambda tab: ", ".join(str(e | (p3 << 12) | (c << 18)) for row in tab for (e, p3, c) in row)
        out.append(f"__device__ const int __attribute__((aligned(16))) kTeamBra_{t}[{len(bra) * 4}] = {{{flat(bra)}}};")
        out.append(f"__device__ const int __attribute__((aligned(16))) kTeamKet_{t}[{len(ket) * 4}] = {{{flat(ket)}}};")
        out.append(f"struct Team_{t} {{")
        out.append(f"    static constexpr int kNodes = {len(nodes)}, kLevels = {len(lev) - 1}, kSlots = {nslots};")
        out.append(f"    static constexpr int kNE = {ne}, kNF = {nf};")
        out.append(f"    static __device__ __forceinline__ const uint4* nodes() {{ return kTeamNodes_{t}; }}")
        out.append(f"    static __device__ __forceinline__ const int* levels() {{ return kTeamLevels_{t}; }}")
        out.append(f"    static __device__ __forceinline__ const unsigned short* targets() {{ return kTeamTargets_{t}; }}")
        out.append(f"    static __device__ __forceinline__ const int* bra() {{ return kTeamBra_{t}; }}")
        out.append(f"    static __device__ __forceinline__ const int* ket() {{ return kTeamKet_{t}; }}")
        out.append("};")
    out.append("}  // namespace qk")
    new = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != new:
        open(path, "w").write(new)
    for cls in TEAM_CLASSES:
        nodes, lev, tslots, nslots, bra, ket, ne, nf = team_tables(cls)
        print(tag(cls), "nodes", len(nodes), "levels", len(lev) - 1, "slots", nslots,
              "max level", max(lev[i + 1] - lev[i] for i in range(len(lev) - 1)))


if __name__ == "__main__":
    main(sys.argv[1])
    if len(sys.argv) > 2:
        import os
        write_team(os.path.join(os.path.dirname(sys.argv[1]), "team_tables.inc"))
        write_units(sys.argv[2])
