#!/usr/bin/env python3
"""Benchmark: QUICK's SCF Fock-build hot path and four-centre gradient on one MI355X.

Workload: (H2O)32, HF/6-31G** (Cartesian d, 800 basis functions), seeded water
cluster (4x4x2 grid, 3.1 A spacing, +-0.3 A jitter, seed 2), Schwarz cutoff
tau = 1e-11 (primitive cutoff 1e-20).  One "step" is one two-electron operator
build G(D) for a fixed density: the density (n x n) and the operator are
resident in HBM; the step runs the density prep, every quartet class kernel and
the fixed-point finalisation.  The density is the converged RHF density of the
reference's own SCF on this geometry (QUICK's Fortran built from source, run to
DENSERMS 1e-7; tests/golden/water32_scf.npz); --density synthetic takes a seeded
synthetic one instead (entries decaying as exp(-0.25 R^2) between function
centres).

Legs:
  * all-FP64 builds (the reference-parity path);
  * mixed precision: quartets with Schwarz bound Y_ij*Y_kl below a threshold in
    FP32, the rest FP64, accumulation in FP64 fixed point; swept over
    thresholds, each checked against the FP64 build (|dE|, max |dF|); the
    headline value is the fastest threshold whose |dE| <= 1 uEh;
  * the four-centre nuclear gradient dE2/dR on (H2O)32 and on n-icosane
    C20H42 (510 bf), seconds per gradient: FP64, and mixed precision with the
    FP32 threshold swept, the fastest whose max |dgrad| against FP64 is
    <= 1e-6 Eh/bohr reported (with Fock build + gradient, FP64 against mixed);
  * (H2O)64 B3LYP/6-31G** (1,600 bf): the two-electron build at exchange
    weight 0.2 (FP64 and mixed), the XC step on the 75 x 302 grid (FP64 MFMA
    contractions) and the four-centre gradient (b3lyp_water64);
  * the (H2O)32 RHF SCF on the device from the core guess (FP64 and mixed Fock
    builds) against the reference's SCF energy, and the one-electron integrals
    and their gradient terms (scf_water32);
  * the reference CPU path (QUICK's Fortran built from source into
    oracle/_ref) timed on a bounded sample of the same build (cpu_baseline).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
                       [--thresholds 1e-3,1e-4,...|none] [--grad-thresholds 1e-1,...]
                       [--density converged|synthetic] [--no-gradient] [--no-scf]
                       [--no-water64] [--copy-kernel-mib N]
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6       # MI355X FP64 vector (and matrix) peak, AMD spec
PEAK_FP32_TFLOPS = 157.3      # MI355X FP32 vector peak, AMD spec
METRIC = "Fock builds/s + gradient s, 32-H2O HF/6-31G**: mixed vs FP64 at ≤1µEh"
DENSITY_SEED = 7


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(symbols, xyz_ang, D, stride=16, timeout=300):
    """The reference CPU path (oracle/_ref/qk_ref_driver, QUICK's Fortran built
    from its sources) on every stride-th outer shell of the same Fock build,
    scaled to a full build.  Returns the cpu_baseline object or None."""
    drv = ROOT / "oracle" / "_ref" / "qk_ref_driver"
    if not drv.exists():
        log("reference driver not built; cpu_baseline is null")
        return None
    from quick_amd.basis import write_quick_basis_file
    from quick_amd.geometry import to_quick_input
    with tempfile.TemporaryDirectory() as td:
        td = Path(td)
        (td / "basis").mkdir()
        write_quick_basis_file(td / "basis" / "6-31GSS")
        (td / "job.in").write_text(to_quick_input(symbols, xyz_ang,
                                                  "HF BASIS=6-31GSS CUTOFF=1.0e-11 DENSERMS=1.0e-7"))
        np.asarray(D, np.float64).ravel(order="F").tofile(td / "D.bin")
        (td / "out").mkdir()
        env = dict(os.environ, QUICK_BASIS=str(td / "basis"), OMP_NUM_THREADS="1")
        t0 = time.time()
        r = subprocess.run(f"ulimit -s unlimited && exec {drv} job.in fock_sample D.bin out {stride}",
                           shell=True, cwd=td, env=env, stdout=subprocess.DEVNULL,
                           stderr=subprocess.PIPE, text=True, timeout=timeout)
        wall = time.time() - t0
        f = td / "out" / "time_sample.bin"
        if not f.exists():
            log(f"reference driver failed (rc={r.returncode}): {r.stderr[-400:]}")
            return None
        t, nsamp, nsh = np.fromfile(f)
    est = t * nsh / nsamp
    log(f"cpu baseline: {int(nsamp)} of {int(nsh)} shells in {t:.2f} s (wall {wall:.1f} s) -> {est:.1f} s/build")
    return {"value": 1.0 / est, "unit": "Fock builds/s", "cores": 1, "kind": "reference",
            "sample": f"QUICK Fortran CPU Fock build (hfoperator get2e) over every {stride}th outer "
                      f"shell, {int(nsamp)} of {int(nsh)}, {t:.1f} s, scaled by {int(nsh)}/{int(nsamp)}",
            "seconds_per_build_est": est}


def scf_leg(b, eng, syms, torch, dev, stream, sptr):
    """(H2O)32 RHF SCF on the device (qk_scf) from the core-Hamiltonian guess to
    DENSERMS 1e-7, with FP64 and with mixed-precision (1e-3) Fock builds,
    against the reference's own SCF energy on this geometry; and the
    one-electron integrals (S, H_core) and their gradient terms on the device,
    timed with HIP events on the launch stream."""
    import ctypes as C
    from quick_amd.fock import QK_MIXED, lib, make_opts
    zc = np.load(ROOT / "tests" / "golden" / "water32_scf.npz")
    Zof = {"H": 1.0, "C": 6.0, "N": 7.0, "O": 8.0}
    eng.set_charges([Zof[s] for s in syms])
    nocc = int(b.nelec) // 2
    out = {"reference_E": float(zc["E_total"]), "reference_cycles": int(zc["scf_cycles"]),
           "reference_cpu_s": float(zc["time_scf"])}
    for lab, opts in (("fp64", make_opts()), ("mixed", make_opts(precision=QK_MIXED, fp32_threshold=1e-3))):
        eng.scf(nocc, opts, max_cycles=3, require_convergence=False)   # warm-up: work lists, solver set-up
        t0 = time.perf_counter()
        _, E, cyc, _ = eng.scf(nocc, opts)
        dt = time.perf_counter() - t0
        out[lab] = {"seconds": dt, "cycles": cyc, "E": E, "abs_dE_vs_reference": abs(E - out["reference_E"])}
        log(f"(H2O)32 SCF {lab}: {dt:.2f} s, {cyc} cycles, E {E:.10f}, |dE| vs reference "
            f"{abs(E - out['reference_E']):.2e}")
    # one-electron legs at the reference's converged density D = 2 C_occ C_occ^T
    # and energy-weighted density W = 2 C_occ diag(eps_occ) C_occ^T
    L, Co = lib(), np.asarray(zc["C_occ"], np.float64)
    eps = np.asarray(zc["eps"], np.float64)[:Co.shape[1]]
    dev_f = lambda M: torch.from_numpy(np.ascontiguousarray(M.ravel(order="F"))).to(dev)
    Dd, Wd = dev_f(2.0 * Co @ Co.T), dev_f(2.0 * (Co * eps) @ Co.T)
    Sd, Hd = torch.empty_like(Dd), torch.empty_like(Dd)
    gd = torch.zeros(3 * len(syms), dtype=torch.float64, device=dev)
    p = lambda t: C.c_void_p(t.data_ptr())
    legs = {"oneint_S_H": lambda: L.qk_oneint_device(eng.h, p(Sd), p(Hd), C.c_void_p(sptr)),
            "oneint_grad": lambda: L.qk_oneint_grad_device(eng.h, p(Dd), p(Wd), p(gd), C.c_void_p(sptr))}
    for lab, fn in legs.items():
        if fn() != 0:
            raise RuntimeError(f"{lab} failed: {L.qk_last_error().decode()}")
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(5):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        out[lab + "_ms"] = e0.elapsed_time(e1) / 5
        log(f"(H2O)32 {lab}: {out[lab + '_ms']:.3f} ms")
    return out


def water64_leg(args, torch, dev, stream, sptr):
    """(H2O)64 B3LYP/6-31G** (1,600 bf, seed 5): the two-electron build at exchange
    weight 0.2 (FP64 and mixed at 1e-3), the XC step on the 75 x 302 product grid
    (FP64 MFMA contractions; a non-negative diagonal density with a known
    integral), and the four-centre gradient (FP64 and mixed at 1e-2)."""
    from quick_amd.basis import build_basis
    from quick_amd.fock import FockEngine, make_opts, QK_FP64, QK_MIXED
    from quick_amd.geometry import CONFIGS, synthetic_density
    from quick_amd.xc import QK_GRID_PRODUCT, XcEngine, make_xc_opts
    syms, xyz = CONFIGS["water64"]()
    b = build_basis(syms, xyz)
    D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, DENSITY_SEED)
    eng = FockEngine(b)
    t0 = time.perf_counter()
    eng.compute_schwarz()
    out = {"atoms": len(syms), "nbasis": int(b.nbasis), "schwarz_s": time.perf_counter() - t0}
    Dd = torch.from_numpy(np.ascontiguousarray(D.ravel(order="F"))).to(dev)
    Gd = torch.empty_like(Dd)
    reps = max(1, min(args.steps, 3))

    def evtime(fn, n):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(n):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    Gs = {}
    for lab, opts in (("fp64", make_opts(hybrid=0.2, precision=QK_FP64)),
                      ("mixed", make_opts(hybrid=0.2, precision=QK_MIXED, fp32_threshold=1e-3))):
        eng.prepare(opts)
        eng.fock_device(Dd.data_ptr(), Gd.data_ptr(), opts, sptr)
        torch.cuda.synchronize()
        msb = evtime(lambda: eng.fock_device(Dd.data_ptr(), Gd.data_ptr(), opts, sptr), reps)
        Gs[lab] = Gd.cpu().numpy().reshape(b.nbasis, b.nbasis, order="F")
        st = eng.stats()
        out[f"fock_{lab}_ms"] = msb
        out[f"fock_{lab}_quartets"] = int(st.quartets)
        log(f"(H2O)64 hybrid 0.2 {lab}: {msb:.2f} ms/build, {st.quartets} quartets")
    out["fock_mixed_abs_dE_vs_fp64"] = abs(0.5 * float(np.sum(D * (Gs["mixed"] - Gs["fp64"]))))
    out["fock_mixed_max_abs_dF_vs_fp64"] = float(np.abs(Gs["mixed"] - Gs["fp64"]).max())
    # XC on the 75 x 302 grid
    rng = np.random.default_rng(3)
    dg = rng.uniform(0.2, 1.0, b.nbasis)
    dg *= b.nelec / dg.sum()
    Dx = torch.from_numpy(np.ascontiguousarray(np.diag(dg).ravel(order="F"))).to(dev)
    Vx = torch.empty_like(Dx)
    xe = XcEngine(b)
    t0 = time.perf_counter()
    npts = xe.build_grid(make_xc_opts(grid=QK_GRID_PRODUCT, nrad=75, nang=302))
    grid_s = time.perf_counter() - t0
    nb, mf, flops = xe.batch_info()
    E, N = xe.xc_device(Dx.data_ptr(), Vx.data_ptr(), sptr)
    torch.cuda.synchronize()
    msx = evtime(lambda: xe.xc_device(Dx.data_ptr(), Vx.data_ptr(), sptr, want_scalars=False), reps)
    out["xc"] = {"grid": "75 radial x 302 Lebedev per atom, SSF weights", "points": int(npts),
                 "grid_build_s": grid_s, "batches": nb, "mean_functions_per_batch": mf, "ms": msx,
                 "mfma_gflop": flops / 1e9, "mfma_tflops_over_whole_step": flops / (msx * 1e-3) / 1e12,
                 "electrons": N, "electrons_exact": int(b.nelec), "E_xc": E}
    log(f"(H2O)64 XC: {npts} points, grid {grid_s:.2f} s, {msx:.2f} ms per evaluation, "
        f"{flops / 1e9:.0f} GFLOP MFMA ({flops / (msx * 1e-3) / 1e12:.1f} TFLOP/s over the step), N {N:.6f}")
    # four-centre gradient
    for lab, opts in (("fp64", make_opts(hybrid=0.2, precision=QK_FP64)),
                      ("mixed", make_opts(hybrid=0.2, precision=QK_MIXED, fp32_threshold=1e-2))):
        g = torch.zeros(3 * eng.natom, dtype=torch.float64, device=dev)
        eng.grad_device(Dd.data_ptr(), g.data_ptr(), opts, sptr)
        torch.cuda.synchronize()
        out[f"grad_{lab}_ms"] = evtime(lambda: eng.grad_device(Dd.data_ptr(), g.data_ptr(), opts, sptr), 1)
        log(f"(H2O)64 gradient {lab}: {out[f'grad_{lab}_ms']:.1f} ms")
    xe.close() if hasattr(xe, "close") else None
    return out


def kernel_of(tag):
    """name of the Fock kernel that evaluates quartet class `tag` (as rocprofv3 lists it, short form)"""
    return "team_kernel" if tag in ("dpdp", "ddpp", "ddds", "dddp", "dddd") else "fock_kernel"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--thresholds", default="1e-3,1e-4,1e-5,1e-6",
                    help="FP32 thresholds of the mixed-precision sweep ('none': FP64 only)")
    ap.add_argument("--no-gradient", action="store_true", help="skip the gradient legs")
    ap.add_argument("--grad-thresholds", default="1e-1,1e-2,1e-3",
                    help="FP32 thresholds of the mixed-precision gradient sweep; the fastest whose max "
                         "|dgrad| against the FP64 gradient is <= 1e-6 Eh/bohr is reported")
    ap.add_argument("--density", choices=("converged", "synthetic"), default="converged",
                    help="(H2O)32 density: the reference's converged RHF density (tests/golden/water32_scf.npz) "
                         "or the seeded synthetic one")
    ap.add_argument("--no-scf", action="store_true", help="skip the (H2O)32 SCF leg")
    ap.add_argument("--no-water64", action="store_true",
                    help="skip the (H2O)64 B3LYP leg (1,600 bf: ERI build at exchange weight 0.2, XC on the "
                         "75 x 302 grid, four-centre gradient)")
    ap.add_argument("--copy-kernel-mib", type=int, default=0,
                    help="only run a plain device-to-device copy of this many MiB (5 times), to "
                         "calibrate the FETCH_SIZE/WRITE_SIZE counters, and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "the multi-GPU sharding is a design note, not an implemented path", file=sys.stderr)
        sys.exit(2)

    import torch
    if args.copy_kernel_mib:
        n = args.copy_kernel_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_kernel_mib} MiB x 5")
        return
    from quick_amd.basis import build_basis
    from quick_amd.fock import FockEngine, make_opts, QK_FP64, QK_MIXED
    from quick_amd.geometry import CONFIGS, synthetic_density

    syms, xyz = CONFIGS["water32"]()
    b = build_basis(syms, xyz)
    if args.density == "converged":
        # the reference's own SCF on this geometry (QUICK's Fortran, DENSERMS 1e-7):
        # D = 2 C_occ C_occ^T, on the coordinates exactly as the reference read them
        zc = np.load(ROOT / "tests" / "golden" / "water32_scf.npz")
        b.xyz = np.asarray(zc["xyz"], np.float64).copy()
        D = 2.0 * zc["C_occ"] @ zc["C_occ"].T
    else:
        D = synthetic_density(b.xyz, b.katom, b.ksumtype, b.nbasis, DENSITY_SEED)
    eng = FockEngine(b)
    t0 = time.time()
    eng.compute_schwarz()
    setup_s = time.time() - t0
    log(f"system: {len(syms)} atoms, {b.nbasis} bf; Schwarz bounds {setup_s:.3f} s")

    dev = torch.device("cuda:0")
    Dd = torch.from_numpy(np.ascontiguousarray(D.ravel(order="F"))).to(dev)
    Gd = torch.empty_like(Dd)
    stream = torch.cuda.current_stream(dev)
    sptr = stream.cuda_stream

    prep_s = {}
    repro = {}

    def timed(opts, label):
        """warmup + `steps` timed builds with `opts` (no per-kernel events inside
        the timed region), then a few builds with per-kernel HIP events for the
        per-class figures; returns ms/build, G, per-kernel rows, stats"""
        tp = time.perf_counter()
        eng.prepare(opts)                     # work lists (once per geometry and option set)
        prep_s[label] = time.perf_counter() - tp
        log(f"{label}: prepare (work lists) {prep_s[label] * 1e3:.1f} ms")
        eng.set_timing(False)
        for _ in range(args.warmup):
            eng.fock_device(Dd.data_ptr(), Gd.data_ptr(), opts, sptr)
        torch.cuda.synchronize()
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        ev0.record(stream)
        for _ in range(args.steps):
            eng.fock_device(Dd.data_ptr(), Gd.data_ptr(), opts, sptr)
        ev1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        ms = ev0.elapsed_time(ev1) / args.steps
        G = Gd.cpu().numpy().reshape(b.nbasis, b.nbasis, order="F")
        # per-class kernel times (HIP events around every class launch)
        eng.set_timing(True)
        for _ in range(max(1, min(args.steps, 5))):
            eng.fock_device(Dd.data_ptr(), Gd.data_ptr(), opts, sptr)
        torch.cuda.synchronize()
        # the fixed-point accumulation makes every build bit-identical
        repro[label] = bool(np.array_equal(G, Gd.cpu().numpy().reshape(b.nbasis, b.nbasis, order="F")))
        log(f"{label}: {args.steps} builds {ms:.3f} ms/build (event), wall {wall * 1e3 / args.steps:.3f} ms/build"
            f"{'' if repro[label] else ' (NOT bit-identical across builds)'}")
        # kernel rows: (kernel, tag, quartets, prim quartets, flops, ms, precision)
        rows = [(f"{kernel_of(c[0])}<{c[0]}>", c[0], c[1], c[2], c[3], c[4], "fp64") for c in eng.class_stats()]
        rows += [(f"{kernel_of(c[0])}_f32<{c[0]}>", c[0], c[1], c[2], c[3], c[4], "fp32")
                 for c in eng.class_stats_fp32()]
        rows = [r for r in rows if r[2]]
        eng.set_timing(False)
        return ms, G, rows, eng.stats()

    def show(rows):
        for r in sorted(rows, key=lambda r: -r[5]):
            log(f"  {r[0]:<24s} {r[2]:>10d} quartets {r[3]:>11d} prim {r[5]:8.3f} ms "
                f"{r[4] / max(r[5], 1e-9) / 1e9:8.2f} TFLOP/s")

    # all-FP64 build (the parity path)
    ms64, G64, rows64, st64 = timed(make_opts(precision=QK_FP64), "fp64")
    show(rows64)
    E64 = 0.5 * float(np.sum(D * G64))

    # mixed precision: FP32 below the Schwarz-bound threshold; the threshold is
    # the largest of the sweep whose energy error against FP64 is <= 1 uEh
    sweep = []
    thresholds = [] if args.thresholds == "none" else [float(t) for t in args.thresholds.split(",")]
    for thr in thresholds:
        msm, Gm, rowsm, stm = timed(make_opts(precision=QK_MIXED, fp32_threshold=thr), f"mixed {thr:.0e}")
        dE = abs(0.5 * float(np.sum(D * Gm)) - E64)
        dF = float(np.abs(Gm - G64).max())
        sweep.append({"fp32_threshold": thr, "ms_per_build": msm, "builds_per_s": 1000.0 / msm,
                      "fp32_quartet_share": stm.quartets_fp32 / max(stm.quartets, 1),
                      "abs_dE_vs_fp64": dE, "max_abs_dF_vs_fp64": dF, "_rows": rowsm, "_st": stm})
        log(f"  fp32 share {stm.quartets_fp32 / max(stm.quartets, 1):.3f} |dE| {dE:.2e} max|dF| {dF:.2e}")
    ok = [r for r in sweep if r["abs_dE_vs_fp64"] <= 1e-6]
    pick = max(ok, key=lambda r: r["builds_per_s"]) if ok else None
    if pick is not None:
        ms, rows, st, precision = pick["ms_per_build"], pick["_rows"], pick["_st"], "mixed"
        show(rows)
    else:
        ms, rows, st, precision = ms64, rows64, st64, "fp64"
    for r in sweep:
        r.pop("_rows"), r.pop("_st")
    thr = pick["fp32_threshold"] if pick is not None else None

    # the roofline kernel: the longest launch of the build's critical path.  In a
    # mixed build the FP64 lists (the strongest pairs) run on a second stream
    # beside the FP32 lists, so their event times overlap other kernels; the
    # FP32 lists' stream is the critical path.  In an FP64 build the classes
    # alternate between two streams, so every launch's event time includes the
    # overlap with the other stream's kernel
    crit = [r for r in rows if r[6] == "fp32"] if precision == "mixed" else rows
    dom = max(crit, key=lambda r: r[5])
    kname, tag, nq, npq, flops, kms, kprec = dom
    peak = PEAK_FP64_TFLOPS if kprec == "fp64" else PEAK_FP32_TFLOPS
    achieved = flops / (kms * 1e-3) / 1e12
    total64 = sum(r[4] for r in rows if r[6] == "fp64")
    total32 = sum(r[4] for r in rows if r[6] == "fp32")
    # mixed-mode bound: time at peak for the FP64 and the FP32 flops
    t_peak_ms = (total64 / (PEAK_FP64_TFLOPS * 1e12) + total32 / (PEAK_FP32_TFLOPS * 1e12)) * 1e3

    # nuclear gradient (four-centre term): the same (H2O)32 system and
    # n-icosane C20H42 (510 bf), dE2/dR at a fixed density (with --density
    # converged, the density of the reference's own SCF on each), in FP64 and in
    # mixed precision, the FP32 threshold autotuned over a sweep: the fastest
    # whose max |dgrad| against the FP64 gradient is <= 1e-6 Eh/bohr
    grads = {}
    gthrs = [float(t) for t in args.grad_thresholds.split(",")]
    for gname in (() if args.no_gradient else ("water32", "icosane")):
        if gname == "water32":
            gb, gD, ge = b, D, eng
        else:
            gsyms, gxyz = CONFIGS[gname]()
            gb = build_basis(gsyms, gxyz)
            if args.density == "converged":
                zi = np.load(ROOT / "tests" / "golden" / "icosane_scf.npz")
                gb.xyz = np.asarray(zi["xyz"], np.float64).copy()
                gD = 2.0 * zi["C_occ"] @ zi["C_occ"].T
            else:
                gD = synthetic_density(gb.xyz, gb.katom, gb.ksumtype, gb.nbasis, DENSITY_SEED)
            ge = FockEngine(gb)
            ge.compute_schwarz()
        gDd = torch.from_numpy(np.ascontiguousarray(gD.ravel(order="F"))).to(dev)
        runs = []
        for gprec, gthr in [("fp64", None)] + [("mixed", t) for t in gthrs]:
            gopts = (make_opts(precision=QK_FP64) if gprec == "fp64"
                     else make_opts(precision=QK_MIXED, fp32_threshold=gthr))
            gout = torch.zeros(3 * ge.natom, dtype=torch.float64, device=dev)
            torch.cuda.synchronize()
            tw = time.perf_counter()
            ge.grad_device(gDd.data_ptr(), gout.data_ptr(), gopts, sptr)   # warm-up (builds the work lists)
            torch.cuda.synchronize()
            first_s = time.perf_counter() - tw
            gval = gout.cpu().numpy()
            ge.set_timing(False)
            nrep = max(1, min(args.steps, 5))
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(nrep):
                ge.grad_device(gDd.data_ptr(), gout.data_ptr(), gopts, sptr)
            e1.record(stream)
            torch.cuda.synchronize()
            gs = e0.elapsed_time(e1) / nrep / 1e3
            ge.set_timing(True)                   # one more, with per-class events
            ge.grad_device(gDd.data_ptr(), gout.data_ptr(), gopts, sptr)
            torch.cuda.synchronize()
            st64, st32 = ge.grad_class_stats(), ge.grad_class_stats(fp32=True)
            ge.set_timing(False)
            lab = gprec if gthr is None else f"mixed {gthr:.0e}"
            for sub, gst in (("", st64), (" fp32", st32)):
                for c in sorted(gst, key=lambda c: -c[3]):
                    if c[1]:
                        log(f"  grad {lab}{sub} class {c[0]}: {c[1]:>10d} quartets {c[2]:>11d} prim {c[3]:8.3f} ms")
            q64, q32 = sum(c[1] for c in st64), sum(c[1] for c in st32)
            r = {"seconds": gs, "atoms": int(ge.natom), "nbasis": int(gb.nbasis), "quartets": int(q64 + q32),
                 "prim_quartets": int(sum(c[2] for c in st64) + sum(c[2] for c in st32)), "precision": gprec,
                 # wall time of the first call: the host builds the gradient's work
                 # lists (once per geometry and option set), then one gradient
                 "first_call_s": first_s}
            if gprec == "mixed":
                r.update({"fp32_threshold": gthr, "fp32_quartet_share": q32 / max(q64 + q32, 1),
                          "max_abs_dgrad_vs_fp64": float(np.abs(gval - runs[0]["_g"]).max())})
            r["_g"] = gval
            runs.append(r)
            log(f"gradient {gname} {lab}: first call {first_s * 1e3:.1f} ms (work lists + one gradient)")
            log(f"gradient {gname} {lab}: {gs * 1e3:.2f} ms ({q64 + q32} quartets"
                + (f", fp32 share {r['fp32_quartet_share']:.3f}, max|dgrad| {r['max_abs_dgrad_vs_fp64']:.2e})"
                   if gprec == "mixed" else ")"))
        for r in runs:
            r.pop("_g")
        f64, mixed = runs[0], runs[1:]
        ok = [r for r in mixed if r["max_abs_dgrad_vs_fp64"] <= 1e-6]
        gpick = min(ok, key=lambda r: r["seconds"]) if ok else None
        grads[gname] = {"seconds": f64["seconds"], "atoms": f64["atoms"], "nbasis": f64["nbasis"],
                        "quartets": f64["quartets"], "prim_quartets": f64["prim_quartets"], "precision": "fp64",
                        "mixed": gpick, "mixed_sweep": mixed,
                        "mixed_vs_fp64_speedup": f64["seconds"] / gpick["seconds"] if gpick else None}

    # the gpu_get2e_-style boundary hands over host arrays: the same build with
    # the density uploaded and the operator downloaded each time (never `value`)
    hopts = make_opts(precision=QK_MIXED, fp32_threshold=thr) if precision == "mixed" else make_opts()
    eng.fock(D, hopts)
    th = time.perf_counter()
    for _ in range(3):
        eng.fock(D, hopts)
    pcie_ms = (time.perf_counter() - th) / 3 * 1e3
    log(f"host-array builds (D up, G down): {pcie_ms:.2f} ms/build")

    scf = None if args.no_scf else scf_leg(b, eng, syms, torch, dev, stream, sptr)
    b3lyp = None if args.no_water64 else water64_leg(args, torch, dev, stream, sptr)

    # HBM bytes per launch of the roofline kernel: FETCH_SIZE + WRITE_SIZE from
    # rocprofv3 --pmc passes over this workload, scaled by a plain copy's
    # readings (profiles/pmc_traffic_water32.json); null when the file has no
    # figure for this kernel at this precision and threshold
    traffic = None
    tfile = ROOT / "profiles" / "pmc_traffic_water32.json"
    if tfile.exists():
        tj = json.loads(tfile.read_text())
        if precision == "mixed" and thr == tj.get("fp32_threshold"):
            traffic = tj["mixed_build"].get(kname, {}).get("bytes")
        elif precision == "fp64":
            traffic = tj["fp64_build"].get(kname, {}).get("bytes")

    base = None if args.no_cpu_baseline else cpu_baseline(syms, xyz, D)
    value = 1000.0 / ms
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "Fock builds/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64" if precision == "mixed" else "f64",
        "data": ("seeded (H2O)32 geometry with the reference's converged RHF density "
                 "(tests/golden/water32_scf.npz), n-icosane with that of the reference's SCF on it "
                 "(tests/golden/icosane_scf.npz); (H2O)64: seeded synthetic density"
                 if args.density == "converged" else
                 "synthetic: seeded (H2O)32 geometry, seeded exp(-0.25 R^2) density"),
        "config": {"workload": "(H2O)32 HF/6-31G** two-electron Fock build G(D), 800 bf, tau 1e-11, "
                               + ("mixed FP32/FP64 ERIs (FP32 below Schwarz bound %g), FP64 accumulation" % thr
                                  if precision == "mixed" else "all-FP64 ERIs"),
                   "atoms": len(syms), "nbasis": int(b.nbasis),
                   "precision": precision, "fp32_threshold": thr, "geometry_seed": 2,
                   "density": args.density, "density_seed": DENSITY_SEED if args.density == "synthetic" else None,
                   "quartets": int(st.quartets),
                   "quartets_fp32": int(st.quartets_fp32), "prim_quartets": int(st.prim_quartets),
                   "E2_half_trDG_fp64": E64,
                   "gradient_s": grads["water32"]["seconds"] if grads else None},
        "gradient": grads,
        "b3lyp_water64": b3lyp,
        "scf_water32": scf,
        "fock_plus_gradient_water32": None if not grads.get("water32", {}).get("mixed") else {
            "fp64_ms": ms64 + grads["water32"]["seconds"] * 1e3,
            "mixed_ms": ms + grads["water32"]["mixed"]["seconds"] * 1e3,
            "mixed_vs_fp64_speedup": (ms64 + grads["water32"]["seconds"] * 1e3) /
                                     (ms + grads["water32"]["mixed"]["seconds"] * 1e3)},
        "fp64": {"builds_per_s": 1000.0 / ms64, "ms_per_build": ms64},
        # host arrays in and out (5 MB up, 5 MB down per build), headline precision
        "pcie_inclusive_ms_per_build": pcie_ms,
        "mixed_vs_fp64_speedup": ms64 / ms,
        # per geometry: Schwarz bounds on the device, then the host's work lists
        # per option set (outside the timed builds)
        "setup_s": {"schwarz": setup_s, "prepare": prep_s},
        "bit_identical_builds": repro,
        "mixed_sweep": sweep,
        "eri_quartets_per_s": st.quartets * value,
        "achieved_tflops_total": (total64 + total32) / (ms * 1e-3) / 1e12,
        "vector_roofline_time_frac": t_peak_ms / ms,
        # the three longest class kernels of the headline build (the first is
        # the roofline kernel above; near-equal ones can swap places run to run)
        "roofline_top3": [{"kernel": r[0], "avg_launch_ms": r[5],
                           "achieved_tflops": r[4] / (r[5] * 1e-3) / 1e12,
                           "frac": r[4] / (r[5] * 1e-3) / 1e12 / (PEAK_FP64_TFLOPS if r[6] == "fp64"
                                                                  else PEAK_FP32_TFLOPS)}
                          for r in sorted(crit, key=lambda r: -r[5])[:3]],
        "roofline": {"bound": "vector", "pipe": "VALU " + kprec, "kernel": kname, "achieved": achieved,
                     "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": traffic, "flops_per_launch": flops, "avg_launch_ms": kms},
        "cpu_baseline": base,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
