"""Headline benchmark: assembly seconds + H-matvec/s of the Helmholtz
combined-field operator on the 81,920-triangle icosphere, mixed (complex64
factors and dense leaves, FP32 accumulation) vs plain FP64 (complex128).

One step = one H-matvec y = C x of the assembled operator (inputs resident in
HBM).  Prints ONE JSON line.  Workload: icosphere level 6, SurfaceL2 order 0,
kappa = 10, intorder 4, leafsize 64, eta 2, ACA eps 1e-6.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "assembly s + H-matvec/s, Helmholtz sphere 80k tris; mixed vs FP64 @1e-6 rel err"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def mie_dudn_soft_sphere(kappa, cos_theta, nmax=None):
    """Normal derivative of the total field on a sound-soft unit sphere for the
    incident plane wave e^{i kappa z}: -(i/kappa) sum_n i^n (2n+1) P_n(cos) / h_n(kappa)."""
    from scipy.special import eval_legendre, spherical_jn, spherical_yn

    nmax = nmax or int(kappa + 40)
    s = np.zeros_like(cos_theta, dtype=np.complex128)
    for n in range(nmax + 1):
        h = spherical_jn(n, kappa) + 1j * spherical_yn(n, kappa)
        s += (1j ** n) * (2 * n + 1) * eval_legendre(n, cos_theta) / h
    return -1j / kappa * s


def burton_miller_rhs(mesh, kappa):
    """rhs_i = int_{T_i} i kappa e^{i kappa z} (n_z - 1) ds with the 6-point rule
    (demos/HelmholtzBurtonMiller.ipynb cell 8, incident wave along +z)."""
    a, b = 0.445948490915965, 0.091576213509771
    wa, wb = 0.223381589678011, 0.109951743655322
    bary = np.array([[1 - 2 * a, a, a], [a, 1 - 2 * a, a], [a, a, 1 - 2 * a],
                     [1 - 2 * b, b, b], [b, 1 - 2 * b, b], [b, b, 1 - 2 * b]])
    w = np.array([wa] * 3 + [wb] * 3)
    P = mesh.xyz[mesh.tri]  # nt x 3 x 3
    pts = np.einsum("qk,tkc->tqc", bary, P)
    n = mesh.normals()
    area = mesh.areas()
    f = 1j * kappa * np.exp(1j * kappa * pts[:, :, 2]) * (n[:, 2:3] - 1.0)
    return (f * w[None, :]).sum(1) * area


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--level", type=int, default=6)
    ap.add_argument("--kappa", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the built-in apply work shapes")
    ap.add_argument("--aca-gram", action="store_true", help="also assemble with NGBEM_ACA=gram (opt-in variant)")
    ap.add_argument("--no-aca-gram", action="store_true", help=argparse.SUPPRESS)  # the default now
    ap.add_argument("--no-block-product", action="store_true",
                    help="skip the Maxwell EFIE 64-right-hand-side block product after the headline workload")
    ap.add_argument("--calib-copy", action="store_true",
                    help="end with one elementwise kernel reading and writing 512 MiB and two 512 MiB streaming "
                         "reads with the apply kernels' load shape (calibrate rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--assembly-profile", type=str,
                    default=os.path.join(ROOT, "profiles", "assembly_helmholtz_cf_L6_fp64_counters.md"),
                    help="rocprof summary with the executed FP64 flops of the assembly kernels")
    ap.add_argument("--traffic", type=str,
                    default=os.path.join(ROOT, "profiles", "hmatvec_helmholtz_cf_L6_traffic.md"),
                    help="summary of the PMC passes with 'bytes_per_launch <kernel>:' lines (complex64 apply)")
    ap.add_argument("--workload", choices=["helmholtz_cf", "maxwell_efie_64rhs", "batch_sl_256"],
                    default="helmholtz_cf",
                    help="helmholtz_cf: Helmholtz combined field on the icosphere of level 6 (81,920 triangles), "
                         "assembly + H-matvecs; maxwell_efie_64rhs: block product of 64 right-hand sides; "
                         "batch_sl_256: 256 Helmholtz single layers on ellipsoids, batched apply + gradient")
    ap.add_argument("--instances", type=int, default=256)
    ap.add_argument("--reserve-gib", type=float, default=0.0,
                    help="map this much device memory into the library's pool before timing (NGBEM_ALLOC=pool)")
    ap.add_argument("--assemble-twice", action="store_true",
                    help="also time a second assembly of the same operator (steady state)")
    ap.add_argument("--threads", type=int, default=8, help="host threads assembling instances concurrently")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU sharding is a design "
              f"note, not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.workload == "maxwell_efie_64rhs":
        bench_maxwell_block(args)
        return
    if args.workload == "batch_sl_256":
        bench_batch(args)
        return

    import torch

    import ngbem_amd as nb
    from ngbem_amd import _capi

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    kappa = args.kappa
    mesh = nb.icosphere(args.level)
    fes = nb.SurfaceL2(mesh, order=0)
    params = dict(intorder=4, leafsize=64, eta=2.0, eps=1e-6)
    log(f"bench: icosphere L{args.level} ({mesh.nt} triangles), kappa={kappa}, {params}")

    # warm the runtime (first HIP call initialises the device) and load the
    # assembly kernels' code objects with a small operator of the same kind
    torch.zeros(1, device=dev)
    if args.reserve_gib > 0:
        nb.reserve(args.reserve_gib)
    small = nb.SurfaceL2(nb.icosphere(2))
    nb.HelmholtzCombinedFieldOperator(small, small, kappa=kappa, mixed=True, **dict(params, leafsize=16))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    C = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=kappa, mixed=True, **params)
    torch.cuda.synchronize()
    t_assembly = time.perf_counter() - t0
    st = C.stats()
    log(f"assembly {t_assembly:.3f} s  near {st['t_near']:.3f} s  far {st['t_far']:.3f} s  "
        f"setup {st['t_setup']:.3f} s  blocks near/far {st['n_near']}/{st['n_far']}  max rank {st['max_rank']}")
    t_assembly2 = None
    if args.assemble_twice:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        C2 = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=kappa, mixed=True, **params)
        torch.cuda.synchronize()
        t_assembly2 = time.perf_counter() - t0
        st2 = C2.stats()
        del C2
        log(f"second assembly {t_assembly2:.3f} s  near {st2['t_near']:.3f} s  far {st2['t_far']:.3f} s  "
            f"setup {st2['t_setup']:.3f} s")

    # the same assembly with the single-pass (Gram) cross approximation,
    # NGBEM_ACA=gram: faster far field, stops that may move by a step
    aca_gram = None
    if args.aca_gram and not args.no_aca_gram:
        os.environ["NGBEM_ACA"] = "gram"
        try:
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            C3 = nb.HelmholtzCombinedFieldOperator(fes, fes, kappa=kappa, mixed=True, **params)
            torch.cuda.synchronize()
            tg = time.perf_counter() - t0
            st3 = C3.stats()
            aca_gram = {"assembly_s": tg, "far_s": st3["t_far"], "far_factor_entries": st3["far_factor_entries"],
                        "far_factor_entries_default": st["far_factor_entries"]}
            del C3
        finally:
            del os.environ["NGBEM_ACA"]
        log(f"assembly with NGBEM_ACA=gram {tg:.3f} s  far {st3['t_far']:.3f} s")

    # the apply's autotuner on the assembled operator (both precisions)
    tune = None
    if not args.no_autotune:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tune = C.Autotune(iters=7)
        torch.cuda.synchronize()
        tune["seconds"] = time.perf_counter() - t0
        log("autotune ({:.2f} s): ".format(tune["seconds"]) + "; ".join(
            f"{k}: far_t rows8 up to {r.get('far_t_choice')}, row pieces of {r.get('rows_choice')} columns, apply "
            f"{r['apply_ms_before']:.3f} -> {r['apply_ms_after']:.3f} ms" for k, r in tune.items() if k != "seconds"))

    x = torch.randn(fes.ndof, dtype=torch.complex128, device=dev)
    y = torch.zeros_like(x)
    M64, M32 = C.mat, C.mat.mixed

    def time_matvecs(M, steps, warmup):
        for _ in range(warmup):
            M.Mult(x, y)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            M.Mult(x, y)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3

    t64 = time_matvecs(M64, args.steps, args.warmup)
    t32 = time_matvecs(M32, args.steps, args.warmup)
    # the same applies replayed from hipGraphs (BEM_APPLY_GRAPH)
    t64g = time_matvecs(M64.graph, args.steps, args.warmup)
    t32g = time_matvecs(M32.graph, args.steps, args.warmup)
    y64 = M64 * x
    y32 = M32 * x
    mv_err = (torch.linalg.norm(y64 - y32) / torch.linalg.norm(y64)).item()
    log(f"graph replay: fp64 {t64g / args.steps * 1e3:.3f} ms  mixed {t32g / args.steps * 1e3:.3f} ms")
    log(f"H-matvec fp64 {t64 / args.steps * 1e3:.3f} ms  mixed {t32 / args.steps * 1e3:.3f} ms  "
        f"mixed rel err {mv_err:.2e}")

    # per-kernel times from HIP events on the launch stream
    def profile(flags, iters=max(5, args.steps // 2)):
        ms = np.zeros(6)
        stream = torch.cuda.current_stream().cuda_stream
        _capi.check(_capi.lib().bem_op_profile_apply(C._h, flags, iters, _capi.dptr(ms), ctypes.c_void_p(stream)))
        return ms

    prof32 = profile(_capi.APPLY_MIXED)
    prof64 = profile(0)
    # the transposed apply (HMatrix::MultTrans; the adjoint costs the same)
    tr32 = profile(_capi.APPLY_MIXED | _capi.APPLY_TRANS)
    tr64 = profile(_capi.APPLY_TRANS)
    log(f"transposed apply: mixed {tr32[5]:.3f} ms  fp64 {tr64[5]:.3f} ms")
    n = fes.ndof
    near = st["near_entries"]
    far_u = far_v = 0
    R_tot = 0
    m_tot = 0
    for b in range(C.num_blocks()):
        bi = C.block_info(b)
        if bi["rank"] >= 0:
            far_u += bi["m"] * bi["rank"]
            far_v += bi["rank"] * bi["n"]
            R_tot += bi["rank"]
            m_tot += bi["m"]

    # algorithmic bytes per launch (each operand once) of the apply kernels; the
    # forward plan reads U inside the row tiles unless NGBEM_MV_PLAN=z (then
    # U t per block in k_mv_far_z, z slices in the tiles)
    zplan = os.environ.get("NGBEM_MV_PLAN", "cols") == "z"

    def kernel_bytes(s):
        kb = {"k_mv_far_t": far_v * s + n * s + R_tot * s}            # V, x, t
        if zplan:
            kb["k_mv_far_z"] = far_u * s + R_tot * s + m_tot * s      # Ut, t, z
            kb["k_mv_rows"] = near * s + n * s + m_tot * s + n * s    # dense, x, z, y
        else:
            kb["k_mv_rows"] = (near + far_u) * s + n * s + R_tot * s + n * s  # dense, U, x, t, y
        return kb

    def apply_bytes(s):
        z = 2 * m_tot * s if zplan else 0
        return (near + far_u + far_v) * s + 2 * n * s + 2 * R_tot * s + z

    kms = {"mixed": dict(zip(["gather", "k_mv_far_t", "k_mv_far_z", "k_mv_rows", "scatter", "apply"], prof32)),
           "fp64": dict(zip(["gather", "k_mv_far_t", "k_mv_far_z", "k_mv_rows", "scatter", "apply"], prof64))}
    kb32, kb64 = kernel_bytes(8), kernel_bytes(16)
    gbs = {"mixed": {k: kb32[k] / (kms["mixed"][k] * 1e-3) / 1e9 for k in kb32},
           "fp64": {k: kb64[k] / (kms["fp64"][k] * 1e-3) / 1e9 for k in kb64}}
    dom = max(kb32, key=lambda k: kms["mixed"][k])  # dominant kernel of the measured (mixed) apply
    # the row tiles run as k_mv_rows_flat (one column list per tile) unless
    # the z plan or NGBEM_MV_ROWS=segments selects the segment loop
    flat = not zplan and os.environ.get("NGBEM_MV_ROWS", "flat") != "segments"
    dom_name = "k_mv_rows_flat" if dom == "k_mv_rows" and flat else dom
    log("mixed: " + "  ".join(f"{k} {kms['mixed'][k]:.3f} ms {gbs['mixed'][k]:.0f} GB/s" for k in kb32))
    log("fp64:  " + "  ".join(f"{k} {kms['fp64'][k]:.3f} ms {gbs['fp64'][k]:.0f} GB/s" for k in kb64))

    out = {
        "metric": METRIC,
        "value": args.steps / t32,
        "unit": "H-matvec/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"Helmholtz combined field K - i kappa V, icosphere L{args.level} "
                               f"({mesh.nt} triangles), SurfaceL2 p0, kappa={kappa}, intorder 4, leafsize 64, "
                               f"eta 2, ACA eps 1e-6; mixed = complex64 dense leaves + factors, FP32 accumulation"},
        "assembly_s": t_assembly,
        "assembly_breakdown_s": {"setup": st["t_setup"], "near": st["t_near"], "far_aca": st["t_far"]},
        "second_assembly_s": t_assembly2,
        "assembly_aca_gram": aca_gram,
        "autotune": tune,
        "graph_replay_ms_per_matvec": {"mixed": t32g / args.steps * 1e3, "fp64": t64g / args.steps * 1e3},
        "fp64_matvec_per_s": args.steps / t64,
        "mixed_vs_fp64_matvec_speedup": t64 / t32,
        "mixed_matvec_rel_err": mv_err,
        "hmatrix": {"near_entries": near, "far_factor_entries": st["far_factor_entries"], "n_near": st["n_near"],
                    "n_far": st["n_far"], "max_rank": st["max_rank"], "nze": st["nze"],
                    "near_kernel_evals": st["near_kernel_evals"], "far_kernel_evals": st["far_kernel_evals"]},
        "near_field_evals_per_s": st["near_kernel_evals"] / st["t_near"] if st["t_near"] > 0 else None,
        "kernel_ms": kms,
        "transposed_apply_ms": {"mixed": float(tr32[5]), "fp64": float(tr64[5]),
                                "stages_mixed": dict(zip(["gather", "k_mv_far_t", "k_mv_far_z", "cols", "scatter",
                                                          "apply"], map(float, tr32)))},
        "kernel_gbs": gbs,
        "roofline": {"kernel": f"{dom_name}<cx<float>,cx<float>>", "bound": "hbm", "achieved": gbs["mixed"][dom],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs["mixed"][dom] / HBM_PEAK_GBS, "traffic": None,
                     "algorithmic_bytes_per_launch": kb32[dom],
                     "fp64_achieved": gbs["fp64"][dom], "fp64_frac": gbs["fp64"][dom] / HBM_PEAK_GBS,
                     "apply_bytes_mixed": apply_bytes(8), "apply_bytes_fp64": apply_bytes(16),
                     "apply_gbs_mixed": apply_bytes(8) / (kms["mixed"]["apply"] * 1e-3) / 1e9,
                     "apply_gbs_fp64": apply_bytes(16) / (kms["fp64"]["apply"] * 1e-3) / 1e9,
                     "apply_frac_mixed": apply_bytes(8) / (kms["mixed"]["apply"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
    }

    # what a plain streaming kernel reaches on this GPU: torch's copy of 1 GiB
    # (bytes read + written / time), for the frac's context
    probe = hbm_probe(dev)
    out["roofline"]["hbm_probe"] = probe
    log(f"HBM probe: torch copy {probe['copy_gbs']:.0f} GB/s, streaming read {probe['read_gbs']:.0f} GB/s")

    if not args.no_solve:
        # Burton-Miller combined-field solve: (1/2 M + C) u = f; FP64 GMRES vs
        # mixed inner GMRES with FP64 iterative refinement
        Mass = nb.MassMatrix(fes)
        f = torch.from_numpy(burton_miller_rhs(mesh, kappa)).to(dev)
        area = torch.from_numpy(mesh.areas()).to(dev)
        # the inverse of the diagonal mass matrix, as the reference's demo
        # (demos/HelmholtzBurtonMiller.ipynb: BilinearForm(u*v*ds,
        # diagonal=True) ... Inverse()); GMRES applies it inside its Arnoldi step
        pre = nb.MassMatrix(fes).Inverse()
        A64 = 0.5 * Mass + M64
        A32 = 0.5 * Mass + M32
        nb.GMRes(A64, f, pre=pre, tol=1e-4, maxsteps=5, printrates=False)  # warm up torch's kernels
        nb.RefinedSolve(A64, A32, f, pre=pre, tol=1e-3, inner_tol=1e-2, maxouter=1)
        # the discrete solution to 1e-12 (untimed): what both forms are measured against
        u_ref = nb.GMRes(A64, f, pre=pre, tol=1e-12, maxsteps=400, printrates=False)

        def timed(fn, reps=5):
            ts, res = [], None
            for _ in range(reps):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                res = fn()
                torch.cuda.synchronize()
                ts.append(time.perf_counter() - t0)
            return float(np.median(ts)), res

        def rel(u):
            return (torch.linalg.norm(u - u_ref) / torch.linalg.norm(u_ref)).item()

        # equal error: both forms to GMRES tolerance 1e-7 (solution error ~1e-7 <= 1e-6)
        t_s64, u64 = timed(lambda: nb.GMRes(A64, f, pre=pre, tol=1e-7, maxsteps=400, printrates=False))
        it64 = nb.GMRes.iterations
        t_s32, u32 = timed(lambda: nb.GMRes(A32, f, pre=pre, tol=1e-7, maxsteps=400, printrates=False))
        it32 = nb.GMRes.iterations
        # the FP64 residual of the mixed solution (one FP64 apply), the check
        # the refinement would start from
        t_res, rres = timed(lambda: torch.linalg.norm(f - nb.krylov._apply(A64, u32)) / torch.linalg.norm(f))
        # to 1e-10: FP64 GMRES against mixed inner GMRES + FP64 refinement
        t_r64, u64b = timed(lambda: nb.GMRes(A64, f, pre=pre, tol=1e-10, maxsteps=400, printrates=False))
        it64b = nb.GMRes.iterations
        t_r32, (u32b, hist) = timed(lambda: nb.RefinedSolve(A64, A32, f, pre=pre, tol=1e-10, inner_tol=1e-5))
        err64, err32, err64b, err32b = rel(u64), rel(u32), rel(u64b), rel(u32b)
        cen = mesh.xyz[mesh.tri].mean(1)
        cos_t = cen[:, 2] / np.linalg.norm(cen, axis=1)
        mie = mie_dudn_soft_sphere(kappa, cos_t)
        u_np = u_ref.cpu().numpy()
        w = mesh.areas()
        mie_err = math.sqrt(np.sum(w * np.abs(u_np - mie) ** 2) / np.sum(w * np.abs(mie) ** 2))
        log(f"solve to 1e-7: fp64 {t_s64 * 1e3:.2f} ms ({it64} its, err {err64:.1e})  mixed {t_s32 * 1e3:.2f} ms "
            f"({it32} its, err {err32:.1e})  speedup {t_s64 / t_s32:.2f}  (+ FP64 residual {t_res * 1e3:.2f} ms: "
            f"{float(rres):.1e})")
        log(f"solve to 1e-10: fp64 {t_r64 * 1e3:.2f} ms ({it64b} its)  mixed+refinement {t_r32 * 1e3:.2f} ms "
            f"({nb.RefinedSolve.inner_iterations} inner its, {len(hist)} outer)  errors {err64b:.1e} / {err32b:.1e}  "
            f"L2 err vs Mie {mie_err:.3e}")
        out["solve"] = {
            "tol_1e-7": {"fp64_s": t_s64, "fp64_iterations": it64, "fp64_rel_err": err64, "mixed_s": t_s32,
                         "mixed_iterations": it32, "mixed_rel_err": err32, "solve_speedup": t_s64 / t_s32,
                         "fp64_residual_check_s": t_res, "mixed_fp64_rel_residual": float(rres),
                         "solve_speedup_with_fp64_residual": t_s64 / (t_s32 + t_res)},
            "tol_1e-10": {"fp64_s": t_r64, "fp64_iterations": it64b, "fp64_rel_err": err64b,
                          "mixed_refined_s": t_r32, "mixed_inner_iterations": nb.RefinedSolve.inner_iterations,
                          "refinement_steps": len(hist), "mixed_rel_err": err32b, "speedup": t_r64 / t_r32},
            "errors_against": "FP64 GMRES solution to 1e-12",
            "rel_l2_err_vs_mie_fp64": mie_err,
            "assemble_plus_solve_s_fp64": t_assembly + t_s64,
            "assemble_plus_solve_s_mixed": t_assembly + t_s32}

    if not args.no_cpu_baseline:
        solved = not args.no_solve
        cb = cpu_baseline(C, mesh, kappa, params, st, f if solved else None, area if solved else None,
                          u64 if solved else None)
        if solved:
            cb["gpu_speedup_assemble_plus_solve_fp64"] = cb["assemble_plus_solve_s"] / (t_assembly + t_s64)
            cb["gpu_speedup_assemble_plus_solve_mixed"] = cb["assemble_plus_solve_s"] / (t_assembly + t_s32)
        out["cpu_baseline"] = cb

    out["roofline_assembly"] = assembly_roofline(C, st, args, kappa)

    if args.traffic and os.path.exists(args.traffic) and args.level == 6 and kappa == 10.0:
        for line in open(args.traffic):  # "bytes_per_launch <kernel>: <bytes>"
            key, _, val = line.partition(":")
            if key.strip() == f"bytes_per_launch {dom_name}":
                out["roofline"]["traffic"] = float(val)
                out["roofline"]["traffic_source"] = os.path.relpath(args.traffic, ROOT)
    if args.calib_copy:
        # counter calibration: one elementwise kernel that reads 512 MiB and
        # writes 512 MiB with ordinary vector loads and stores
        a = torch.ones(128 << 20, dtype=torch.float32, device=dev)
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        torch.mul(a, 2.0, out=b)
        torch.cuda.synchronize()
        # and the library's streaming read with the apply kernels' load shape
        # (one 8- or 16-byte element per lane, 64 consecutive per wave load):
        # 512 MiB each, 4096 workgroups (k_probe_read<cx<float>> / <cx<double>>)
        part = torch.empty(4096, dtype=torch.complex128, device=dev)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        for esz in (8, 16):
            _capi.check(_capi.lib().bem_probe_read(ctypes.c_void_p(a.data_ptr()), (512 << 20) // esz, esz,
                                                    ctypes.c_void_p(part.data_ptr()), 4096, stream), "bem_probe_read")
        torch.cuda.synchronize()
    if not args.no_block_product:
        # the many-right-hand-side configuration (Maxwell EFIE, 64 plane-wave
        # loads per block product on the matrix cores), after the headline
        # operator's memory is released
        C = M64 = M32 = A64 = A32 = None  # noqa: F841  (drop every reference to the headline operator)
        import gc

        gc.collect()
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        bp = bench_maxwell_block(argparse.Namespace(level=args.level, steps=10, warmup=2), emit=False)
        out["block_product_maxwell_efie_64rhs"] = {
            "config": bp["config"]["workload"], "ms_fp32": bp["ms_per_step"],
            "ms_fp64": 1e3 / bp["fp64_block_per_s"], "tflops_fp32": bp["roofline"]["achieved"],
            "frac_fp32_mfma": bp["roofline"]["frac"], "tflops_fp64": bp["roofline"]["fp64_achieved"],
            "frac_fp64_mfma": bp["roofline"]["fp64_achieved"] / bp["roofline"]["fp64_peak"],
            "block_vs_vector_speedup": bp["block_vs_vector_speedup"],
            "fp32_vs_fp64_rel_err": bp["fp32_vs_fp64_rel_err"], "assembly_s": bp["assembly_s"]}
    print(json.dumps(out), flush=True)


def hbm_probe(dev, reps=10):
    """Streaming bandwidth of torch's copy kernel: 1 GiB read and 1 GiB
    written per call, (bytes read + written) / time over `reps` calls."""
    import torch
    a = torch.ones(256 << 20, dtype=torch.float32, device=dev)
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    gbs = 2 * a.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    # a read-only stream with the apply kernels' load shape (bem_probe_read:
    # one 8-byte element per lane, 4096 workgroups looping over 1 GiB)
    from ngbem_amd import _capi
    part = torch.empty(4096, dtype=torch.complex128, device=dev)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def probe():
        _capi.check(_capi.lib().bem_probe_read(ctypes.c_void_p(a.data_ptr()), a.numel() // 2, 8,
                                                ctypes.c_void_p(part.data_ptr()), 4096, stream), "bem_probe_read")

    probe()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        probe()
    e1.record()
    torch.cuda.synchronize()
    rgbs = a.numel() * 4 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b, part
    torch.cuda.empty_cache()
    return {"copy_gbs": gbs, "read_gbs": rgbs,
            "how": "torch copy_ of 1 GiB float32 (read + written bytes), and bem_probe_read of 1 GiB "
                   "(8-byte loads, 4096 workgroups); 10 calls each, HIP events"}


def bench_maxwell_block(args, emit=True):
    """Maxwell EFIE single layer, HDivSurface p0, kappa 5, intorder 4,
    leafsize 40, eta 2, ACA eps 1e-5; one step = one block product of 64
    right-hand sides (MFMA path), FP64 and FP32 copies, against 64 vector
    products."""
    import torch

    import ngbem_amd as nb

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    level = args.level
    kappa = 5.0
    nrhs = 64
    mesh = nb.icosphere(level)
    hd = nb.HDivSurface(mesh)
    params = dict(intorder=4, leafsize=40, eta=2.0, eps=1e-5)
    log(f"bench maxwell: icosphere L{level}, {hd.ndof} RWG dofs, kappa {kappa}, {params}, {nrhs} rhs")
    torch.zeros(1, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    V = nb.MaxwellSingleLayerPotentialOperator(hd, kappa, mixed=True, **params)
    torch.cuda.synchronize()
    t_assembly = time.perf_counter() - t0
    st = V.stats()
    log(f"assembly {t_assembly:.3f} s  near {st['t_near']:.3f} s  far {st['t_far']:.3f} s  "
        f"blocks near/far {st['n_near']}/{st['n_far']}  max rank {st['max_rank']}")
    g = torch.Generator(device="cpu").manual_seed(0)
    X = torch.randn(nrhs, hd.ndof, dtype=torch.complex128, generator=g).to(dev)
    Y = torch.zeros_like(X)

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    t64 = timed(lambda: V.mat.Mult(X, Y), args.steps, args.warmup)
    t32 = timed(lambda: V.mat.mixed.Mult(X, Y), args.steps, args.warmup)
    y = torch.zeros(hd.ndof, dtype=torch.complex128, device=dev)

    def per_vector(M):
        for r in range(nrhs):
            M.Mult(X[r], y)

    tv64 = timed(lambda: per_vector(V.mat), max(2, args.steps // 10), 1)
    tv32 = timed(lambda: per_vector(V.mat.mixed), max(2, args.steps // 10), 1)
    Y64 = V.mat * X
    Y32 = V.mat.mixed * X
    err = (torch.linalg.norm(Y64 - Y32) / torch.linalg.norm(Y64)).item()
    entries = st["near_entries"] + st["far_factor_entries"]
    flops = 8.0 * nrhs * entries  # complex multiply-add = 8 real flops
    log(f"block apply fp64 {t64 * 1e3:.3f} ms ({flops / t64 / 1e12:.1f} TFLOP/s)  "
        f"fp32 {t32 * 1e3:.3f} ms ({flops / t32 / 1e12:.1f} TFLOP/s); 64 vector applies fp64 {tv64 * 1e3:.3f} ms "
        f"fp32 {tv32 * 1e3:.3f} ms; fp32 vs fp64 rel err {err:.2e}")
    out = {
        "metric": "block H-matvec/s (64 right-hand sides), Maxwell EFIE single layer",
        "value": 1.0 / t32, "unit": "block H-matvec/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t32 * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"Maxwell single layer, icosphere L{level} ({mesh.nt} triangles, {hd.ndof} "
                               f"HDivSurface p0 dofs), kappa 5, intorder 4, leafsize 40, eta 2, ACA eps 1e-5, "
                               f"{nrhs} right-hand sides per product"},
        "assembly_s": t_assembly,
        "fp64_block_per_s": 1.0 / t64,
        "per_vector_ms": {"fp64": tv64 * 1e3, "fp32": tv32 * 1e3},
        "block_vs_vector_speedup": {"fp64": tv64 / t64, "fp32": tv32 / t32},
        "fp32_vs_fp64_rel_err": err,
        "roofline": {"bound": "mfma", "achieved": flops / t32 / 1e12, "peak": 157.3, "unit": "TFLOP/s",
                     "frac": flops / t32 / 1e12 / 157.3, "traffic": None,
                     "flops_per_product": flops, "fp64_achieved": flops / t64 / 1e12, "fp64_peak": 78.6,
                     "note": "whole block product (gather, V X, row tiles, scatter) timed"},
        "hmatrix": {"near_entries": st["near_entries"], "far_factor_entries": st["far_factor_entries"],
                    "n_near": st["n_near"], "n_far": st["n_far"], "max_rank": st["max_rank"]},
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def bench_batch(args):
    """Helmholtz single layers on ellipsoids (semi-axes U[0.5, 2], level
    U{3, 4, 5}, kappa U[1, 8], seed 0), SurfaceL2 p0, intorder 4, leafsize 40,
    eta 2, ACA eps 1e-6.  All geometries in one operator (one assembly, each
    geometry with its own wavenumber); one step = one apply of the whole batch
    replayed from a HIP graph; also one gradient step of ||V(kappa) x - t||^2
    w.r.t. x and kappa (adjoint apply + d/dkappa operator)."""
    import torch

    import ngbem_amd as nb

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    rng = np.random.default_rng(0)
    n_inst = args.instances
    axes = rng.uniform(0.5, 2.0, (n_inst, 3))
    levels = rng.integers(3, 6, n_inst)
    kappas = rng.uniform(1.0, 8.0, n_inst)
    meshes = [nb.ellipsoid(int(l), tuple(a)) for l, a in zip(levels, axes)]
    spaces = [nb.SurfaceL2(m) for m in meshes]
    ntri = sum(m.nt for m in meshes)
    log(f"bench batch: {n_inst} instances, levels {np.bincount(levels, minlength=6)[3:].tolist()} (L3/L4/L5), "
        f"{ntri} triangles in all")
    torch.zeros(1, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    slb = nb.SingleLayerBatch(spaces, kappas, threads=args.threads, intorder=4, leafsize=40, eta=2.0, eps=1e-6)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    log(f"assembly of the {n_inst}-geometry single layer and its d/dkappa operator: {t_asm:.2f} s")
    n = slb.V.width
    g = torch.Generator(device="cpu").manual_seed(1)
    x = torch.randn(n, dtype=torch.complex128, generator=g).to(dev)
    y = torch.zeros_like(x)
    B = slb.V

    def timed(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / 1e3 / steps

    t_eager = timed(lambda: B.Mult(x, y), args.steps, args.warmup)
    # the same apply captured once in a HIP graph (torch.cuda.CUDAGraph) and replayed
    y_ref = (B * x).clone()
    yg = torch.zeros_like(x)
    s_cap = torch.cuda.Stream()
    s_cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s_cap):
        B.Mult(x, yg)  # warm the capture stream
    torch.cuda.current_stream().wait_stream(s_cap)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        B.Mult(x, yg)
    t_apply = timed(lambda: graph.replay(), args.steps, args.warmup)
    graph_err = (torch.linalg.norm(yg - y_ref) / torch.linalg.norm(y_ref)).item()
    xr = torch.randn(n, dtype=torch.float64, generator=g).to(dev)
    tgt = torch.randn(n, dtype=torch.complex128, generator=g).to(dev)
    kap = torch.tensor(kappas, dtype=torch.float64, device=dev)

    def grad_step():
        xx = xr.clone().requires_grad_(True)
        kk = kap.clone().requires_grad_(True)
        loss = ((slb(xx, kk) - tgt).abs() ** 2).sum()
        loss.backward()
        return xx.grad, kk.grad

    t_grad = timed(grad_step, max(3, args.steps // 5), 1)
    log(f"batched apply {t_apply * 1e3:.3f} ms (graph replay, rel. diff to direct {graph_err:.1e}), "
        f"{t_eager * 1e3:.3f} ms (direct launches); forward+backward {t_grad * 1e3:.3f} ms")
    out = {
        "metric": "batched H-matvec/s over 256 Helmholtz single layers (one launch sequence)",
        "value": 1.0 / t_apply, "unit": "batched H-matvec/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_apply * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"{n_inst} Helmholtz single layers on ellipsoids, semi-axes U[0.5,2], level U{{3,4,5}}, "
                               f"kappa U[1,8] (seed 0), SurfaceL2 p0, intorder 4, leafsize 40, eta 2, eps 1e-6; "
                               f"{ntri} triangles, {n} dofs in all"},
        "assembly_s": t_asm, "apply_direct_launch_ms": t_eager * 1e3, "grad_step_ms": t_grad * 1e3,
        "graph_replay_rel_diff": graph_err,
    }
    print(json.dumps(out), flush=True)


def RefinedSolve_iters(nb):
    return getattr(nb.RefinedSolve, "inner_iterations", -1)


FP64_PEAK_TFLOPS = 78.6


def assembly_roofline(C, st, args, kappa):
    """Assembly kernels against their bounds: executed FP64 flops (from the
    committed rocprof counter summary of this workload) over the kernels'
    device seconds measured now with HIP events (bem_stats), and for the cross
    approximation its HBM bytes (measured traffic and the algorithmic
    workspace reads) over its seconds.  Algorithmic ACA bytes: step k of a
    block reads the k earlier rows of V and columns of U once (the residual
    row and column) and writes one of each: 16 (m' + n') (r (r - 1) / 2 + r)
    bytes per block of rank r on m' x n' quadrature points."""
    ntq = 6  # points of the intorder-4 triangle rule
    acab = 0.0
    for b in range(C.num_blocks()):
        bi = C.block_info(b)
        r = bi["rank"]
        if r > 0:
            acab += 16.0 * ntq * (bi["m"] + bi["n"]) * (r * (r - 1) / 2 + r)
    t_nr, t_ns, t_aca = st["t_kernel_near_regular"], st["t_kernel_near_singular"], st["t_kernel_aca"]
    out = {"kernel_seconds": {"k_near_regular": t_nr, "k_near_singular": t_ns, "k_aca": t_aca},
           "near_evals_per_s": st["near_kernel_evals"] / (t_nr + t_ns) if t_nr + t_ns > 0 else None,
           "aca_algorithmic_bytes": acab,
           "aca_algorithmic_gbs": acab / t_aca / 1e9 if t_aca > 0 else None}
    flops, hbm = {}, {}
    if args.assembly_profile and os.path.exists(args.assembly_profile) and args.level == 6 and kappa == 10.0:
        for line in open(args.assembly_profile):  # "fp64_flops <kernel>: <flops>" / "hbm_bytes <kernel>: <bytes>"
            key, _, val = line.partition(":")
            if key.startswith("fp64_flops "):
                flops[key.split()[1]] = float(val)
            elif key.startswith("hbm_bytes "):
                hbm[key.split()[1]] = float(val)
        out["source"] = os.path.relpath(args.assembly_profile, ROOT)
    for k, t in (("k_near_regular", t_nr), ("k_near_singular", t_ns), ("k_aca", t_aca)):
        if k in flops and t > 0:
            a = flops[k] / t / 1e12
            out[k] = {"bound": "fp64" if k != "k_aca" else "hbm", "fp64_flops": flops[k], "fp64_achieved": a,
                      "fp64_peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "fp64_frac": a / FP64_PEAK_TFLOPS}
    if "k_aca" in out and "k_aca" in hbm and t_aca > 0:
        out["k_aca"].update(traffic=hbm["k_aca"], achieved_gbs=hbm["k_aca"] / t_aca / 1e9,
                            peak_gbs=HBM_PEAK_GBS, frac=hbm["k_aca"] / t_aca / 1e9 / HBM_PEAK_GBS)
    log("assembly kernels: " + ", ".join(f"{k} {t * 1e3:.1f} ms" for k, t in out["kernel_seconds"].items())
        + "".join(f"; {k} {out[k]['fp64_achieved']:.2f} TFLOP/s FP64 ({out[k]['fp64_frac']:.3f})"
                  for k in ("k_near_regular", "k_near_singular", "k_aca") if k in out)
        + (f"; k_aca {out['k_aca']['achieved_gbs']:.0f} GB/s HBM" if "achieved_gbs" in out.get("k_aca", {}) else ""))
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def numpy_gmres(apply, b, tol, maxsteps):
    """GMRES (classical Gram-Schmidt twice, no restart) on host vectors for the
    CPU baseline solve; returns (x, iterations)."""
    n = b.shape[0]
    beta = np.linalg.norm(b)
    V = np.zeros((maxsteps + 1, n), dtype=np.complex128)
    H = np.zeros((maxsteps + 1, maxsteps), dtype=np.complex128)
    V[0] = b / beta
    g = np.zeros(maxsteps + 1, dtype=np.complex128)
    g[0] = beta
    cs = np.zeros(maxsteps, dtype=np.complex128)
    sn = np.zeros(maxsteps, dtype=np.complex128)
    k = 0
    for k in range(maxsteps):
        w = apply(V[k])
        for _ in range(2):
            h = V[: k + 1].conj() @ w
            w = w - h @ V[: k + 1]
            H[: k + 1, k] += h
        H[k + 1, k] = np.linalg.norm(w)
        V[k + 1] = w / H[k + 1, k]
        for i in range(k):
            hi, hj = H[i, k], H[i + 1, k]
            H[i, k] = np.conj(cs[i]) * hi + np.conj(sn[i]) * hj
            H[i + 1, k] = -sn[i] * hi + cs[i] * hj
        t = np.hypot(abs(H[k, k]), abs(H[k + 1, k]))
        cs[k], sn[k] = H[k, k] / t, H[k + 1, k] / t
        H[k, k], H[k + 1, k] = t, 0.0
        g[k + 1] = -sn[k] * g[k]
        g[k] = np.conj(cs[k]) * g[k]
        if abs(g[k + 1]) <= tol * beta:
            break
    k += 1
    y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
    return y @ V[:k], k


def cpu_baseline(C, mesh, kappa, params, st, f=None, area=None, u_gpu=None):
    """The oracle (C++ FP64 restatement of the reference, OpenMP on the host
    threads of one GPU's share): the whole assembly of the same operator
    (cluster trees, partition, every block), its blocks compared with the
    GPU's, the H-matvec on the identical H-matrix (the GPU's blocks copied
    in; median of 20) and a GMRES solve of the Burton-Miller system with that
    apply (tolerance 1e-7, as the GPU solves)."""
    from oracle import oracle as O

    threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or min(16, os.cpu_count() or 1)
    model = cpu_model()
    log(f"cpu baseline: {model}, {threads} threads; assembling the whole operator on the oracle")
    t0 = time.perf_counter()
    ref = O.OracleOp(mesh, O.HELMHOLTZ_CF, kappa=kappa, trial=O.P0, test=O.P0, **params)
    t_tree = time.perf_counter() - t0
    t_blocks = ref.assemble(nthreads=threads)
    t_asm = t_tree + t_blocks
    log(f"cpu baseline: assembly {t_asm:.2f} s (trees + partition {t_tree:.2f} s, blocks {t_blocks:.2f} s)")
    # the GPU's blocks against the oracle's: every dense block, a seeded 2% of the low-rank ones
    nb_ = C.num_blocks()
    infos = ref.blocks()
    far_ids = [i for i, b in enumerate(infos) if not b["near"] and not b["fallback"]]
    rng = np.random.default_rng(0)
    far_sample = set(rng.choice(far_ids, size=max(1, len(far_ids) // 50), replace=False).tolist()) if far_ids else set()
    near_err, far_err, same_rank, nfs = 0.0, 0.0, 0, 0
    eps = params["eps"]
    for b in range(nb_):
        d = C.block_data(b)
        if b in far_sample or not isinstance(d, tuple):
            o = ref.get_block(b)
            if isinstance(d, tuple):
                A, A_ref = d[0] @ d[1], o[0] @ o[1]
                far_err = max(far_err, float(np.linalg.norm(A - A_ref) / np.linalg.norm(A_ref)))
                same_rank += d[0].shape[1] == o[0].shape[1]
                nfs += 1
            else:
                near_err = max(near_err, float(np.abs(d - o).max() / np.abs(o).max()))
        if isinstance(d, tuple):
            ref.set_block(b, np.concatenate([d[0].ravel(), d[1].ravel()]), d[0].shape[1])
        else:
            ref.set_block(b, d.ravel(), -1)
    log(f"cpu baseline: GPU blocks vs oracle: dense max rel err {near_err:.1e}, low-rank ({nfs} sampled) max rel "
        f"err {far_err:.1e} (10 eps = {10 * eps:.0e}), same rank {same_rank}/{nfs}")
    x = np.random.default_rng(0).standard_normal(mesh.nt) + 0j
    ref.apply(x, nthreads=threads)
    ts = []
    for _ in range(20):
        ref.apply(x, nthreads=threads)
        ts.append(ref.last_apply_seconds)
    mv_s = float(np.median(ts))
    out = {"value": 1.0 / mv_s, "unit": "H-matvec/s", "cores": threads, "kind": "port", "cpu_model": model,
           "sample": f"oracle FP64 (OpenMP, {threads} threads): whole assembly of the same operator, H-matvec on the "
                     f"identical {nb_}-block H-matrix (median of 20), GMRES solve to 1e-7 with that apply",
           "assembly_s": t_asm, "assembly_trees_s": t_tree, "assembly_blocks_s": t_blocks,
           "hmatvec_s": mv_s,
           "gpu_vs_oracle_blocks": {"dense_max_rel_err": near_err, "lowrank_sampled": nfs,
                                    "lowrank_max_rel_err": far_err, "lowrank_same_rank": same_rank}}
    if f is not None:
        fa = f.cpu().numpy()
        ar = area.cpu().numpy()

        def pa(v):
            return 0.5 * v + ref.apply(v, nthreads=threads) / ar  # diag(1/area) (1/2 M + C) v

        t0 = time.perf_counter()
        u, its = numpy_gmres(pa, fa / ar, 1e-7, 200)
        t_solve = time.perf_counter() - t0
        out.update(solve_s=t_solve, solve_iterations=its, assemble_plus_solve_s=t_asm + t_solve)
        if u_gpu is not None:
            ug = u_gpu.cpu().numpy()
            out["solve_rel_diff_vs_gpu_fp64"] = float(np.linalg.norm(u - ug) / np.linalg.norm(ug))
        log(f"cpu baseline: H-matvec {mv_s * 1e3:.1f} ms; GMRES {t_solve:.2f} s ({its} its); assembly + solve "
            f"{t_asm + t_solve:.1f} s")
    return out


if __name__ == "__main__":
    main()
