#!/usr/bin/env python3
"""Benchmark: SP2 density-matrix solves on one MI355X.

Workload: prg_sp2_alg2 on the dense FP64 model tight-binding Hamiltonian,
N = 8192 orbitals, occupation N/2, idempotency tolerance 1e-10, stopping on
the limdiff rule ("Abs"), seed 0.  One step = one full solve (Gershgorin
bounds, normalisation, purification loop, final 2X) with H already resident
in HBM.  The FP64 MFMA path and the mixed-precision path are both timed;
each solve's density is checked against an FP64 eigendecomposition
(||D - D_diag||_F, |Tr D - N|, band-energy error) outside the timed region.

Prints one JSON line (rank 0, one GPU).
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "qmd-progress_amd"))

PEAK = {  # MI355X dense peaks
    "hbm_gbs": 8000.0,
    "fp64_tflops": 78.6,
    "fp16_tflops": 2500.0,
    "int8_tops": 5000.0,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=8192)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--precision", default="auto",
                    help="fp64 | f16split | bf16split | int8crt (every --moduli count) | "
                         "int8crt<m> | auto (all)")
    ap.add_argument("--moduli", default="10,11",
                    help="comma-separated moduli counts of the int8 Chinese-remainder products "
                         "(6..14); each runs as mode int8crt<m>")
    ap.add_argument("--cpu-iters", type=int, default=5,
                    help="full-size SP2 iterations timed for the CPU baseline sample (>= 5, "
                         "after one warm-up; the median is used)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="dense",
                    choices=["dense", "graph", "gpt0", "fermi", "adjoint"],
                    help="dense: prg_sp2_alg2 at --n (the headline); graph: "
                         "prg_subgraphSP2Loop on the water-box system; gpt0: "
                         "prg_build_densityGP_T0 on the same system; fermi: "
                         "prg_implicit_fermi (kT 0.05 eV, 8 levels) at --n; adjoint: "
                         "forward + reverse SP2 sweep for dTr(D A)/dH at --n")
    ap.add_argument("--natoms", type=int, default=10000)
    ap.add_argument("--nparts", type=int, default=512)
    ap.add_argument("--graph-threshold", type=float, default=0.11,
                    help="|D| threshold of the density graph (graph workload)")
    ap.add_argument("--water", default="gapped", choices=["gapless", "gapped"],
                    help="graph workload: on-site energies (-0.5, 0, 0.5) eV at half filling "
                         "(gapless) or (-8, 4, 4) eV with one occupied orbital per atom "
                         "(a 5 eV gap, so the thresholded density graph can hold D)")
    ap.add_argument("--graph-precision", default="fp64", choices=["fp64", "f16split", "bf16split"],
                    help="graph workload: products of the batched subgraph kernel")
    ap.add_argument("--products", default="fp64",
                    help="fermi workload: arithmetic of the symmetric products, fp64, "
                         "int8crt<m> (int8 CRT products with m moduli) or bf16split "
                         "(BF16 hi+lo split products, FP32 accumulation; pass an --ns-tol "
                         "above their noise floor, e.g. 1e-4)")
    ap.add_argument("--mu-steps", type=int, default=None,
                    help="fermi workload: fixed Newton steps on mu (0: until |Tr P - N/2| "
                         "<= 1e-6; default 0, and 4 with --products bf16split)")
    ap.add_argument("--ns-tol", type=float, default=1e-10,
                    help="fermi workload: Newton-Schulz tolerance ||I - Y Inv||_F")
    ap.add_argument("--adjoint-of", default="sp2", choices=["sp2", "fermi"],
                    help="adjoint workload: the SP2 recursion or the fixed-mu implicit-Fermi map")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 256 MiB device copy first (reference kernel for the "
                         "FETCH_SIZE / WRITE_SIZE counter scale)")
    return ap.parse_args()


def model_hamiltonian_torch(n, seed, device):
    """Same generator as the oracle's model_hamiltonian (numpy PCG64 stream),
    built on the host once and copied to HBM."""
    import numpy as np
    import torch
    rng = np.random.default_rng(seed)
    z = rng.standard_normal((n, n))
    idx = np.arange(n)
    dist = np.abs(idx[:, None] - idx[None, :])
    off = -np.exp(-dist / 8.0) * (1.0 + 0.1 * z)
    del z
    off[dist > 64] = 0.0
    del dist
    np.fill_diagonal(off, 0.0)
    h = 0.5 * (off + off.T)
    del off
    h[idx, idx] = rng.uniform(-1.0, 1.0, n)
    return torch.from_numpy(h).to(device)


def errors_vs_diag(h, rho, evals, evecs, n):
    """||D - D_diag||_F, |Tr D - N|, |Tr(H D) - Tr(H D_diag)| (FP64, on the GPU)."""
    import torch
    nocc = n // 2
    q = evecs[:, :nocc]
    ddiag = 2.0 * (q @ q.T)
    err = torch.linalg.norm(rho - ddiag).item()
    tr = abs(torch.trace(rho).item() - n)
    eb = abs(torch.sum(h * rho).item() - 2.0 * evals[:nocc].sum().item())
    return err, tr, eb


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU per run: multi-GPU runs are not supported",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import numpy as np
    import torch
    import progress_amd as P

    if args.copy_calibration:
        dev = torch.device("cuda:0")
        src = torch.ones(32 * 1024 * 1024, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "graph":
        graph_bench(args, P, torch, np)
        return
    if args.workload == "gpt0":
        gpt0_bench(args, P, torch, np)
        return
    if args.workload == "fermi":
        fermi_bench(args, P, torch, np)
        return
    if args.workload == "adjoint":
        adjoint_bench(args, P, torch, np)
        return
    n = args.n
    dev = torch.device("cuda:0")
    ld = P.padded_ld(n)
    h = model_hamiltonian_torch(n, args.seed, dev)
    hpad = torch.zeros((ld, ld), dtype=torch.float64, device=dev)
    hpad[:n, :n] = h  # symmetric: row/column-major views coincide
    rho = torch.zeros_like(hpad)

    # the library's default moduli count, fixed before any solve: the headline
    # mode (the D_diag comparison below only gates it)
    default_moduli = P.get_int8_moduli()
    moduli = [int(m) for m in str(args.moduli).split(",") if m]
    if default_moduli not in moduli:
        moduli.append(default_moduli)
    modes = {"fp64": (P.PREC_FP64, 0), "f16split": (P.PREC_F16SPLIT, 0),
             "bf16split": (P.PREC_BF16SPLIT, 0)}
    for m in moduli:
        modes[f"int8crt{m}"] = (P.PREC_INT8CRT, m)
    if args.precision == "auto":
        run_modes = list(modes)
    elif args.precision == "int8crt":
        run_modes = [f"int8crt{m}" for m in moduli]
    else:
        run_modes = [args.precision]

    evals = evecs = None
    if not args.no_check:
        t0 = time.time()
        evals, evecs = torch.linalg.eigh(h)
        torch.cuda.synchronize()
        t_eigh = time.time() - t0
        gap = (evals[n // 2] - evals[n // 2 - 1]).item()
    else:
        t_eigh = None
        gap = None

    results = {}
    for name in run_modes:
        mode, nmod = modes[name]
        if nmod:
            P.set_int8_moduli(nmod)
        try:
            for _ in range(args.warmup):
                P.sp2_alg2_device(hpad, rho, n, 0.5, 1, 200, "Abs", 1e-10, precision=mode)
        except P.ProgressError as e:
            results[name] = {"error": str(e)}
            continue
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        iters = []
        for _ in range(args.steps):
            it, pp = P.sp2_alg2_device(hpad, rho, n, 0.5, 1, 200, "Abs", 1e-10, precision=mode)
            iters.append(it)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ms = 1e3 * dt / args.steps
        it = iters[-1]
        alg_flops = 2.0 * float(n) ** 3 * it
        r = {
            "solves_per_s": args.steps / dt,
            "ms_per_solve": ms,
            "iterations": it,
            "alg_tflops": alg_flops / (ms * 1e-3) / 1e12,
        }
        # kernel-level timing of the product launches (separate untimed pass)
        P.step_timing(True)
        P.sp2_alg2_device(hpad, rho, n, 0.5, 1, 200, "Abs", 1e-10, precision=mode)
        ms_sum, cnt = P.step_timing()
        pms_sum, pcnt = P.int8_product_timing()
        P.step_timing(False)
        if cnt:
            r["step_ms"] = ms_sum / cnt
            r["step_launches"] = cnt
            # the product kernel: the whole step for fp64 / f16split; the
            # int8 CRT product kernel alone (its FP64 update kernel apart)
            kms = (pms_sum / pcnt) if (nmod and pcnt) else ms_sum / cnt
            r["product_kernel_ms"] = kms
            r["product_kernel_launches"] = pcnt if nmod else cnt
            r["product_kernel_fp64_equiv_tflops"] = 2.0 * float(n) ** 3 / (kms * 1e-3) / 1e12
            ops = executed_ops(ld, "int8crt" if nmod else name, nmod)
            r["product_kernel_mfma_ops"] = ops
            r["product_kernel_mfma_tops"] = ops / (kms * 1e-3) / 1e12
            if nmod:
                r["moduli"] = nmod
                r["update_kernel_ms"] = r["step_ms"] - kms
                r["int8_config"] = P.int8_waves()
        if evals is not None:
            rr = rho[:n, :n]
            err, tr, eb = errors_vs_diag(h, rr, evals, evecs, n)
            r.update({"fro_err_vs_diag": err, "trace_err": tr, "band_energy_err": eb,
                      "meets_1e-6": bool(err <= 1e-6)})
        results[name] = r

    # headline: int8 CRT products with the library's default moduli count
    # (chosen before the solve); the check against D_diag is a pass/fail gate
    # (fp64 when it fails).  The fastest mode meeting the bar on this input is
    # reported beside it, not as the headline: only D_diag selects it.
    ok = [k for k, v in results.items() if "solves_per_s" in v and v.get("meets_1e-6", True)]
    fastest_ok = max(ok, key=lambda k: results[k]["solves_per_s"]) if ok else None
    head = f"int8crt{default_moduli}"
    if head not in ok:
        head = "fp64"
    hr = results.get(head, {})
    kind = "int8crt" if head.startswith("int8crt") else head
    hmod = hr.get("moduli")
    peak_tf = {"fp64": PEAK["fp64_tflops"], "f16split": PEAK["fp16_tflops"],
               "bf16split": PEAK["fp16_tflops"], "int8crt": PEAK["int8_tops"]}[kind]
    kms = hr.get("product_kernel_ms")
    alg = (2.0 * float(n) ** 3 / (kms * 1e-3) / 1e12) if kms else None
    # the symmetric square's own count: N (N + 1) / 2 distinct outputs of 2N flop
    alg_sym = (float(n) ** 2 * (n + 1) / (kms * 1e-3) / 1e12) if kms else None
    kern = hr.get("product_kernel_mfma_tops")
    roofline = {
        "bound": "mfma",
        "kernel": {"fp64": ("sp2_step_kernel (FP64 MFMA, fused SP2 update)"
                            if P.sp2_split_slices(n) == 1 else
                            f"sp2_split_product_kernel + sp2_split_epilogue_kernel (FP64 MFMA, "
                            f"K in {P.sp2_split_slices(n)} slices; the pair's time)"),
                   "f16split": "sp2_f16split_kernel (FP16 MFMA, 3 products)",
                   "bf16split": "sp2_f16split_kernel<bf16> (BF16 MFMA, 3 products)",
                   "int8crt": f"{crt_kernel_name(hr.get('int8_config'))} (int8 MFMA, {hmod} "
                              "moduli x upper-triangle 256x256 tiles)"}[kind],
        "achieved": alg,
        "peak": peak_tf,
        "unit": "TFLOP/s",
        "frac": (alg / peak_tf) if alg else None,
        "frac_symmetric": (alg_sym / peak_tf) if alg_sym else None,
        "traffic": traffic_bytes(kind, n, hmod, hr.get("int8_config")),
        "executed": kern,
        "executed_frac": (kern / peak_tf) if kern else None,
        "note": "achieved = algorithmic 2N^3 flop of the X^2 the launch produces / the "
                "kernel's HIP-event average duration on the library stream (frac_symmetric: "
                "the symmetric square's own count N^2 (N + 1), at most 1 for a kernel that "
                "computes each distinct output once); peak = the dense "
                "MFMA peak of the type that executes it (FP64 78.6 TF, FP16 2.5 PF, int8 "
                "5.0 POP/s).  executed = MFMA operations the launch issues (upper-triangle "
                "tiles only; one pass per modulus for int8crt, 3 for f16split) / the same "
                "time.  traffic = HBM bytes per launch from the committed FETCH_SIZE / "
                "WRITE_SIZE counter summary of this workload (null when none matches)",
    }

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, args.seed, args.cpu_iters, results.get("fp64", {}).get("iterations"))

    fp64 = results.get("fp64", {})
    out = {
        "metric": "SP2 solves/s + MFMA TFLOP/s, mixed vs FP64 at ||D-D_diag||_F<=1e-6, N=8192",
        "value": hr.get("solves_per_s"),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": hr.get("ms_per_solve"),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "f16split": "f16 (hi+lo split), f64 bookkeeping",
                  "bf16split": "bf16 (hi+lo split), f64 bookkeeping",
                  "int8crt": "int8 residues (exact), f64 reconstruction"}[kind],
        "data": "synthetic",
        "config": {"workload": f"prg_sp2_alg2 dense model tight-binding H, N={n}, occ=N/2, "
                               f"idemtol=1e-10, seed={args.seed}",
                   "n": n, "headline_mode": head, "moduli": hmod,
                   "moduli_rule": "library default, fixed before the solve; D_diag gates only",
                   "fastest_mode_meeting_bar": fastest_ok, "gap_eV": gap,
                   "eigh_s": t_eigh},
        "tflops": hr.get("alg_tflops"),
        "modes": results,
        "mixed_vs_fp64": {k: (v["solves_per_s"] / fp64["solves_per_s"])
                          for k, v in results.items()
                          if "solves_per_s" in v and "solves_per_s" in fp64 and k != "fp64"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def batched_executed_flops(m, panel):
    """MFMA flops one iteration of the batched SP2 kernel executes on a part
    of padded size m: the upper-triangle tiles over panels of `panel` (the
    last one 64 wide when panel does not divide m); a diagonal tile leaves
    its strictly lower quadrant to the mirror write (3/4 of its product)."""
    widths = [min(panel, m - s) for s in range(0, m, panel)]
    f = 0.0
    for b, wb in enumerate(widths):
        for a in range(b + 1):
            wa = widths[a]
            f += 2.0 * wa * wb * m * (0.75 if a == b else 1.0)
    return f


def batched_tile_count(m, panel):
    """Upper-triangle tiles of a part of padded size m over panels of `panel`."""
    k = len(range(0, m, panel))
    return k * (k + 1) // 2


def progress(msg):
    """One line on stderr per setup phase or step, so that a long untimed
    setup is visible while it runs (the JSON line stays alone on stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def graph_bench(args, P, torch, np):
    """Graph-partitioned SP2 (prg_subgraphSP2Loop) on the water-box system:
    natoms atoms x 3 orbitals.  Setup (untimed): a dense SP2 solve of the whole
    system on the GPU gives the branch sequence and the density D; the graph
    is |D| >= threshold, split into nparts parts by the BFS partitioner.
    Timed: whole prg_subgraphSP2Loop calls with H, the graph and rho resident
    in HBM (library handles): core+halo extraction on the GPU, the host batch
    plan, gather + normalise of every core+halo block, the lock-step batched
    purification, the core-row scatter; the device phases are reported apart.
    Accuracy: the assembled density against the dense solve."""
    from progress_amd import workloads as W
    t0 = time.time()
    gapped = args.water == "gapped"
    onsite = (-8.0, 4.0, 4.0) if gapped else (-0.5, 0.0, 0.5)
    bndfil = 1.0 / 3.0 if gapped else 0.5
    rows, cols, vals, n, box = W.water_box(natoms=args.natoms, onsite=onsite, seed=args.seed)
    t_gen = time.time() - t0
    progress(f"water box generated: {n} orbitals, {len(vals)} entries, {t_gen:.1f} s")
    dev = torch.device("cuda:0")
    ld = P.padded_ld(n)
    hd = torch.zeros((ld, ld), dtype=torch.float64, device=dev)
    hd[torch.from_numpy(rows).to(dev), torch.from_numpy(cols).to(dev)] = \
        torch.from_numpy(vals).to(dev)
    rd = torch.zeros_like(hd)
    P.set_int8_moduli(max(int(m) for m in str(args.moduli).split(",") if m))
    t0 = time.time()
    it, pp = P.sp2_alg2_device(hd, rd, n, bndfil, 1, 200, "Abs", 1e-10,
                               precision=P.PREC_INT8CRT)
    torch.cuda.synchronize()
    t_dense = time.time() - t0
    progress(f"dense int8 CRT solve: {it} iterations, {t_dense:.1f} s")
    lo, hi = P.gershgorin_device(hd, n)
    d_dense = rd[:n, :n]
    g = (d_dense.abs() >= args.graph_threshold).to(torch.float64).cpu().numpy()
    h = hd[:n, :n].cpu().numpy()
    progress("graph |D| >= threshold copied to the host")
    gp = P.bfs_partition(g.T, args.nparts)
    progress(f"BFS partition: {gp.totalParts} parts")
    # library handles built once; their device copies stay resident, so each
    # timed call is the whole prg_subgraphSP2Loop: core+halo extraction on the
    # GPU, the host batch plan, gather, batched purification, scatter
    mh, mg, mr = P.Matrix.from_numpy(h), P.Matrix.from_numpy(g), P.Matrix.zeros(n)
    del g
    progress("library handles of H, the graph and rho built")
    times = []
    walls = []
    gprec = {"fp64": P.PREC_FP64, "f16split": P.PREC_F16SPLIT,
             "bf16split": P.PREC_BF16SPLIT}[args.graph_precision]
    P.set_precision(gprec)
    try:
        for k in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            _, vv = P.subgraph_sp2_loop(mh, mg, gp, pp, it, lo, hi, out=mr)
            wall = time.perf_counter() - t0
            progress(f"subgraph loop call {k}: {wall * 1e3:.1f} ms")
            ga, lp, sc = P.subgraph_timing()
            if k >= args.warmup:
                times.append((ga + lp + sc) * 1e-3)
                walls.append(wall)
                last = (ga, lp, sc)
    finally:
        P.set_precision(P.PREC_FP64)
    rho = mr.numpy()
    progress("accuracy against the dense solve")
    ls = [gp.lsize[i] for i in range(gp.totalParts)]
    core = [gp.llsize[i] for i in range(gp.totalParts)]
    dt = sum(walls) / len(walls)
    dev_s = sum(times) / len(times)
    alg = sum(2.0 * float(m) ** 3 for m in ls) * it
    panel, depth, kslab, launch = P.subgraph_config()
    mpad = [max(64, ((m + 63) // 64) * 64) for m in ls]
    executed = sum(batched_executed_flops(m, panel) for m in mpad) * it
    if args.graph_precision != "fp64":
        executed *= 3.0  # hi.hi, hi.lo and lo.hi 16-bit products per tile
    loop_s = last[1] * 1e-3
    dd = d_dense.cpu().numpy()
    err = float(np.linalg.norm(rho - dd))
    rel = err / float(np.linalg.norm(dd))
    tr_err = abs(float(np.trace(rho)) - 2.0 * bndfil * n)  # rho = 2 X, Tr X = bndfil N
    # band energy Tr(rho H) against the dense solve's (rho = 2 X on both sides)
    eb_err = abs(float(np.sum(rho * h)) - float(np.sum(dd * h)))
    achieved = executed / loop_s / 1e12
    cpu = None if args.no_cpu else graph_cpu_baseline(h, gp, pp, it, lo, hi)
    kdesc = {"fp64": "batched FP64 MFMA", "f16split": "batched FP16 hi+lo split MFMA",
             "bf16split": "batched BF16 hi+lo split MFMA"}[args.graph_precision]
    out = {
        "metric": f"graph-partitioned SP2 solves/s (prg_subgraphSP2Loop, {kdesc})",
        "value": 1.0 / dt,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "f16split": "f16 (hi+lo split), f64 update",
                  "bf16split": "bf16 (hi+lo split), f64 update"}[args.graph_precision],
        "data": "synthetic",
        "config": {"workload": f"water box ({args.water}: on-site {onsite} eV, occupation "
                               f"{bndfil:.4f} N): {args.natoms} atoms x 3 orbitals = {n} orbitals, "
                               f"box {box:.2f} A, hopping -exp(-r/1.5) (r < 4 A), graph "
                               f"|D| >= {args.graph_threshold}, {args.nparts} BFS parts, "
                               f"double-jump halos, seed {args.seed}",
                   "n": n, "iterations": it, "parts": gp.totalParts,
                   "lsize_min": min(ls), "lsize_mean": float(np.mean(ls)),
                   "lsize_p90": float(np.percentile(ls, 90)), "lsize_max": max(ls),
                   "core_mean": float(np.mean(core))},
        "phases_ms": {"gather": last[0], "loop": last[1], "scatter": last[2],
                      "device_total": dev_s * 1e3,
                      "extraction_and_plan": (dt - dev_s) * 1e3},
        "setup_s": {"generate": t_gen, "dense_int8crt_solve": t_dense},
        "alg_tflops": alg / dt / 1e12,
        "accuracy_vs_dense": {"fro": err, "rel_fro": rel, "trace_err": tr_err,
                              "band_energy_err": eb_err},
        "tiles": {"panel": panel, "prefetch_slabs": depth, "k_slab": kslab,
                  "count": sum(batched_tile_count(m, panel) for m in mpad),
                  "launch": {0: "eager", 1: "captured hipGraph", 2: "hipGraph replay"}[launch]},
        "roofline": {"bound": "mfma",
                     "kernel": (f"batched_sp2_kernel (FP64 MFMA, {panel}-wide panels)"
                                if args.graph_precision == "fp64" else
                                f"batched split-product kernel ({args.graph_precision}, 3 "
                                f"16-bit products per tile)"),
                     "achieved": achieved,
                     "peak": PEAK["fp64_tflops" if args.graph_precision == "fp64" else
                                  "fp16_tflops"],
                     "unit": "TFLOP/s",
                     "frac": achieved / PEAK["fp64_tflops" if args.graph_precision == "fp64"
                                             else "fp16_tflops"],
                     "traffic": traffic_bytes("graph", n, None, args.water)
                     if (panel, depth, kslab, args.graph_precision) == (64, 2, 16, "fp64")
                     else None,
                     "algorithmic_tflops": alg / loop_s / 1e12,
                     "note": "executed MFMA work of the padded upper-triangle tiles per "
                             "iteration / loop device time"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


def gpt0_bench(args, P, torch, np):
    """prg_build_densityGP_T0 on the water-box system: the graph is |D| >=
    threshold of a dense int8-CRT solve (setup, untimed); each timed call
    partitions it (BFS bisection on the host), extracts the core+halo lists on
    the GPU, solves every part by batched eigendecomposition (kT = 0.1 eV
    Fermi densities), assembles the core rows and takes the reference's Ef
    step.  H, the graph and rho stay resident in library handles.  Accuracy:
    the assembled density against the dense solve (T = 0)."""
    from progress_amd import workloads as W
    rows, cols, vals, n, box = W.water_box(natoms=args.natoms, seed=args.seed)
    dev = torch.device("cuda:0")
    ld = P.padded_ld(n)
    hd = torch.zeros((ld, ld), dtype=torch.float64, device=dev)
    hd[torch.from_numpy(rows).to(dev), torch.from_numpy(cols).to(dev)] = \
        torch.from_numpy(vals).to(dev)
    rd = torch.zeros_like(hd)
    P.set_int8_moduli(max(int(m) for m in str(args.moduli).split(",") if m))
    P.sp2_alg2_device(hd, rd, n, 0.5, 1, 200, "Abs", 1e-10, precision=P.PREC_INT8CRT)
    torch.cuda.synchronize()
    d_dense = rd[:n, :n]
    g = (d_dense.abs() >= args.graph_threshold).to(torch.float64).cpu().numpy()
    h = hd[:n, :n].cpu().numpy()
    # Ef starts from the midgap of a quarter-size box of the same density (the
    # Fermi level is intensive; an eigendecomposition at the full size is out
    # of the solver's reach); the calls take the routine's own Ef steps
    r4, c4, v4, n4, _ = W.water_box(natoms=max(1, args.natoms // 4), seed=args.seed)
    h4 = torch.zeros((n4, n4), dtype=torch.float64, device=dev)
    h4[torch.from_numpy(r4).to(dev), torch.from_numpy(c4).to(dev)] = torch.from_numpy(v4).to(dev)
    w4 = torch.linalg.eigvalsh(h4)
    ef0 = 0.5 * (w4[n4 // 2 - 1] + w4[n4 // 2]).item()
    del h4, w4
    mh, mg, mr = P.Matrix.from_numpy(h), P.Matrix.from_numpy(g), P.Matrix.zeros(n)
    del g, h
    walls = []
    ef = ef0
    for k in range(args.warmup + args.steps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        _, ef = P.build_density_gp_t0(mh, mg, 0.5, ef, args.nparts, out=mr)
        if k >= args.warmup:
            walls.append(time.perf_counter() - t0)
    dt = sum(walls) / len(walls)
    rho = mr.numpy()
    dd = d_dense.cpu().numpy()
    err = float(np.linalg.norm(rho - dd))
    _line("prg_build_densityGP_T0 calls/s (batched rocSOLVER eigensolves + FP64 MFMA)",
          1.0 / dt, "calls/s", dt, args, "f64",
          {"workload": f"water box: {args.natoms} atoms x 3 orbitals = {n} orbitals, graph |D| "
                       f">= {args.graph_threshold}, {args.nparts} BFS parts, kT 0.1 eV, "
                       f"Ef from {ef0:.6f} (midgap of a quarter-size box), seed {args.seed}",
           "n": n,
           "ef_final": ef},
          {"accuracy_vs_dense_t0": {"fro": err, "rel_fro": err / float(np.linalg.norm(dd)),
                                    "trace_err": abs(float(np.trace(rho)) - n)},
           "cpu_baseline": None, "roofline": None})


def _line(metric, value, unit, dt, args, dtype, config, extra, higher=True):
    out = {"metric": metric, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dt * 1e3, "higher_is_better": higher,
           "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
           "config": config}
    out.update(extra)
    print(json.dumps(out))


def fermi_bench(args, P, torch, np):
    """prg_implicit_fermi (recursive Fermi-operator expansion, CG / Newton-Schulz
    inverses) on the model Hamiltonian at --n, beta = 20 eV^-1 (kT 0.05),
    8 levels, k = 2, occupation N/2, its symmetric products in FP64, as int8
    CRT products (--products int8crt<m>) or as BF16 hi+lo split products
    (--products bf16split).  The library matrix handles stay resident across
    the timed calls (H and P in HBM).  Accuracy: P against the Fermi function
    of an FP64 eigendecomposition at the converged mu, and (int8 or BF16
    products) against the FP64-product solve at the same tolerance."""
    n = args.n
    h = model_hamiltonian_torch(n, args.seed, torch.device("cuda:0"))
    hh = h.cpu().numpy()
    mh = P.Matrix.from_numpy(hh)
    mp = P.Matrix.zeros(n)
    beta, nsteps, nocc = 20.0, 8, 0.5 * n
    L = P.lib()
    w = torch.linalg.eigvalsh(h)
    mu = 0.5 * (w[n // 2 - 1] + w[n // 2]).item()
    products = args.products
    moduli = None
    if products.startswith("int8crt") and products[7:].isdigit():
        moduli = int(products[7:])
    elif products not in ("fp64", "bf16split"):
        raise SystemExit(f"--products: fp64, int8crt<m> or bf16split, not {products}")

    # osteps = 0: Newton steps on mu until |Tr P - nocc| <= 1e-6.  BF16 split
    # products carry about 16 significant bits, so the trace cannot reach that
    # limit: their solves (and the FP64 solve they are compared with) take four
    # Newton steps on mu instead, unless --mu-steps says otherwise.
    osteps = args.mu_steps if args.mu_steps is not None else (4 if products == "bf16split" else 0)

    def solve(mu0, method):
        L.prg_implicit_fermi(mh.h, mp.h, nsteps, 2, nocc, mu0, beta, method, osteps, 1e-6, 0.0,
                             args.ns_tol)
        P._check()
        progress(f"fermi: {products} solve done")
        return L.prg_implicit_fermi_last_mu()

    method = 1  # Newton-Schulz inverses (CG fallback)
    p_fp64 = None
    if products != "fp64":
        P.set_product_precision(P.PREC_FP64)
        solve(mu, method)
        p_fp64 = torch.from_numpy(mp.numpy()).cuda()
        if moduli is not None:
            P.set_int8_moduli(moduli)
            P.set_product_precision(P.PREC_INT8CRT)
        else:
            P.set_product_precision(P.PREC_BF16SPLIT)
    try:
        for _ in range(args.warmup):
            solve(mu, method)
        torch.cuda.synchronize()
        P.gemm_counters(reset=True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            mu1 = solve(mu, method)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        nprod, alg, exe, i8 = (v / args.steps for v in P.gemm_counters(reset=True, int8=True))
    finally:
        P.set_product_precision(P.PREC_FP64)
    pm = torch.from_numpy(mp.numpy()).cuda()
    wv, q = torch.linalg.eigh(h)
    f = 1.0 / (1.0 + torch.exp(beta * (wv - mu1)))
    pref = (q * f) @ q.T
    err = torch.linalg.norm(pm - pref).item()
    tr = abs(torch.trace(pm).item() - nocc)
    # band energy 2 Tr(P H) (spin-paired occupation) against the eigendecomposition's
    eb = abs(2.0 * torch.sum(pm * h).item() - 2.0 * torch.sum(f * wv).item())
    acc = {"fro_err_vs_fermi_eigh": err, "trace_err": tr, "band_energy_err": eb}
    if p_fp64 is not None:
        acc["fro_diff_vs_fp64_products"] = torch.linalg.norm(pm - p_fp64).item()
    if products == "bf16split":
        # executed BF16 MFMA flops: three products over the upper-triangle
        # 128 x 128 tiles of the padded size per symmetric product
        ld = P.padded_ld(n)
        nt = ld // 128
        ex16 = nprod * 3 * nt * (nt + 1) / 2 * 2.0 * 128 * 128 * ld
        roof = {"bound": "mfma", "kernel": "split_gemm_kernel<bf16> (BF16 MFMA, 3 products)",
                "achieved": ex16 / dt / 1e12, "peak": PEAK["fp16_tflops"],
                "unit": "TFLOP/s", "frac": ex16 / dt / 1e12 / PEAK["fp16_tflops"],
                "traffic": None,
                "note": "BF16 MFMA flops executed by the solve's split products / "
                        "whole-solve wall time (plane splits, host recurrences and syncs "
                        "included: a lower bound of the kernel rate)"}
        metric = "implicit Fermi solves/s (prg_implicit_fermi, BF16 hi+lo split products)"
        dtype = "bf16/f32/f64"
    elif moduli is None:
        roof = {"bound": "mfma", "kernel": "gemm_kernel (FP64 MFMA)",
                "achieved": exe / dt / 1e12, "peak": PEAK["fp64_tflops"],
                "unit": "TFLOP/s", "frac": exe / dt / 1e12 / PEAK["fp64_tflops"],
                "traffic": None,
                "note": "MFMA flops executed by the solve's products / whole-solve "
                        "wall time (host recurrences, reductions and syncs included: "
                        "a lower bound of the kernel rate)"}
        metric = "implicit Fermi solves/s (prg_implicit_fermi, FP64 MFMA)"
        dtype = "f64"
    else:
        roof = {"bound": "mfma", "kernel": "int8 CRT product kernels (crt_product_kernel / crt_product_persistent_kernel, configuration timed per shape; int8 MFMA)",
                "achieved": i8 / dt / 1e12, "peak": PEAK["int8_tops"], "unit": "TOP/s",
                "frac": i8 / dt / 1e12 / PEAK["int8_tops"], "traffic": None,
                "note": "int8 MFMA operations executed by the solve's CRT products / "
                        "whole-solve wall time (splits, reconstructions, host recurrences "
                        "included: a lower bound of the kernel rate)"}
        metric = "implicit Fermi solves/s (prg_implicit_fermi, int8 CRT products)"
        dtype = "int8/f64"
    cpu = None if args.no_cpu else products_cpu_baseline(hh, nprod, args.cpu_iters)
    _line(metric, 1.0 / dt, "solves/s", dt, args, dtype,
          {"workload": f"prg_implicit_fermi model tight-binding H, N={n}, beta=20 (kT 0.05 eV), "
                       f"8 levels, k=2, method 1 (Newton-Schulz, tol {args.ns_tol:g}), "
                       + (f"{osteps} Newton steps on mu, " if osteps else "")
                       + f"seed={args.seed}", "n": n, "products": products,
           "mu_start": mu, "mu": mu1},
          {"accuracy": acc, "products_per_solve": nprod, "alg_tflops": alg / dt / 1e12,
           "cpu_baseline": cpu, "roofline": roof})


def adjoint_bench(args, P, torch, np):
    """Forward + reverse-mode adjoint of L = Tr(D A) on the model Hamiltonian
    at --n (A banded, bandwidth 32, N(0,1) entries) through the torch.autograd
    binding (FP64): SP2 with checkpoints every 8 steps, or (--adjoint-of
    fermi) the fixed-mu implicit-Fermi map (8 levels, beta 20, mu at the
    spectrum's midgap) with one central finite difference along A as a check."""
    if args.adjoint_of == "fermi":
        return fermi_adjoint_bench(args, P, torch, np)
    from progress_amd.autograd import sp2_density
    n = args.n
    dev = torch.device("cuda:0")
    h = model_hamiltonian_torch(n, args.seed, dev)
    g = torch.Generator(device="cpu").manual_seed(args.seed + 1)
    a = torch.randn((n, n), generator=g, dtype=torch.float64)
    idx = torch.arange(n)
    a[(idx[:, None] - idx[None, :]).abs() > 32] = 0.0
    a = (0.5 * (a + a.T)).to(dev)

    def step():
        hv = h.clone().requires_grad_(True)
        d = sp2_density(hv, 0.5, 1e-10, 8)
        loss = torch.sum(d * a)
        loss.backward()
        return loss.item(), hv.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    P.gemm_counters(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lv, grad = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    nprod, alg, exe = (v / args.steps for v in P.gemm_counters(reset=True))
    # symmetry of the gradient and the first-order response cross-check
    sym = torch.linalg.norm(grad - grad.T).item() / torch.linalg.norm(grad).item()
    _line("SP2 forward + adjoint solves/s (dTr(D A)/dH, FP64)", 1.0 / dt, "solves/s", dt, args,
          "f64",
          {"workload": f"sp2_density forward + backward, model H N={n}, A banded bw 32, "
                       f"checkpoints every 8 steps, seed={args.seed}", "n": n},
          {"loss": lv, "grad_asymmetry": sym, "reverse_products_per_solve": nprod,
           "reverse_alg_tflops": alg / dt / 1e12, "cpu_baseline": None, "roofline": None})


def fermi_adjoint_bench(args, P, torch, np):
    from progress_amd.autograd import implicit_fermi_density
    n = args.n
    dev = torch.device("cuda:0")
    h = model_hamiltonian_torch(n, args.seed, dev)
    g = torch.Generator(device="cpu").manual_seed(args.seed + 1)
    a = torch.randn((n, n), generator=g, dtype=torch.float64)
    idx = torch.arange(n)
    a[(idx[:, None] - idx[None, :]).abs() > 32] = 0.0
    a = (0.5 * (a + a.T)).to(dev)
    w = torch.linalg.eigvalsh(h)
    mu = 0.5 * (w[n // 2 - 1] + w[n // 2]).item()
    nsteps, beta, tol = 8, 20.0, args.ns_tol

    def step():
        hv = h.clone().requires_grad_(True)
        p = implicit_fermi_density(hv, nsteps, mu, beta, tol)
        loss = torch.sum(p * a)
        loss.backward()
        return loss.item(), hv.grad

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    P.gemm_counters(reset=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        lv, grad = step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    nprod, alg, exe = (v / args.steps for v in P.gemm_counters(reset=True))
    sym = torch.linalg.norm(grad - grad.T).item() / torch.linalg.norm(grad).item()
    # one central difference along A / ||A|| (step 1e-4) against <G, A> / ||A||
    ld = P.padded_ld(n)
    e = a / torch.linalg.norm(a)
    eps = 1e-4
    vals = []
    for sgn in (1.0, -1.0):
        hp = torch.zeros((ld, ld), dtype=torch.float64, device=dev)
        hp[:n, :n] = h + sgn * eps * e
        pp = torch.zeros_like(hp)
        P.implicit_fermi_device(hp, pp, n, nsteps, mu, beta, tol)
        vals.append(torch.sum(pp[:n, :n] * a).item())
    fd = (vals[0] - vals[1]) / (2.0 * eps)
    an = torch.sum(grad * e).item()
    _line("implicit Fermi forward + adjoint solves/s (dTr(P A)/dH, fixed mu, FP64)", 1.0 / dt,
          "solves/s", dt, args, "f64",
          {"workload": f"implicit_fermi_density forward + backward, model H N={n}, A banded "
                       f"bw 32, 8 levels, beta 20, mu midgap {mu:.6f}, Newton-Schulz tol "
                       f"{tol:g}, seed={args.seed}", "n": n},
          {"loss": lv, "grad_asymmetry": sym, "products_per_solve": nprod,
           "alg_tflops": alg / dt / 1e12,
           "fd_check": {"direction": "A/||A||", "step": eps, "fd": fd, "adjoint": an,
                        "rel_err": abs(fd - an) / abs(an)},
           "cpu_baseline": None, "roofline": None})


def executed_ops(ld, mode, moduli):
    """MFMA operations of one symmetric SP2 product launch: upper-triangle
    128x128 tiles (fp64, f16split, bf16split) or 256x256 tiles over the residue planes'
    leading dimension rounded up to 256 (int8crt)."""
    tm = 256 if mode == "int8crt" else 128
    kd = ((ld + 255) // 256) * 256 if mode == "int8crt" else ld
    nt = kd // tm
    per_pass = nt * (nt + 1) // 2 * 2.0 * tm * tm * kd
    return per_pass * {"fp64": 1, "f16split": 3, "bf16split": 3, "int8crt": moduli}[mode]


def crt_kernel_name(config):
    """The int8 product kernel a configuration code launches (9xx: the
    persistent one-workgroup-per-CU form; 8xx: one workgroup per item)."""
    if config is not None and int(config) >= 900:
        return "crt_product_persistent_kernel"
    return "crt_product_kernel"


def traffic_bytes(mode, n, moduli, config=None):
    """HBM bytes per launch of the dominant kernel from profiles/traffic.json
    (FETCH_SIZE and WRITE_SIZE per dispatch, scaled by the calibration copy),
    when that file holds this kernel at this size (and int8 configuration,
    where the entry names one); otherwise None."""
    path = ROOT / "profiles" / "traffic.json"
    if not path.exists():
        return None
    for e in json.loads(path.read_text()).get("kernels", []):
        if e.get("mode") == mode and e.get("n") == n and e.get("moduli", moduli) == moduli \
                and e.get("config", config) == config:
            return e["fetch_kib"] * e["fetch_scale"] * 1024 + e["write_kib"] * e["write_scale"] * 1024
    return None


def graph_cpu_baseline(h, gp, pp, iters, lo, hi, stride=8, budget_s=20.0):
    """The oracle's per-part prg_sp2_submatrix_inplace restatement (numpy /
    OpenBLAS dgemm) on every stride-th part of the partitioning (stopping
    once budget_s seconds of CPU solves are spent), all iterations, scaled to
    the whole partitioning by sum n^3 of the parts it solved."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import progress_oracle as O
    try:
        from threadpoolctl import threadpool_info
        threads = max(i.get("num_threads", 1) for i in threadpool_info()
                      if i.get("internal_api") == "openblas") if threadpool_info() else os.cpu_count()
    except Exception:
        threads = os.cpu_count()
    halos = gp.halos()
    sample = []
    vv = [0.0] * iters
    t = 0.0
    for p in range(0, len(halos), stride):
        if t >= budget_s:
            break
        idx = np.asarray(halos[p])
        sub = h[np.ix_(idx, idx)]
        t0 = time.perf_counter()
        O.sp2_submatrix(sub, pp, iters, vv, lo, hi, gp.llsize[p])
        t += time.perf_counter() - t0
        sample.append(p)
    cube = [float(len(x)) ** 3 for x in halos]
    scale = sum(cube) / sum(cube[p] for p in sample)
    solve_s = t * scale
    return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 per-part SP2 (numpy/OpenBLAS dgemm) on {len(sample)} of "
                      f"{len(halos)} parts (every {stride}th until {budget_s:.0f} s were "
                      f"spent), {iters} iterations each, "
                      f"{t:.2f} s, scaled by sum n^3 ({scale:.2f}x)"}


def products_cpu_baseline(h, nprod, iters):
    """The oracle's FP64 products (numpy / OpenBLAS dgemm) at the full N: `iters`
    products of the normalised H timed, scaled to the solve's product count
    (the recurrences and element-wise updates of the solve are not counted,
    which favours the CPU)."""
    import numpy as np
    try:
        from threadpoolctl import threadpool_info
        threads = max(i.get("num_threads", 1) for i in threadpool_info()
                      if i.get("internal_api") == "openblas") if threadpool_info() else os.cpu_count()
    except Exception:
        threads = os.cpu_count()
    x = np.asfortranarray(h / np.abs(h).sum(axis=0).max())
    t0 = time.perf_counter()
    for _ in range(iters):
        x = x @ x
    t = (time.perf_counter() - t0) / iters
    n = h.shape[0]
    solve_s = t * nprod
    return {"value": 1.0 / solve_s, "unit": "solves/s", "cores": threads, "kind": "port",
            "sample": f"{iters} FP64 products at N={n} (numpy/OpenBLAS dgemm, {t:.2f} s each), "
                      f"scaled to the solve's {nprod:.0f} products",
            "tflops": 2.0 * float(n) ** 3 / t / 1e12}


def host_description():
    """CPU model, logical CPUs and the BLAS numpy calls (threadpoolctl)."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    blas, threads = "unknown", os.cpu_count()
    try:
        from threadpoolctl import threadpool_info
        libs = [i for i in threadpool_info() if i.get("user_api") == "blas"]
        if libs:
            blas = f"{libs[0].get('internal_api')} {libs[0].get('version')}"
            threads = libs[0].get("num_threads", threads)
    except Exception:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "blas": blas, "blas_threads": threads}


def cpu_baseline(n, seed, iters, gpu_iters):
    """The oracle's FP64 prg_sp2_alg2 restatement (numpy dgemm on the host
    BLAS) on a bounded sample of the same workload: the normalisation, one
    untimed warm-up iteration, then `iters` (>= 5) full-size N x N iterations
    timed one by one; solves/s = 1 / (normalisation + median iteration x the
    GPU run's iteration count)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import numpy as np
    import progress_oracle as O
    iters = max(5, iters)
    host = host_description()
    h = O.model_hamiltonian(n, seed)
    occ = 0.5 * n
    t0 = time.perf_counter()
    x = O.normalize(h)
    t_norm = time.perf_counter() - t0
    trx = float(np.trace(x))

    def step(x, trx):
        # src/prg_sp2_mod.F90:321-349: X^2, limdiff, branch, recursed trace
        x2 = x @ x
        trx2 = float(np.trace(x2))
        limdiff = abs(trx2 - occ) - abs(2.0 * trx - trx2 - occ)
        if limdiff >= 1e-10:
            return 2.0 * x - x2, 2.0 * trx - trx2
        return x2, trx2

    x, trx = step(x, trx)  # warm-up
    times = []
    for _ in range(iters):
        t0 = time.perf_counter()
        x, trx = step(x, trx)
        times.append(time.perf_counter() - t0)
    t_it = float(np.median(times))
    it_total = gpu_iters or 66
    solve_s = t_norm + it_total * t_it
    out = {"value": 1.0 / solve_s, "unit": "solves/s", "cores": host["blas_threads"],
           "kind": "port",
           "sample": f"oracle FP64 SP2 alg2 (numpy dgemm, {host['blas']}, "
                     f"{host['blas_threads']} threads), N={n}: normalisation "
                     f"({t_norm:.2f} s) + 1 warm-up + {iters} timed full-size iterations, "
                     f"median {t_it:.3f} s (min {min(times):.3f}, max {max(times):.3f}), "
                     f"scaled to {it_total} iterations",
           "tflops": 2.0 * float(n) ** 3 / t_it / 1e12,
           "iteration_s": times}
    out.update(host)
    return out


if __name__ == "__main__":
    main()
