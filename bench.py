#!/usr/bin/env python3
"""Benchmark of the 2.5D FE forward + sensitivity path on one MI355X.

Metric: modelled quadrupoles/s (forward: assembly -> factor -> substitution ->
back-transform -> voltages) on the 240x80-element / 19,521-node / 96-electrode /
3,090-quadrupole workload, complex FP64; plus the adjoint-gradient time versus
the explicit-Jacobian path on the same model.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Prints one JSON line.  A step is one full forward over all 14 x 96 (k, l)
systems with the conductivities already resident in device memory.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
PEAK_HBM_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_baseline(w, nsys, K, sample=64):
    """Time the reference (oracle/_ref: the reference's Fortran routines built
    from its source) on a bounded sample of (k, l) systems; fall back to the
    C restatement when the reference build is absent."""
    import numpy as np
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    cores = max(1, min(cores, K))   # the reference runs at most kwnanz threads
    sample = min(sample, nsys)
    stride = max(1, nsys // sample)
    try:
        from oracle import ref
        if not ref.available():
            raise ImportError
        r = ref.Reference(w.mesh)
        r.load_data(w.strnr, w.vnr)
        r.rwaven()
        r.precal()
        first = 1
        t0 = time.perf_counter()
        ref.forward_sample(w.sigma, first, first + sample - 1, cores)
        dt = time.perf_counter() - t0
        kind = "reference"
    except ImportError:
        from oracle import cpu
        o = cpu.Oracle(w.mesh)
        o.rwaven(w.strnr, w.vnr)
        o.precal()
        t0 = time.perf_counter()
        o.forward(w.sigma, nthreads=cores, sys_range=(0, sample))
        dt = time.perf_counter() - t0
        kind = "port"
    del stride, np
    t_full = dt * nsys / sample
    return {
        "value": w.nanz / t_full,
        "unit": "quadrupoles/s",
        "cores": cores,
        "nproc": os.cpu_count(),
        "kind": kind,
        "sample": f"{sample} of {nsys} (k,l) systems (assemble+scale+chol+vre), "
                  f"{dt:.2f} s wall on {cores} OpenMP threads (the reference uses at most one per "
                  f"wavenumber) of {os.cpu_count()} host cores, extrapolated linearly to the full forward",
    }


def k20_rule(g):
    """20 wavenumbers built the way rwaven builds its set (src/rwaven.f90:
    129-151) -- Gauss-Legendre on [0, 1] mapped to k0 x^2 with weights 2 k0 w
    x, then 4 Gauss-Laguerre points mapped to k0 (x + 1) with weights k0 w e^x,
    k0 = 1 / (2 amin) -- with 16 Legendre points instead of the rule's count."""
    import numpy as np
    k0 = float(g.kwn[-4] / (np.polynomial.laguerre.laggauss(4)[0][0] + 1.0))
    x, wt = np.polynomial.legendre.leggauss(16)
    xs, ws = 0.5 * (x + 1.0), 0.5 * wt
    xl, wl = np.polynomial.laguerre.laggauss(4)
    kwn = np.concatenate([k0 * xs ** 2, k0 * (xl + 1.0)])
    kwnwi = np.concatenate([2.0 * k0 * ws * xs, k0 * wl * np.exp(xl)])
    return kwn, kwnwi


def uneven_batch(args):
    """COUNT independent models of uneven mesh and electrode counts (nx x nz
    quadrilaterals, nx ~ U[30,160], nz ~ U[15,50], 24-72 electrodes): one
    Gauss-Newton step each -- forward, -ln V, data-misfit gradient, explicit J
    and 20 CG iterations on J^H W J + lam R (lam = 10).  Three paths over the
    same prepared models (handles, element matrices and observed data are
    set up before the timed regions):

    * sequential: the plain per-model path, one model at a time (forward,
      explicit J, the data term J^H W r from J, launch-by-launch CG);
    * threaded: every model on its own handle and stream, ``threads`` in
      flight, adjoint gradient, graph-captured CG;
    * persistent: one crt_forward_many over all models (the factorisations of
      every model fed to persistent launches from one work queue), then the
      adjoint gradient, J and the graph-captured CG per model, threaded.

    The value is the persistent path's model steps per second."""
    import numpy as np
    from titatest_amd import batch, meshgen
    ws = meshgen.uneven_batch(count=args.uneven_batch, seed=1006)
    mods = []
    for i, w in enumerate(ws):
        nx = len(np.unique(w.mesh.sx)) - 1
        mods.append(batch.prepare(w, nx, w.mesh.elanz // nx, seed=1003 + i))
    log(f"prepared {len(mods)} models")
    for m in mods:  # warm-up: buffers, graph capture
        batch.gn_step(m, use_graph=True)
        batch.gn_step(m, use_graph=False)
    batch.run_batch_persistent(mods, threads=args.threads)
    batch.run_sequential(mods)
    t_seq, t_thr, t_per = [], [], []
    for rep in range(max(1, args.steps // 2)):
        t_seq.append(batch.run_sequential(mods)[1])
        t_thr.append(batch.run_batch(mods, threads=args.threads, use_graph=True)[1])
        t_per.append(batch.run_batch_persistent(mods, threads=args.threads)[1])
        log(f"round {rep + 1}: sequential {t_seq[-1] * 1e3:.1f} ms, threaded {t_thr[-1] * 1e3:.1f} ms, "
            f"persistent {t_per[-1] * 1e3:.1f} ms")
    ts, tt, tp = min(t_seq), min(t_thr), min(t_per)
    cells = sum(w.mesh.elanz for w in ws)
    quads = sum(w.nanz for w in ws)
    print(json.dumps({
        "metric": "Gauss-Newton steps/s over an uneven batch (gradient + 20 CG iterations), FP64, 1 GPU",
        "value": len(ws) / tp, "unit": "model steps/s", "n_gpus": 1, "steps": len(t_per),
        "warmup": 1, "ms_per_step": tp * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic (seeded random-field models, generated meshes)",
        "config": {"workload": f"{len(ws)} uneven models, {cells} cells, {quads} quadrupoles in total",
                   "threads": args.threads},
        "sequential_per_model_ms": ts * 1e3, "threaded_ms": tt * 1e3, "persistent_ms": tp * 1e3,
        "persistent_over_sequential": ts / tp, "threaded_over_sequential": ts / tt,
        "target_over_sequential": 20.0,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample", type=int, default=256)
    ap.add_argument("--skip-adjoint", action="store_true")
    ap.add_argument("--skip-dc", action="store_true", help="leave out the DC-twin forward comparison")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (calibrates HBM byte counters)")
    ap.add_argument("--uneven-batch", type=int, default=0, metavar="COUNT",
                    help="instead of the forward: one Gauss-Newton step on each of COUNT uneven models "
                         "(adjoint gradient + 20 CG iterations), graph-captured CG against the per-launch path")
    ap.add_argument("--threads", type=int, default=8, help="models in flight for --uneven-batch")
    ap.add_argument("--tri", action="store_true",
                    help="the terrain-following triangular mesh instead: 257x100 nodes (25,700), 50,688 "
                         "triangles, mb 101, 128 electrodes, 7,074 quadrupoles, mixed sides and bottom")
    ap.add_argument("--k20", action="store_true",
                    help="the 240x80 workload with 20 wavenumbers (the reference's quadrature construction "
                         "with 16 Gauss-Legendre points forced instead of its rule's 10): 1,920 systems, "
                         "a stress variant without a parity counterpart")
    ap.add_argument("--autotune", action="store_true",
                    help="choose the factor workgroup size and back-substitution depth by timing them first "
                         "(titatest_amd/autotune.py)")
    ap.add_argument("--dirichlet", action="store_true",
                    help="the same 240x80 model with Dirichlet sides and bottom: one shared factor per "
                         "wavenumber, then a forward and a backward sweep per source (kompb)")
    ap.add_argument("--recompute-passes", type=int, default=7, metavar="K",
                    help="also time forward + adjoint gradient on a second handle that holds K wavenumbers' "
                         "potentials at a time (crt_set_recompute; 0 leaves it out)")
    ap.add_argument("--recompute-budget", type=float, default=20.0, metavar="GIB",
                    help="device budget for factors in flight on the recomputing handle (crt_set_device_budget)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        nbytes = args.copy_probe << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nbytes, "dispatches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import numpy as np
    import torch

    from titatest_amd import crt, meshgen

    if not torch.cuda.is_available():
        sys.exit("bench.py needs a HIP device")
    if args.uneven_batch:
        uneven_batch(args)
        return
    dev = torch.device("cuda:0")

    if args.tri:
        w = meshgen.topo_triangles()
    else:
        w = meshgen.large_grf(side_type=13 if args.dirichlet else 11)
    m = w.mesh
    free_start = torch.cuda.mem_get_info()[0]
    g = crt.CRT(m)
    g.rwaven(w.strnr, w.vnr)
    if args.k20:
        g.set_wavenumbers(*k20_rule(g))
    g.precal()
    tuned = None
    if args.autotune:
        from titatest_amd import autotune
        choice, times = autotune.tune(g, w.sigma, force=True)
        tuned = {"factor_threads": choice[0], "backsolve_depth": choice[1],
                 "ms": {f"{t}x{d}": v for (t, d), v in times.items()}}
        log(f"autotune: {tuned}")
    n, mb, E, K, N = m.sanz, m.mb, m.eanz, g.K, w.nanz
    nsys = K * E
    nfac = K if args.dirichlet else nsys   # factorisations per forward
    log(f"workload {'topo_triangles' if args.tri else 'large_grf'}: n={n} mb={mb} E={E} K={K} N={N} "
        f"systems={nsys} factors={nfac}")
    sigma_dev = torch.tensor(w.sigma, dtype=torch.complex128, device=dev)
    st = np.ascontiguousarray(w.strnr, np.int32)
    vn = np.ascontiguousarray(w.vnr, np.int32)
    volt = torch.empty(N, dtype=torch.complex128, device=dev)
    sigmaa = torch.empty(N, dtype=torch.complex128, device=dev)

    def step():
        rc = g.lib.crt_forward_dev(g.h, sigma_dev.data_ptr())
        if rc:
            raise crt.CRTError("forward_dev", rc)
        rc = g.lib.crt_voltages_dev(g.h, N, st.ctypes.data, vn.ctypes.data, 0, volt.data_ptr(),
                                    sigmaa.data_ptr())
        if rc:
            raise crt.CRTError("voltages_dev", rc)

    torch.cuda.synchronize()
    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    g.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        step()
        log(f"step {i + 1}/{args.steps} done")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    times = g.stage_times()
    g.set_timing(False)
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = N / (ms_per_step / 1e3)

    # the factor kernel: 4 n mb^2 flops per band factorisation plus 8 n mb
    # for the forward sweep fused into it; L written once.  A forward may take
    # several launches (factors in flight are capped by the device budget):
    # per launch = per step / launches per step.
    fac_ms, fac_n = times.get("factor_solve", (float("nan"), 1))
    fac_avg_s = fac_ms / fac_n / 1e3
    launches = fac_n / args.steps
    flops = nfac * (4.0 * n * mb * mb + 8.0 * n * mb) / launches
    achieved = flops / fac_avg_s / 1e12
    lbytes = nfac * 16.0 * n * (mb + 1) / launches

    # HBM bytes per factor launch from the committed rocprofv3 FETCH_SIZE /
    # WRITE_SIZE passes of this workload (scaled by the copy-kernel calibration)
    traffic = None
    tf = ROOT / "profiles" / "factor_traffic_large_grf.json"
    if tf.exists() and not (args.dirichlet or args.tri):
        try:
            traffic = json.loads(tf.read_text())["factor"]["hbm_bytes"]
        except (ValueError, KeyError):
            traffic = None

    result = {
        "metric": "modelled quadrupoles/s and adjoint-gradient ms vs explicit J, FP64, 1 GPU",
        "value": value,
        "unit": "quadrupoles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded Gaussian random-field model, generated mesh)",
        "config": {"workload": (f"terrain-following triangles, {m.elanz} elements, {n} nodes, mb {mb}, "
                                f"{E} electrodes, {N} dipole-dipole quadrupoles, {K} wavenumbers x {E} sources, "
                                "mixed sides and bottom" if args.tri else
                                "240x80 quads, 19521 nodes, mb 82, 96 electrodes, "
                                "3090 dipole-dipole quadrupoles, 14 wavenumbers x 96 sources"
                                + (", Dirichlet sides and bottom (one factor per wavenumber)"
                                   if args.dirichlet else ", mixed sides and bottom")
                                + (", 20-wavenumber stress variant (no parity counterpart)" if args.k20 else "")),
                   "systems": nsys, "factorisations": nfac, "nodes": n, "bandwidth": mb, "electrodes": E,
                   "quadrupoles": N, "wavenumbers": K},
        "roofline": {"bound": "mfma", "pipe": "FP64 vector ALU (v_fma_f64); MI355X's FP64 vector and "
                     "matrix peaks are both 78.6 TFLOP/s", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                     "kernel": "crt_factor_kernel", "avg_launch_ms": fac_ms / fac_n,
                     "launches_per_step": launches,
                     "algorithmic_flops": flops, "algorithmic_bytes": lbytes},
        "stage_ms_per_step": {k: v[0] / args.steps for k, v in times.items()},
    }
    if tuned is not None:
        result["config"]["tuning"] = tuned

    if not args.skip_adjoint:
        # config with data: the adjoint gradient of the weighted misfit vs the
        # explicit-J data term, both after the same forward
        rng = np.random.default_rng(1004)
        g.forward(w.sigma)
        v1, sa1 = g.bvolti(st, vn)
        dat = sa1 + 0.02 * rng.standard_normal(N)
        wgt = 1.0 / (0.02 + 1e-4 / np.abs(v1)) ** 2
        r = dat - sa1
        g.set_timing(True)
        ga = None
        for _ in range(3):
            ga = g.adjoint_grad(wgt, r, mode=1)
        t_adj = g.stage_times()
        g.set_timing(True)
        gj = None
        for _ in range(3):
            g.sensitivities(1, to_host=False)
            gj = g.jhw_r(wgt, r)
        t_exp = g.stage_times()
        g.set_timing(False)
        adj_ms = sum(v[0] for v in t_adj.values()) / 3
        exp_ms = sum(v[0] for v in t_exp.values()) / 3
        result["adjoint"] = {
            "adjoint_gradient_ms": adj_ms,
            "explicit_j_ms": exp_ms,
            "explicit_over_adjoint": exp_ms / adj_ms,
            "forward_ms": ms_per_step,
            "rel_diff": float(np.max(np.abs(ga - gj)) / np.max(np.abs(gj))),
            "kernels_ms": {**{k: v[0] / v[1] for k, v in t_adj.items()},
                           **{k: v[0] / v[1] for k, v in t_exp.items()}},
        }
        log(f"adjoint {adj_ms:.3f} ms, explicit J {exp_ms:.3f} ms")

        if args.recompute_passes > 0:
            # the same forward + adjoint gradient holding a pass of
            # wavenumbers' potentials at a time, against the stored mode
            # above: wall time and device bytes the handle holds (free device
            # memory before the handle and after its adjoint; the stored
            # handle also holds its explicit J)
            torch.cuda.synchronize()
            stored_bytes = free_start - torch.cuda.mem_get_info()[0]
            free0 = torch.cuda.mem_get_info()[0]
            gr = crt.CRT(m)
            gr.rwaven(w.strnr, w.vnr)
            if args.k20:
                gr.set_wavenumbers(*k20_rule(gr))
            gr.precal()
            gr.set_recompute(args.recompute_passes)
            gr.set_device_budget(int(args.recompute_budget * (1 << 30)))

            def fwd_adj(h):
                def fwd():
                    rc = h.lib.crt_forward_dev(h.h, sigma_dev.data_ptr())
                    if rc:
                        raise crt.CRTError("forward_dev", rc)
                    rc = h.lib.crt_voltages_dev(h.h, N, st.ctypes.data, vn.ctypes.data, 0, volt.data_ptr(),
                                                sigmaa.data_ptr())
                    if rc:
                        raise crt.CRTError("voltages_dev", rc)
                fwd()
                h.adjoint_grad(wgt, r, mode=1)  # the pass buffers exist before the timed calls
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                fwd()
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                gg = h.adjoint_grad(wgt, r, mode=1)
                t2 = time.perf_counter()
                return (t1 - t0) * 1e3, (t2 - t1) * 1e3, gg

            rf_ms, ra_ms, gr_grad = fwd_adj(gr)
            torch.cuda.synchronize()
            rec_bytes = free0 - torch.cuda.mem_get_info()[0]
            sf_ms, sa_ms, _ = fwd_adj(g)
            result["recompute"] = {
                "wavenumbers_per_pass": args.recompute_passes,
                "factor_budget_bytes": int(args.recompute_budget * (1 << 30)),
                "forward_ms": rf_ms, "adjoint_gradient_ms": ra_ms,
                "stored_forward_ms": sf_ms, "stored_adjoint_gradient_ms": sa_ms,
                "device_bytes": rec_bytes, "stored_device_bytes": stored_bytes,
                "potential_bytes_per_copy": 16 * min(args.recompute_passes, K) * E * n,
                "stored_potential_bytes_per_copy": 16 * K * E * n,
                "rel_diff": float(np.max(np.abs(gr_grad - ga)) / np.max(np.abs(ga))),
            }
            log(f"recompute passes of {args.recompute_passes}: forward {rf_ms:.2f} ms, adjoint {ra_ms:.2f} ms, "
                f"{rec_bytes / 1e9:.2f} GB held (stored {stored_bytes / 1e9:.2f} GB)")
            gr.close()

    if not args.skip_dc:
        # the DC twin (real-entry scaled rows, factor and substitutions) on the
        # same model against the complex forward: wall time per forward (host
        # sigma in, nothing copied out) and HIP-event stage times
        def per_forward(fn, reps=3):
            fn(w.sigma)
            torch.cuda.synchronize()
            g.set_timing(True)
            t0 = time.perf_counter()
            for _ in range(reps):
                fn(w.sigma)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) * 1e3 / reps
            stages = {k: v[0] / reps for k, v in g.stage_times().items()}
            g.set_timing(False)
            return dt, stages

        dc_ms, dc_st = per_forward(g.forward_dc)
        cx_ms, cx_st = per_forward(g.forward)
        result["dc"] = {"forward_ms": dc_ms, "complex_forward_ms": cx_ms, "dc_over_complex": dc_ms / cx_ms,
                        "stage_ms": dc_st, "complex_stage_ms": cx_st}
        log(f"DC forward {dc_ms:.2f} ms, complex forward {cx_ms:.2f} ms ({dc_ms / cx_ms:.3f})")

    if not args.no_cpu_baseline and not (args.dirichlet or args.tri or args.k20):
        log("cpu baseline sample")
        result["cpu_baseline"] = cpu_baseline(w, nsys, K, args.cpu_sample)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
