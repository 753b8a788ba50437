"""Benchmark of the exact-diagonalization hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload qed4x3|qed4x2] [--no-cpu]
                    [--serve-requests 512] [--serve-rate 20] [--plain-sample 32] [--no-serve]
    python bench.py --workload su2c3 [--sites 20] [--dyn-steps 200]
    python bench.py --workload cpu_qed4x3
    python bench.py --workload momentum
    python bench.py --workload matfree
    python bench.py --workload serving_graphs
    python bench.py --workload realgauge

Workload (default): 2D U(1) QED, 4x3 periodic lattice, staggered fermions,
E in {-1,0,1}, g=1, m=0.5, Gauss-law + zero-charge sector (N = 28,904,615).
The matrix is built on the GPU (sector enumeration + CSR assembly), then one
"step" is one Lanczos matvec y = H x (complex128 CSR SpMV) over the whole
sector, inputs resident in HBM.  `value` is the SpMV's algorithmic bandwidth

    B = nnz*(16 + 4) + (N+1)*8 + N*(16 + 16)   bytes per launch

divided by the mean launch time (HIP events around K launches on the stream
the kernel runs on).  The JSON line also reports the build times, the FP64
and mixed-precision lowest-eigenvalue solves, the FP32 SpMV rate and the
CPU baseline (the C/OpenMP restatement of the reference's csr_spmv timed on
the host cores over the same matrix).

Serving leg (the "served GS solves/s + p99 latency vs plain FP64" half of the
metric): Z3 pure gauge theory on 3x3, 4x3 and 4x4 periodic lattices
(zero-flux sectors: 6,561 / 177,147 / 14,348,907 states), H = -J sum(U_plaq +
h.c.) - h sum(E + E^dag), J = 1, h ~ U[0.3, 3]; a stream of 512 requests with
Poisson arrivals (rate 20/s, seed 1), lattice drawn uniformly, each asking for
the lowest 4 eigenvalues at tol 1e-8.  Served: resident bases and
dictionary-coded multi-term matrices (assembled once per lattice), one
continuous-batching server per lattice on its own stream, admission by device
memory footprint, warm starts from cached Ritz blocks.  The whole stream is
served (latency, and capacity with every request present at t = 0).  Plain:
one request at a time, each re-assembling H(J, h) and running the FP64 block
Lanczos cold.  Served and plain are compared on the same requests: every
stride-th request of the stream (--plain-sample of them, 64 by default, so the
lattice mix is the stream's), at their stream arrival times (solves/s, p50,
p99) and all present at t = 0 (capacity).  |dE| of the served against the
plain eigenvalues is reported over that sample.

Host-core side of the default line: the C/OpenMP restatement of csr_spmv
(-march=native, built on the host) over the same matrix on all host threads
and on one thread, with the host CPU model; and the full QED 4x3 build +
lowest eigenvalue on the host cores beside the GPU's (--no-host-e0 skips it).

Workload su2c3 (dynamics): 1D SU(2), spin-1/2 rishons with staggered
two-colour matter, 20 sites, OBC, N_tot = 20, g = 1, m = 1 (N = 377,379,369
states).  Quench from the bare vacuum ("V"); exp(-iH dt) psi by Krylov steps
with a 30-vector Lanczos basis, dt = 0.05 to t = 10 (200 steps), once with the
FP64 products and once with complex64 gathered vectors (FP64 values, sums and
recurrence).  The matrix is dictionary-coded and its value array released
(5 bytes per nonzero) so that the Krylov basis fits in HBM next to it.  The
line reports the build times, seconds per step for both precisions and, per
step, the site occupations and the electric energy of both runs with their
difference.

Workload realgauge: QED 4x3 PBC solved in the real gauge (D^H H D with
D = diag(i^(sum_j a_j N_j)), float64 values) beside the complex FP64 solve: the
real matrix's assembly, its SpMV on real vectors, and the FP64 and mixed
lowest-eigenvalue solves with their |dE0| against the complex one.

Workload cpu_qed4x3: the full build + lowest-eigenvalue path of QED 4x3 PBC on the
host cores (oracle/refcpu: C/OpenMP sector enumeration and two-pass triplets,
SciPy coalescing into CSR, eigsh on the C SpMV), beside the same path on the
GPU, for the build+E0 ratio.
"""
import argparse
import json
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "ed-lgt_amd"):
    sys.path.insert(0, str(p))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
WORKLOADS = {
    "qed4x3": dict(lvals=[4, 3], g=1.0, m=0.5, desc="2D U(1) QED 4x3 PBC, staggered fermions, E in {-1,0,1}, g=1, m=0.5"),
    "qed4x2": dict(lvals=[4, 2], g=1.0, m=0.5, desc="2D U(1) QED 4x2 PBC, staggered fermions, E in {-1,0,1}, g=1, m=0.5"),
}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def warm_up():
    """One tiny build + solve (QED 2x2 PBC) so that the one-time loading of
    the library's GPU code objects is not timed: build + E0 is quoted warm,
    the reference's own benchmark convention (its README times the second
    run)."""
    import torch
    from edlgt_amd import QED_Model
    m = QED_Model(spin=1, pure_theory=False, lvals=[2, 2], has_obc=[False, False])
    m.build_Hamiltonian(1.0, 0.5)
    m.diagonalize_Hamiltonian(1, "sparse")
    torch.cuda.synchronize()


def spmv_bytes(n, nnz, s_v=16, s_x=16):
    return nnz * (s_v + 4) + (n + 1) * 8 + n * (s_x + 16)


STREAM_VARIANTS = [(512, 2048, 4), (256, 1024, 4), (512, 4096, 8), (1024, 4096, 4), (256, 2048, 8)]


def launch_desc(tune):
    """Autotuner result as a readable launch shape."""
    g, grid = int(tune[0]), int(tune[1])
    if g == 0 and grid < 0:
        bs, tile, u = STREAM_VARIANTS[-grid - 1]
        return {"kernel": "csr_stream", "threads": bs, "tile_nnz": tile, "gathers_in_flight": u}
    if g == 0:
        return {"kernel": "csr_stream", "threads": 512, "tile_nnz": 2048, "gathers_in_flight": 4}
    return {"kernel": "vector_csr", "lanes_per_row": g, "workgroups": grid}


def time_launches(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(steps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / steps


def real_gauge_leg(model, N, nnz, e64, args):
    """QED with staggered matter solved in the real gauge: D^H H D (float64
    values) built from the same term program, its SpMV timed on real vectors,
    then the FP64 and mixed lowest-eigenvalue solves (eigenvectors mapped back
    through D).  The headline SpMV line stays the complex128 matrix's."""
    import torch
    from edlgt_amd import PREC_MIXED
    progress("real-gauge build and solves")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    rg = model.real_gauge_matrix()
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    if rg is None:
        return None
    Hr = rg[0]
    tune = Hr.autotune(n_rhs=1)
    Hr.set_launch(*tune[:2])
    gen = torch.Generator(device="cuda").manual_seed(0)
    xr = torch.randn(N, dtype=torch.float64, device="cuda", generator=gen)
    yr = torch.empty_like(xr)
    ms = time_launches(lambda: Hr.matvec(xr, yr), args.steps, args.warmup)
    B = spmv_bytes(N, nnz, s_v=8, s_x=8)
    out = dict(assembly_s=round(t_build, 3), spmv_ms=round(ms, 4), spmv_alg_bytes=B,
               spmv_alg_GBs=round(B / (ms * 1e-3) / 1e9, 1),
               spmv_frac=round(B / (ms * 1e-3) / HBM_PEAK_GBS / 1e9, 4))
    del xr, yr
    for name, prec in (("fp64", None), ("mixed", PREC_MIXED)):
        kw = {} if prec is None else dict(precision=prec)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model.diagonalize_Hamiltonian(1, "sparse", real_gauge=True, **kw)
        torch.cuda.synchronize()
        out[f"lanczos_{name}_s"] = round(time.perf_counter() - t0, 3)
        out[f"lanczos_{name}_matvecs"] = [int(v) for v in model.H.matvecs]
        out[f"dE0_{name}_vs_complex_fp64"] = abs(float(model.H.energies[0]) - e64)
    model.H.Npsi = model.H.GSpsi = None
    model._rg = None
    torch.cuda.empty_cache()
    return out


def cpu_baseline(H, x_dev, max_seconds=12.0, one_thread_seconds=8.0):
    """oracle/refcpu csr_spmv (row-parallel, complex128; built -march=native on
    this host) on the host cores over the same CSR, then on one thread;
    returns GB/s by the same algorithmic-bytes formula, with the host's CPU."""
    from oracle import refcpu
    native = refcpu.use_native()
    rp, col, val = H.to_host()
    x = x_dev.cpu().numpy()
    y = np.empty_like(x)
    B = spmv_bytes(H.n, H.nnz)

    def timed(limit, max_reps):
        refcpu.spmv(rp, col, val, x, y)  # warm
        reps, t0 = 0, time.perf_counter()
        while True:
            refcpu.spmv(rp, col, val, x, y)
            reps += 1
            el = time.perf_counter() - t0
            if el > limit or reps >= max_reps:
                return reps, el / reps

    threads = refcpu.threads()
    reps, t = timed(max_seconds, 50)
    refcpu.set_threads(1)
    reps1, t1 = timed(one_thread_seconds, 5)
    refcpu.set_threads(threads)
    model, sockets, cores, cpus = refcpu.host_cpu()
    return dict(value=round(B / t / 1e9, 2), unit="GB/s", cores=threads, kind="port",
                sample=f"{reps} full complex128 CSR SpMV passes over the same matrix "
                       f"(N={H.n}, nnz={H.nnz}), {t*1e3:.1f} ms each, {threads} threads; "
                       f"{reps1} passes on 1 thread, {t1*1e3:.0f} ms each",
                one_thread_GBs=round(B / t1 / 1e9, 2),
                host=dict(model=model, sockets=sockets, physical_cores=cores, logical_cpus=cpus),
                build="gcc -O3 -fopenmp " + ("-march=native (built on this host)" if native else
                                             "-march=x86-64-v3 (no compiler on this host)"))


def host_build_e0():
    """The full QED 4x3 build + lowest-eigenvalue path on the host cores
    (oracle/refcpu: C/OpenMP sector and two-pass triplets, SciPy coalescing,
    eigsh on the C SpMV), timed once."""
    from oracle import refcpu
    t0 = time.perf_counter()
    e_cpu, N, nnz, T = refcpu.qed_build_and_e0([4, 3], [False, False], 1.0, 0.5, log=progress)
    return dict(build_plus_E0_s=round(time.perf_counter() - t0, 2), threads=refcpu.threads(), E0=e_cpu,
                phases_s={k: round(v, 2) for k, v in T.items()}, sector_dim=N, nnz=nnz)


def cpu_build_e0(args):
    """QED 4x3 PBC (N = 28,904,615) build + E0: host cores (refcpu restatement) vs GPU."""
    import torch
    from edlgt_amd import QED_Model
    from oracle import refcpu
    progress("GPU build + E0")
    warm_up()
    t0 = time.perf_counter()
    model = QED_Model(spin=1, pure_theory=False, lvals=[4, 3], has_obc=[False, False])
    model.build_Hamiltonian(1.0, 0.5)
    model.diagonalize_Hamiltonian(1, "sparse")
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t0
    e_gpu = float(model.H.energies[0])
    del model
    torch.cuda.empty_cache()
    progress(f"GPU {t_gpu:.2f}s; CPU path on {refcpu.threads()} threads")
    t0 = time.perf_counter()
    e_cpu, N, nnz, T = refcpu.qed_build_and_e0([4, 3], [False, False], 1.0, 0.5, log=progress)
    t_cpu = time.perf_counter() - t0
    out = {"metric": "QED 4x3 PBC build + lowest eigenvalue, wall time", "value": round(t_gpu, 3), "unit": "s",
           "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": round(1e3 * t_gpu, 1),
           "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "c128",
           "data": "synthetic: matrix built from the model", "config": {"workload": WORKLOADS["qed4x3"]["desc"],
                                                                      "sector_dim": N, "nnz": nnz},
           "gpu": {"build_plus_E0_s": round(t_gpu, 3), "E0": e_gpu},
           "cpu_baseline": {"kind": "port", "cores": refcpu.threads(), "build_plus_E0_s": round(t_cpu, 2),
                            "phases_s": {k: round(v, 2) for k, v in T.items()}, "E0": e_cpu,
                            "sample": "full QED 4x3 path once: sector, triplets, SciPy CSR, eigsh(k=1, SA)"},
           "speedup": round(t_cpu / t_gpu, 1), "dE0_gpu_vs_cpu": abs(e_gpu - e_cpu)}
    print(json.dumps(out), flush=True)


def su2_dynamics(args):
    """1D SU(2) quench dynamics at full size (see the module docstring)."""
    import torch
    from edlgt_amd import PREC_MIXED, SU2_Model
    Ls = args.sites
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mod = SU2_Model(spin=0.5, pure_theory=False, sectors=[Ls], lvals=[Ls], has_obc=[True])
    torch.cuda.synchronize()
    t_sector = time.perf_counter() - t0
    # stored CSR, then coded here so that assembly and coding are timed apart
    # (the model's automatic choice makes the same coded matrix in one step)
    mod.matrix_free = False
    t0 = time.perf_counter()
    H = mod.build_Hamiltonian(1.0, 1.0)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    t0 = time.perf_counter()
    n_dict = H.compress()
    H.release_values()
    torch.cuda.synchronize()
    t_code = time.perf_counter() - t0
    x = torch.randn(H.n, dtype=torch.complex128, device="cuda")
    y = torch.empty_like(x)
    g, grid, ms_tune = H.autotune()
    ms_spmv = time_launches(lambda: H.matvec(x, y), 10, 2)
    del x, y
    torch.cuda.empty_cache()
    names = ["T2_px"] + [f"N_{l}" for l in ("tot", "single", "pair", "zero")]
    mod.get_observables(names)
    psi0 = mod.get_qmb_state_from_configs([mod.overlap_QMB_state("V")])
    tl = np.arange(args.dyn_steps + 1) * 0.05
    runs = {}
    chunk = 20
    for prec, name in ((0, "fp64"), (PREC_MIXED, "mixed")):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        psi, margs, errs = psi0, [], []
        for c0 in range(0, args.dyn_steps, chunk):
            n = min(chunk, args.dyn_steps - c0)
            # the time line of a call is measured from the state passed in
            mod.time_evolution_Hamiltonian(psi, tl[:n + 1], m=30, reorth=False, store_states=False,
                                           precision=prec)
            psi = mod.H.psi_final
            margs.extend(list(mod.H.marginals[(0 if c0 == 0 else 1):]))
            errs.extend(list(mod.H.krylov_err))
            print(f"{name}: {c0 + n}/{args.dyn_steps} steps, {time.perf_counter() - t0:.1f}s", file=sys.stderr,
                  flush=True)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        mod.H.krylov_err = np.asarray(errs)
        obs = {"E2": [], "N_tot": [], "N_single": [], "N_pair": [], "N_zero": []}
        even = mod.lattice.even_mask()
        for k in range(len(tl)):
            marg = margs[k]
            v = {n: mod.local_expvals(marg, n) for n in names}
            obs["E2"].append(float(v["T2_px"][:-1].mean()))
            obs["N_single"].append(float(v["N_single"].mean()))
            npair = 0.5 * v["N_pair"][even].mean() + 0.5 * v["N_zero"][~even].mean()
            obs["N_pair"].append(float(npair))
            obs["N_zero"].append(float(0.5 * v["N_zero"][even].mean() + 0.5 * v["N_pair"][~even].mean()))
            obs["N_tot"].append(float(v["N_single"].mean() + 2 * npair))
        runs[name] = dict(seconds=round(el, 2), s_per_step=round(el / args.dyn_steps, 4),
                          krylov_err_max=float(mod.H.krylov_err.max()), obs=obs)
    err = {k: [abs(a - b) for a, b in zip(runs["mixed"]["obs"][k], runs["fp64"]["obs"][k])]
           for k in runs["fp64"]["obs"]}
    out = {
        "metric": "Krylov exp(-iH dt) step time (m=30), 1D SU(2) quench",
        "value": runs["fp64"]["s_per_step"], "unit": "s/step", "n_gpus": 1, "steps": args.dyn_steps,
        "warmup": 0, "ms_per_step": round(1e3 * runs["fp64"]["s_per_step"], 2), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "c128",
        "data": "synthetic: bare-vacuum quench built on the GPU from the model",
        "config": {"workload": f"1D SU(2) j=1/2 with matter, {Ls} sites OBC, N_tot={Ls}, g=1, m=1, "
                               "m_krylov=30, dt=0.05", "sector_dim": H.n, "nnz": H.nnz},
        "build": dict(sector_s=round(t_sector, 3), assembly_s=round(t_asm, 3), code_s=round(t_code, 3),
                      n_dict=n_dict, coded_spmv_ms=round(ms_spmv, 3), coded_variant=-grid - 1),
        "fp64": {k: v for k, v in runs["fp64"].items() if k != "obs"},
        "mixed": {k: v for k, v in runs["mixed"].items() if k != "obs"},
        "max_abs_err_mixed_vs_fp64": {k: max(v) for k, v in err.items()},
        "per_step": {"t": [round(float(t), 4) for t in tl], "fp64": runs["fp64"]["obs"],
                     "abs_err_mixed": err},
    }
    print(json.dumps(out), flush=True)


def coupling_sweep():
    """QED 4x2 PBC with matter (N = 140,647), m = 0.5, g in {0.5, 1, 2}, lowest 4
    eigenvalues: one build + FP64 Lanczos per coupling (plain) vs one
    multi-term assembly solved for all couplings by one server (batched), and
    the same server on the real-gauge multi-term matrix."""
    import torch
    from edlgt_amd import QED_Model, Server
    gs = [0.5, 1.0, 2.0]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    mod = QED_Model(spin=1, pure_theory=False, lvals=[4, 2], has_obc=[False, False])
    torch.cuda.synchronize()
    t_sector = time.perf_counter() - t0
    plain, t0 = {}, time.perf_counter()
    for g in gs:
        H = mod.build_Hamiltonian(g, 0.5)
        ev, _, _, _ = H.lanczos(4, tol=1e-10, want_vectors=False)
        plain[g] = ev
    torch.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    t0 = time.perf_counter()
    H3 = mod.build_multi_term(0.5)
    res = Server(H3, n_eigs=4, tol=1e-10, max_basis=64, max_slots=3).solve_all(
        [(i, mod.multi_term_coefficients(g), 7 + i) for i, g in enumerate(gs)])
    torch.cuda.synchronize()
    t_batch = time.perf_counter() - t0
    # the same sweep on the real-gauge multi-term matrix (float64 values and vectors)
    t0 = time.perf_counter()
    H3r = mod.build_multi_term(0.5, real_gauge=True)
    res_r = Server(H3r, n_eigs=4, tol=1e-10, max_basis=64, max_slots=3).solve_all(
        [(i, mod.multi_term_coefficients(g), 7 + i) for i, g in enumerate(gs)])
    torch.cuda.synchronize()
    t_batch_r = time.perf_counter() - t0
    del H3r
    # the same couplings squared (the configuration's g^2/2 E^2 and 1/(2 g^2)
    # plaquette formula in the API's g), served on the same multi-term matrix
    gsq = [g * g for g in gs]
    res2 = Server(H3, n_eigs=4, tol=1e-10, max_basis=64, max_slots=3).solve_all(
        [(i, mod.multi_term_coefficients(g), 17 + i) for i, g in enumerate(gsq)])
    return dict(workload="QED 4x2 PBC with matter, m=0.5, g in {0.5,1,2}, lowest 4 eigenvalues, tol 1e-10",
                sector_dim=mod.sector_dim, sector_s=round(t_sector, 3), plain_s=round(t_plain, 3),
                batched_s=round(t_batch, 3), batched_real_gauge_s=round(t_batch_r, 3),
                max_dE_real_gauge_vs_plain=max(float(np.max(np.abs(res_r[i][0] - plain[g])))
                                               for i, g in enumerate(gs)),
                E0_per_site={str(g): float(plain[g][0] / 8) for g in gs},
                E0_per_site_g_squared={str(g): float(res2[i][0][0] / 8) for i, g in enumerate(gsq)},
                max_dE_batched_vs_plain=max(float(np.max(np.abs(res[i][0] - plain[g]))) for i, g in enumerate(gs)))


def serve_leg(args):
    """Served vs plain lowest-eigenvalue solves on the Z3 request stream."""
    import torch
    from edlgt_amd import Zn_Model
    from edlgt_amd.serving import SolveService, latency_summary, poisson_stream, serve_plain
    lats = {"3x3": [3, 3], "4x3": [4, 3], "4x4": [4, 4]}
    models, csrs = {}, {}
    t0 = time.perf_counter()
    for name, lv in lats.items():
        models[name] = Zn_Model(3, lvals=lv, has_obc=[False, False], flux_sector=(0, 0))
        csrs[name] = models[name].build_multi_term()
        csrs[name].compress()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t0
    reqs = poisson_stream(args.serve_requests, args.serve_rate, list(lats),
                          lambda rng: [1.0, rng.uniform(0.3, 3.0)], seed=1)
    svc = SolveService(csrs, n_eigs=4, tol=1e-8, max_basis=args.serve_basis, max_slots=16, warm_cache=64)
    res = svc.run(reqs)
    served = latency_summary(res, svc.wall)
    served["mean_matvecs"] = {p: round(float(np.mean([r["matvecs"] for r in res.values() if r["problem"] == p])), 1)
                              for p in lats}
    served["all_converged"] = all(r["converged"] for r in res.values())
    served["peak_device_GB"] = round(svc.peak_bytes / 1e9, 2)
    del svc
    torch.cuda.empty_cache()
    # capacity: the same requests all present at t = 0 (fresh servers and
    # caches), so the rate is the servers' own, not the arrival rate
    progress("served stream done; capacity run")
    burst = [dict(r, arrival=0.0) for r in reqs]
    svc = SolveService(csrs, n_eigs=4, tol=1e-8, max_basis=args.serve_basis, max_slots=16, warm_cache=64)
    bres = svc.run(burst)
    capacity = dict(requests=len(bres), wall_s=round(svc.wall, 3), solves_per_s=round(len(bres) / svc.wall, 2),
                    all_converged=all(r["converged"] for r in bres.values()))
    del svc
    torch.cuda.empty_cache()

    def solve_one(r):
        H = models[r["problem"]].build_Hamiltonian(r["coef"][0], r["coef"][1])
        ev, _, _, _ = H.lanczos(4, tol=1e-8, max_basis=args.plain_basis, want_vectors=False)
        return ev

    # the same requests through both paths: every stride-th request of the
    # whole stream (so the three lattice sizes appear as in the stream), at
    # their stream arrival times, and all present at t = 0 (capacity)
    ordered = sorted(reqs, key=lambda r: r["arrival"])
    stride = max(1, len(ordered) // args.plain_sample)
    sample = ordered[::stride][:args.plain_sample]
    progress(f"served {served['requests']} requests in {served['wall_s']:.1f}s; "
             f"same-sample comparison on {len(sample)} requests")
    svc = SolveService(csrs, n_eigs=4, tol=1e-8, max_basis=args.serve_basis, max_slots=16, warm_cache=64)
    sres = svc.run(sample)
    served_s = latency_summary(sres, svc.wall)
    del svc
    svc = SolveService(csrs, n_eigs=4, tol=1e-8, max_basis=args.serve_basis, max_slots=16, warm_cache=64)
    cres = svc.run([dict(r, arrival=0.0) for r in sample])
    served_cap = len(cres) / svc.wall
    del svc
    torch.cuda.empty_cache()
    pres = serve_plain(solve_one, sample)
    plain = latency_summary(pres, serve_plain.wall)
    busy = sum(r["done"] - r["admitted"] for r in pres.values())
    plain_cap = len(pres) / busy
    d = [float(np.max(np.abs(sres[i]["evals"] - pres[i]["evals"]))) for i in pres]
    d0 = [float(abs(sres[i]["evals"][0] - pres[i]["evals"][0])) for i in pres]
    mix = {p: sum(1 for r in sample if r["problem"] == p) for p in lats}
    same = dict(
        requests=len(sample), stride=stride, lattice_mix=mix,
        arrival_rate_per_s=round(args.serve_rate / stride, 3),
        served=dict(solves_per_s=round(served_s["solves_per_s"], 3), p50_ms=round(served_s["p50_ms"], 1),
                    p99_ms=round(served_s["p99_ms"], 1), mean_ms=round(served_s["mean_ms"], 1)),
        plain=dict(solves_per_s=round(plain["solves_per_s"], 3), p50_ms=round(plain["p50_ms"], 1),
                   p99_ms=round(plain["p99_ms"], 1), mean_ms=round(plain["mean_ms"], 1)),
        capacity_solves_per_s=dict(served=round(served_cap, 2), plain=round(plain_cap, 3),
                                   ratio=round(served_cap / plain_cap, 2)),
        p50_ratio=round(plain["p50_ms"] / served_s["p50_ms"], 2),
        p99_ratio=round(plain["p99_ms"] / served_s["p99_ms"], 2),
        max_dE_served_vs_plain=max(d), max_dE0_served_vs_plain=max(d0))
    return dict(
        workload="Z3 pure gauge 3x3/4x3/4x4 PBC zero-flux sectors (6561/177147/14348907 states), J=1, "
                 "h~U[0.3,3], lowest 4 eigenvalues at tol 1e-8",
        stream=dict(requests=args.serve_requests, poisson_rate_per_s=args.serve_rate, seed=1),
        lanczos_basis=dict(served=args.serve_basis, plain=args.plain_basis),
        setup_s=round(t_setup, 3), served=served, served_capacity=capacity,
        same_sample=same,
        speedup_capacity_same_sample=same["capacity_solves_per_s"]["ratio"],
        speedup_capacity_vs_plain_service_rate=round(capacity["solves_per_s"] / plain_cap, 2))


def serving_graphs_ab(args):
    """The served Z3 stream with each step's fused-SpMM chain replayed from
    HIP graphs against the same chain launched eagerly
    (EDLGT_SERVER_GRAPHS=1 / 0, read when a server is created), two
    repetitions: per-lattice capacity, whole-stream capacity, the 64-request
    same-sample burst and the stream's p50 / p99."""
    import os
    import torch
    from edlgt_amd import Zn_Model
    from edlgt_amd.serving import SolveService, latency_summary, poisson_stream
    lats = {"3x3": [3, 3], "4x3": [4, 3], "4x4": [4, 4]}
    csrs = {}
    for name, lv in lats.items():
        csrs[name] = Zn_Model(3, lvals=lv, has_obc=[False, False], flux_sector=(0, 0)).build_multi_term()
        csrs[name].compress()
    reqs = poisson_stream(args.serve_requests, args.serve_rate, list(lats),
                          lambda rng: [1.0, rng.uniform(0.3, 3.0)], seed=1)
    ordered = sorted(reqs, key=lambda r: r["arrival"])
    sample = ordered[::max(1, len(ordered) // args.plain_sample)][:args.plain_sample]

    def service(problems):
        return SolveService(problems, n_eigs=4, tol=1e-8, max_basis=args.serve_basis, max_slots=16, warm_cache=64)

    runs = []
    for rep in range(2):
        for mode in ("1", "0"):
            os.environ["EDLGT_SERVER_GRAPHS"] = mode
            row = dict(graphs=mode == "1", rep=rep)
            for p in lats:
                sub = [dict(r, arrival=0.0) for r in reqs if r["problem"] == p]
                svc = service({p: csrs[p]})
                svc.run(sub)
                row[f"{p}_alone_solves_per_s"] = round(len(sub) / svc.wall, 2)
            svc = service(csrs)
            res = svc.run([dict(r, arrival=0.0) for r in reqs])
            row["capacity_solves_per_s"] = round(len(res) / svc.wall, 2)
            svc = service(csrs)
            res = svc.run([dict(r, arrival=0.0) for r in sample])
            row["sample_burst_solves_per_s"] = round(len(res) / svc.wall, 2)
            svc = service(csrs)
            res = svc.run(reqs)
            m = latency_summary(res, svc.wall)
            row.update(stream_p50_ms=round(m["p50_ms"], 1), stream_p99_ms=round(m["p99_ms"], 1))
            progress(json.dumps(row))
            runs.append(row)
            del svc
            torch.cuda.empty_cache()
    os.environ.pop("EDLGT_SERVER_GRAPHS", None)
    print(json.dumps({"metric": "served Z3 stream, graph-replayed vs eager step chain", "runs": runs}), flush=True)


def momentum_sectors(args):
    """Z3 pure gauge 4x4 PBC zero-flux sector (N = 14,348,907), J = 1, h = 0.7:
    full-sector build + lowest 4 eigenvalues vs the zero-momentum sector of
    the 16 unit translations (translation table, orbit scan, B^H H B) + its
    lowest 4; the ground state lies at k = 0."""
    import torch
    from edlgt_amd import Zn_Model

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        out = fn()
        torch.cuda.synchronize()
        return out, time.perf_counter() - t0

    progress("Z3 4x4 full sector")
    mod, t_sec = timed(lambda: Zn_Model(3, lvals=[4, 4], has_obc=[False, False], flux_sector=(0, 0)))
    H, t_build = timed(lambda: mod.build_Hamiltonian(1.0, 0.7))
    N, nnz = H.n, H.nnz
    (ev, _, _, mv), t_solve = timed(lambda: H.lanczos(4, tol=1e-10, want_vectors=False))
    del H
    mod.H = None
    torch.cuda.empty_cache()
    progress(f"full: N={N} build {t_build:.2f}s solve {t_solve:.2f}s; k=(0,0) sector")
    mod2, _ = timed(lambda: Zn_Model(3, lvals=[4, 4], has_obc=[False, False], flux_sector=(0, 0)))
    _, t_mb = timed(lambda: mod2.set_momentum_sector([1, 1], [0, 0]))
    Hk, t_kbuild = timed(lambda: mod2.build_Hamiltonian(1.0, 0.7))
    (evk, _, _, mvk), t_ksolve = timed(lambda: Hk.lanczos(4, tol=1e-10, want_vectors=False))
    out = {"metric": "momentum sector (k=0) build + lowest 4 eigenvalues, wall time", "value": round(t_mb + t_kbuild + t_ksolve, 3),
           "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic: matrix built from the model",
           "config": {"workload": "Z3 pure gauge 4x4 PBC zero-flux sector, J=1, h=0.7, unit translations",
                      "sector_dim": N, "nnz": nnz, "k_sector_dim": Hk.n, "k_nnz": Hk.nnz},
           "full": {"build_s": round(t_build, 3), "solve_s": round(t_solve, 3), "matvecs": int(np.sum(mv)),
                    "evals": [float(e) for e in ev]},
           "k0": {"projector_s": round(t_mb, 3), "build_s": round(t_kbuild, 3), "solve_s": round(t_ksolve, 3),
                  "matvecs": int(np.sum(mvk)), "evals": [float(e) for e in evk]},
           "dE0_k0_vs_full": abs(float(evk[0] - ev[0]))}
    print(json.dumps(out), flush=True)


def real_gauge_workload(args):
    """QED 4x3 PBC (N = 28,904,615) solved in the real gauge beside the complex
    matrix: the complex FP64 lowest eigenvalue, then the real matrix D^H H D,
    its SpMV on real vectors and its FP64 and mixed solves (real_gauge_leg)."""
    import torch
    from edlgt_amd import QED_Model
    wl = WORKLOADS["qed4x3"]
    warm_up()
    progress("QED 4x3: complex build and FP64 solve")
    model = QED_Model(spin=1, pure_theory=False, lvals=wl["lvals"], has_obc=[False, False])
    H = model.build_Hamiltonian(wl["g"], wl["m"])
    N, nnz = H.n, H.nnz
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model.diagonalize_Hamiltonian(1, "sparse")
    torch.cuda.synchronize()
    t64 = time.perf_counter() - t0
    e64 = float(model.H.energies[0])
    rg = real_gauge_leg(model, N, nnz, e64, args)
    # Krylov exp(-iH dt) steps from a random state, complex matrix and real gauge
    gen = torch.Generator(device="cuda").manual_seed(1)
    psi0 = torch.randn(N, dtype=torch.complex128, device="cuda", generator=gen)
    psi0 /= torch.linalg.vector_norm(psi0)
    tl = np.arange(4) * 0.05
    dyn = {}
    finals = {}
    for rgf in (False, True):
        model.time_evolution_Hamiltonian(psi0, tl[:2], real_gauge=rgf, m=30, store_states=False)  # warm session
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model.time_evolution_Hamiltonian(psi0, tl, real_gauge=rgf, m=30, store_states=False)
        torch.cuda.synchronize()
        key = "real_gauge" if rgf else "complex"
        dyn[f"{key}_s_per_step"] = round((time.perf_counter() - t0) / (len(tl) - 1), 4)
        finals[key] = model.H.psi_final.clone()
        model.H.release_krylov()
    dyn["max_abs_diff_final_state"] = float((finals["complex"] - finals["real_gauge"]).abs().max())
    dyn["steps"] = int(len(tl) - 1)
    dyn["dt"] = 0.05
    rg["krylov_dynamics"] = dyn
    out = {"metric": "real-gauge FP64 lowest-eigenvalue solve time", "value": rg["lanczos_fp64_s"], "unit": "s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic (default_rng(0) start vector)",
           "config": {"workload": wl["desc"], "sector_dim": N, "stored_nnz": nnz},
           "complex_fp64": {"lanczos_s": round(t64, 3), "E0": e64}, "real_gauge": rg}
    print(json.dumps(out), flush=True)


def matrix_free(args):
    """QED 4x3 PBC (N = 28,904,615): the matrix-free operator's product (entries
    generated from the term program in every product) against the stored CSR
    SpMV of the same Hamiltonian, and the lowest eigenvalue through it."""
    import torch
    from edlgt_amd import QED_Model
    progress("QED 4x3: stored and matrix-free operators")
    ms = QED_Model(spin=1, pure_theory=False, lvals=[4, 3], has_obc=[False, False])
    ms.matrix_free = False
    Hs = ms.build_Hamiltonian(1.0, 0.5)
    mf = QED_Model(spin=1, pure_theory=False, lvals=[4, 3], has_obc=[False, False])
    mf.matrix_free = True
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Hf = mf.build_Hamiltonian(1.0, 0.5)
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    N, nnz = Hs.n, Hs.nnz
    gen = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(N, dtype=torch.complex128, device="cuda", generator=gen)
    ys, yf = torch.empty_like(x), torch.empty_like(x)
    ms_s = time_launches(lambda: Hs.matvec(x, ys), 10, 2)
    ms_f = time_launches(lambda: Hf.matvec(x, yf), args.steps if args.steps < 10 else 5, 1)
    rel = float((ys - yf).abs().max() / ys.abs().max())
    del ms.H, Hs
    torch.cuda.empty_cache()
    progress(f"matrix-free product {ms_f:.1f} ms (stored {ms_s:.2f} ms); lowest eigenvalue")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev, _, _, mv = Hf.lanczos(1, tol=1e-10, want_vectors=False)
    torch.cuda.synchronize()
    t_solve = time.perf_counter() - t0
    B = spmv_bytes(N, nnz)
    out = {"metric": "matrix-free product time (CSR-equivalent GB/s beside it)", "value": round(ms_f, 2),
           "unit": "ms", "n_gpus": 1, "steps": 5, "warmup": 1, "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "c128", "data": "synthetic (random complex128 vector, seed 0)",
           "config": {"workload": WORKLOADS["qed4x3"]["desc"], "sector_dim": N, "stored_nnz": nnz},
           "matrix_free": {"build_s": round(t_build, 3), "product_ms": round(ms_f, 2),
                           "csr_equivalent_GBs": round(B / (ms_f * 1e-3) / 1e9, 1),
                           "device_bytes_per_product": N * (8 + 16 + 16) + N * 16,
                           "max_rel_diff_vs_stored": rel, "E0": float(ev[0]), "lanczos_s": round(t_solve, 2),
                           "matvecs": int(np.sum(mv))},
           "stored": {"product_ms": round(ms_s, 3), "csr_bytes": int(nnz * 20 + (N + 1) * 8)}}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="qed4x3", choices=sorted(WORKLOADS) + ["su2c3", "cpu_qed4x3", "momentum", "matfree", "serving_graphs",
                                                                  "realgauge"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host-e0", action="store_true",
                    help="skip the host-core build + lowest eigenvalue of QED 4x3 (about two minutes)")
    ap.add_argument("--no-solve", action="store_true",
                    help="SpMV line only: no Lanczos solves, coupling sweep or serving leg")
    ap.add_argument("--sites", type=int, default=20)
    ap.add_argument("--dyn-steps", type=int, default=200)
    ap.add_argument("--serve-requests", type=int, default=512)
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--plain-sample", type=int, default=64,
                    help="requests in the same-sample served/plain comparison (every stride-th of the stream)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-basis", type=int, default=160,
                    help="Lanczos basis size of the served solves (measured on the stream: p99 "
                         "3.8 s at 128, 3.4 s at 160, 3.7 s at 192)")
    ap.add_argument("--plain-basis", type=int, default=64,
                    help="Lanczos basis size of the plain solves, the best of 20 (eigsh's "
                         "default), 32, 48, 64 and 96 measured on the stream's first requests")
    ap.add_argument("--calibrate", action="store_true",
                    help="also launch plain device copies of a known size (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): no multi-GPU path is "
              "built; one GPU runs every kernel at its full single-shard shape.", file=sys.stderr)
        sys.exit(2)
    if args.workload == "su2c3":
        return su2_dynamics(args)
    if args.workload == "cpu_qed4x3":
        return cpu_build_e0(args)
    if args.workload == "momentum":
        return momentum_sectors(args)
    if args.workload == "matfree":
        return matrix_free(args)
    if args.workload == "realgauge":
        return real_gauge_workload(args)
    if args.workload == "serving_graphs":
        return serving_graphs_ab(args)
    import torch
    from edlgt_amd import QED_Model, PREC_MIXED
    from edlgt_amd import _lib as L

    wl = WORKLOADS[args.workload]
    warm_up()
    progress(f"building {wl['desc']}")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    model = QED_Model(spin=1, pure_theory=False, lvals=wl["lvals"], has_obc=[False, False])
    torch.cuda.synchronize()
    t_sector = time.perf_counter() - t0
    t0 = time.perf_counter()
    H = model.build_Hamiltonian(wl["g"], wl["m"])
    torch.cuda.synchronize()
    t_assembly = time.perf_counter() - t0
    N, nnz = H.n, H.nnz

    progress(f"N={H.n} nnz={H.nnz}: sector {t_sector:.2f}s assembly {t_assembly:.2f}s; autotuning")
    # launch-shape autotuner (lanes per row x workgroups), timed on the device
    tune = H.autotune(n_rhs=1)
    tune32 = H.autotune(n_rhs=1, precision=L.PREC_MIXED)
    H.set_launch(*tune[:2])

    gen = torch.Generator(device="cuda").manual_seed(0)
    x = torch.randn(N, dtype=torch.complex128, device="cuda", generator=gen)
    y = torch.empty_like(x)
    ms = time_launches(lambda: H.matvec(x, y), args.steps, args.warmup)
    B = spmv_bytes(N, nnz)
    gbs = B / (ms * 1e-3) / 1e9

    if args.calibrate:
        # 5 elementwise passes over N complex128: 16*N bytes read, 16*N written each
        for _ in range(5):
            torch.neg(x, out=y)
        torch.cuda.synchronize()

    # mixed precision: FP32 values and FP32 gathered vector, FP64 accumulation
    x32 = x.to(torch.complex64)
    lib = L.load()

    def mixed():  # uses the FP32-tuned launch shape
        H.set_launch(*tune32[:2])
        L.check(lib.edlgt_spmm(H.handle, L.PREC_MIXED, 1, x32.data_ptr(), y.data_ptr(), L.VALUE_C64,
                               L.stream_handle()))

    mixed()
    ms32 = time_launches(mixed, args.steps, args.warmup)
    H.set_launch(*tune[:2])
    gbs32 = spmv_bytes(N, nnz, s_v=8, s_x=8) / (ms32 * 1e-3) / 1e9

    progress(f"SpMV {ms:.3f} ms ({gbs:.0f} GB/s); mixed {ms32:.3f} ms; solving")
    extra = {}
    if not args.no_solve:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        model.diagonalize_Hamiltonian(1, "sparse")
        torch.cuda.synchronize()
        t_fp64 = time.perf_counter() - t0
        e64 = float(model.H.energies[0])
        mv64 = [int(v) for v in model.H.matvecs]
        t0 = time.perf_counter()
        model.diagonalize_Hamiltonian(1, "sparse", precision=PREC_MIXED)
        torch.cuda.synchronize()
        t_mix = time.perf_counter() - t0
        emix = float(model.H.energies[0])
        extra = dict(E0_fp64=e64, E0_per_site=e64 / np.prod(wl["lvals"]), lanczos_fp64_s=round(t_fp64, 3),
                     lanczos_fp64_matvecs=mv64, lanczos_mixed_s=round(t_mix, 3),
                     lanczos_mixed_matvecs=[int(v) for v in model.H.matvecs],
                     dE0_mixed_vs_fp64=abs(emix - e64),
                     build_plus_E0_s=round(t_sector + t_assembly + t_fp64, 3))
        extra["real_gauge"] = real_gauge_leg(model, N, nnz, e64, args)

    cpu = None
    if not args.no_cpu:
        progress("CPU baseline")
        cpu = cpu_baseline(H, x)
        torch.cuda.empty_cache()
    progress("coded SpMV")

    # dictionary-coded values: same products bit for bit, 5 bytes per entry
    y_ref = y.clone()
    H.matvec(x, y_ref)
    t0 = time.perf_counter()
    n_dict = H.compress()
    torch.cuda.synchronize()
    t_code = time.perf_counter() - t0
    ms_coded = time_launches(lambda: H.matvec(x, y), args.steps, args.warmup)
    coded = dict(n_dict=n_dict, code_bytes=H.code_bytes, compress_s=round(t_code, 4),
                 spmv_coded_ms=round(ms_coded, 4), bitwise_equal=bool(torch.equal(y, y_ref)),
                 spmv_coded_alg_GBs=round(spmv_bytes(N, nnz, s_v=H.code_bytes) / (ms_coded * 1e-3) / 1e9, 2))
    if not args.no_solve:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ev, _, _, mvc = H.lanczos(1, tol=1e-12, want_vectors=False)
        torch.cuda.synchronize()
        t_lc = time.perf_counter() - t0
        coded.update(lanczos_coded_s=round(t_lc, 3), lanczos_coded_matvecs=int(mvc[0]),
                     dE0_coded_vs_fp64=float(abs(ev[0] - e64)),
                     build_code_E0_s=round(t_sector + t_assembly + t_code + t_lc, 3))
    del model, H, x, y, y_ref, x32
    torch.cuda.empty_cache()

    serving = sweep = None
    if not args.no_solve:  # the coupling sweep and the served stream are solves too
        progress("QED 4x2 coupling sweep")
        sweep = coupling_sweep()
    if not args.no_serve and not args.no_solve:
        progress("serving leg")
        serving = serve_leg(args)

    if cpu is not None and not args.no_host_e0 and args.workload == "qed4x3" and not args.no_solve:
        progress("host-core build + E0 of the same sector")
        he = host_build_e0()
        he["gpu_build_plus_E0_s"] = extra["build_plus_E0_s"]
        he["speedup"] = round(he["build_plus_E0_s"] / extra["build_plus_E0_s"], 1)
        he["dE0_gpu_vs_host"] = abs(he["E0"] - extra["E0_fp64"])
        cpu["build_plus_E0"] = he

    traffic = None
    tf = ROOT / "profiles" / f"spmv_{args.workload}_traffic.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get("hbm_bytes_per_launch")
    out = {
        "metric": "Lanczos SpMV GB/s (% HBM peak)",
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (random complex128 input vector, seed 0); matrix built on the GPU from the model",
        "config": {"workload": wl["desc"], "sector_dim": N, "nnz": nnz, "step": "one CSR SpMV y=Hx",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "algorithmic_bytes": B},
        "cpu_baseline": cpu,
        "extra": dict(sector_s=round(t_sector, 3), assembly_s=round(t_assembly, 3),
                      spmv_launch=launch_desc(tune), spmv_mixed_launch=launch_desc(tune32),
                      spmv_mixed_ms=round(ms32, 4), spmv_mixed_GBs=round(gbs32, 2),
                      pct_hbm_peak=round(100 * gbs / HBM_PEAK_GBS, 2), coded=coded, **extra),
        "coupling_sweep": sweep,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
