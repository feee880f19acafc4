#!/usr/bin/env python3
"""Benchmark of the MI355X assembly path on the headline workload.

Workload (one "step"): SimpleAssembler::BuildVector + BuildMatrix of the hex8 gradient-damage
integrand (GradientDamage<3>, 3 displacement + 1 nonlocal-strain dof per node, 2x2x2 Gauss) on
the structured 100 x 100 x 100 mesh of a 100 mm cube: 1,000,000 cells, 4,121,204 dofs,
436,334,416 nonzeros. E = 30000, nu = 0.2, k = 10, kappa0 = 1e-4, beta = 100, alpha = 0.99,
c = 4 mm^2; u ~ U[-5e-3, 5e-3] mm, eeq ~ U[0, 5e-4], kappa ~ U[0, 3e-4] per IP (seeded).
Inputs and outputs are resident in HBM; the timed region is K steps, each the pre-pass over the
(cell, IP) states and the node kernel that writes R and K.

Prints one JSON line (rank 0). `value` is cells/s; `roofline` prices the assembly kernel's
algorithmic bytes per launch against 8 TB/s HBM; `fp64` prices the measured FP64 flops of the step
against 78.6 TFLOP/s; `cpu_baseline` times the CPU restatement of
NuTo's algorithm (the oracle, OpenMP, triplets + setFromTriplets) on a bounded sample.
"""
import argparse
import json
import os
import sys
import time

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
HERE = os.path.dirname(os.path.abspath(__file__))
# HBM traffic of the assembly kernel per launch, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes
# scaled by the same counters on a copy kernel of known size (written by --copy-calibration runs)
TRAFFIC_FILE = os.path.join(HERE, "profiles", "kGdmHex8Node_hex8_gdm_100cube_traffic.txt")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def algorithmic_bytes(nnz, n_rows, n_nodes, n_cells, nip, nn_cell):
    """Compulsory bytes of one R+K assembly: K and R written once; coordinates, u, eeq, history and
    connectivity read once (index arrays of the pattern are not counted)."""
    return (8 * nnz + 8 * n_rows + 8 * (3 + 3 + 1) * n_nodes + 8 * n_cells * nip + 4 * n_cells * nn_cell)


def measured_counter(n, key):
    """A per-step figure from the committed counter summary (None when absent or for another size)."""
    if n != 100 or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        if line.startswith(key + ":"):
            return float(line.split(":")[1])
    return None


def measured_traffic(n):
    """HBM bytes per step (FETCH_SIZE + WRITE_SIZE of both kernels, calibrated on a copy kernel)."""
    return measured_counter(n, "traffic_bytes_per_launch")


def copy_calibration(mib):
    """A plain device copy of known size (reads and writes `mib` MiB each), the reference point for
    scaling FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": 8 * n, "copy_calibration_bytes_written": 8 * n}), flush=True)


def copy_bandwidth(mib=2048, reps=10):
    """Measured device-to-device copy bandwidth (read + write bytes / time, median of reps), the
    practical HBM ceiling to set beside the 8 TB/s peak."""
    import numpy as np
    import torch

    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for a, z in evs:
        a.record()
        dst.copy_(src)
        z.record()
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(z) for a, z in evs]))
    del src, dst
    torch.cuda.empty_cache()
    return 2.0 * 8 * n / (ms * 1e-3) / 1e9


def solve_bench(ctx, X, C, ranks, nn, K, R, iters, length):
    """ConstrainedSystemSolver::SolveTrialState on the assembled system: bottom face clamped, the top
    face's vertical displacement stepped by 1e-3 mm, eeq free; BiCGSTAB with the diagonal
    preconditioner for a fixed number of iterations (the tolerance is out of reach), timed after one
    untimed solve that builds the constraint matrix. A solve also allocates its Krylov vectors and
    forms the preconditioner; the per-iteration time is the difference of a solve with `iters`
    iterations and one with 2, divided by iters - 2 (each the fastest of 3 solves: the allocations of
    a solve take a variable time)."""
    import numpy as np
    import torch

    import nuto_amd as N

    xs = np.zeros((nn, 3))
    xs[ranks.ravel()] = X[C.ravel()]
    bottom = np.nonzero(np.abs(xs[:, 2]) < 1e-9)[0]
    top = np.nonzero(np.abs(xs[:, 2] - length) < 1e-9)[0]
    eqs = [(int(3 * r + c), []) for r in bottom for c in range(3)] + [(int(3 * r + 2), []) for r in top]
    ctx.set_constraints(0, eqs)
    rhs = np.zeros(len(eqs))
    rhs[3 * len(bottom):] = 1e-3
    rhs_d = torch.from_numpy(rhs).cuda()
    x = torch.empty(ctx.n_rows, dtype=torch.float64, device="cuda")
    ctx.solve_trial_state(K, R, None, rhs_d, method=N.SOLVER_BICGSTAB, tol=1e-300, max_iter=2, out=x)
    torch.cuda.synchronize()

    def timed(n_it):
        t0 = time.perf_counter()
        _, inf = ctx.solve_trial_state(K, R, None, rhs_d, method=N.SOLVER_BICGSTAB, tol=1e-300, max_iter=n_it,
                                       out=x)
        torch.cuda.synchronize()
        return 1000.0 * (time.perf_counter() - t0), inf

    iters = max(iters, 3)
    ms2, info2 = min((timed(2) for _ in range(3)), key=lambda v: v[0])
    ms, info = min((timed(iters) for _ in range(3)), key=lambda v: v[0])
    # the node product reads K's values once, each node's column list once for its 4 rows (3 u + 1
    # eeq), gathers x and writes y once
    spmv_bytes = 8 * ctx.nnz + 4 * ctx.nnz // 4 + 16 * ctx.n_rows
    per_it = (ms - ms2) / max(info["iterations"] - info2["iterations"], 1)
    return {"workload": f"hex8 gradient damage {len(C)} cells: bottom face clamped, top face stepped, "
                        "BiCGSTAB + diagonal preconditioner (SolveTrialState), fixed iterations",
            "equations": len(eqs), "independent_dofs": info["independent_dofs"],
            "iterations": info["iterations"], "ms": ms, "ms_per_iteration": per_it,
            "setup_ms": ms2 - info2["iterations"] * per_it,
            "k_products_per_iteration": 2,
            "k_product_GBps_lower_bound": 2 * spmv_bytes / (per_it * 1e-3) / 1e9,
            "relative_residual": info["relative_residual"]}


def adjoint_ratio(n, steps, warmup):
    """Adjoint vs forward on the differentiation workload: hex8 gradient damage on an n^3 mesh of an
    n mm cube with per-cell fields E ~ LN(30000, 0.1), f_t ~ LN(3, 0.1), G_f ~ LN(0.1, 0.15);
    forward = fused R+K assembly, adjoint = nuto_vjp producing K^T lambda and dL/dtheta."""
    import numpy as np
    import torch

    import nuto_amd as N
    from nuto_amd import synthetic as S

    X, C = N.structured_mesh(N.HEX8, n, n, n, (float(n), float(n), float(n)))
    ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
    ctx = N.Context()
    ctx.set_cells(N.HEX8, X, C)
    ctx.set_dof_type(0, N.HEX8, 3, 3 * nn, N.cell_dofs(ranks, 3))
    ctx.set_dof_type(1, N.HEX8, 1, nn, N.cell_dofs(ranks, 1))
    ctx.set_integrand(N.gradient_damage_params())
    ctx.set_rule(0, 2)
    nnz, n_rows = ctx.build_pattern()
    nc = C.shape[0]
    E, ft, gf = S.material_fields(nc)
    ctx.set_material_fields(E, ft, gf)
    u = torch.from_numpy(S.uniform(S.SEED_U, 3 * nn, -5e-3, 5e-3)).cuda()
    e = torch.from_numpy(S.uniform(S.SEED_E, nn, 0.0, 5e-4)).cuda()
    kap = torch.from_numpy(S.uniform(S.SEED_KAPPA, nc * 8, 0.0, 3e-4)).cuda()
    lam = torch.from_numpy(S.uniform(0x3001, n_rows, -1.0, 1.0)).cuda()
    R = torch.empty(n_rows, dtype=torch.float64, device="cuda")
    K = torch.empty(nnz, dtype=torch.float64, device="cuda")

    def timed(fn):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in evs:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(b) for a, b in evs]))

    fwd = timed(lambda: ctx.build_vector_matrix(u, e, kap, R, K))
    adj = timed(lambda: ctx.vjp(u, e, kap, lam))
    return {"workload": f"hex8 gradient damage {n}x{n}x{n} cells, per-cell lognormal (E, f_t, G_f)",
            "forward_ms": fwd, "adjoint_ms": adj, "ratio": adj / fwd,
            "adjoint_outputs": "K^T lambda (u, eeq) and dR/dtheta^T lambda (3 per cell)"}


def pcie_inclusive(ctx, u, e, kap, R, K, reps=3):
    """The same step for a caller that holds u, eeq and kappa in host memory and wants R and K
    back there: pinned host buffers, copies in, the assembly, copies out (never the headline value;
    the boundary takes device pointers and K is meant to stay resident for a device solver)."""
    import numpy as np
    import torch

    hu, he, hk = (t.cpu().pin_memory() for t in (u, e, kap))
    hR = torch.empty(R.numel(), dtype=torch.float64, pin_memory=True)
    hK = torch.empty(K.numel(), dtype=torch.float64, pin_memory=True)
    times = []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t = time.perf_counter()
        u.copy_(hu, non_blocking=True)
        e.copy_(he, non_blocking=True)
        kap.copy_(hk, non_blocking=True)
        ctx.build_vector_matrix(u, e, kap, R, K)
        hR.copy_(R, non_blocking=True)
        hK.copy_(K, non_blocking=True)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
    ms = 1000.0 * float(np.median(times[1:]))
    bytes_h2d = 8 * (u.numel() + e.numel() + kap.numel())
    bytes_d2h = 8 * (R.numel() + K.numel())
    return {"ms": ms, "cells_per_s": ctx.n_cells / (ms * 1e-3), "bytes_to_device": bytes_h2d,
            "bytes_to_host": bytes_d2h, "host_link_GBps": (bytes_h2d + bytes_d2h) / (ms * 1e-3) / 1e9,
            "note": "pinned host buffers; copies, the assembly and copies back in one timed region"}


def batched(n_meshes, steps, warmup, form=-1, tune=True):
    """Mixed-type batch: meshes with log-uniform element counts in [1e3, 5e4], tet4 / tet10 / hex8 /
    hex20 drawn uniformly, jittered grids (10 % of the spacing), local isotropic damage. Compares the
    tuned batched launch (graph replay) with one launch per mesh on the same data. form >= 0 forces
    the cell-matrix kernel of the two-phase launches (0 columns, 1 / 2 Gram, 3 Gram on MFMA); without
    tune the two-phase launch per element type (mode 5, whole groups) is timed as it is."""
    import numpy as np
    import torch

    import nuto_amd as N
    from nuto_amd import synthetic as S
    from nuto_amd import workloads as W

    t0 = time.time()
    meshes = W.mixed_batch(n_meshes)
    print(f"batch: {n_meshes} meshes generated ({time.time() - t0:.1f} s)", file=sys.stderr, flush=True)
    b = N.Batch()
    b.set_integrand(N.local_damage_params())
    for ct, dt, X, C, ranks, nn in meshes:
        b.add_mesh(ct, dt, X, C, ranks, nn)
    sizes = b.build()
    cells = sum(m[3].shape[0] for m in meshes)
    u = torch.from_numpy(S.uniform(S.SEED_U, sizes["u"], -5e-3, 5e-3)).cuda()
    kap = torch.from_numpy(S.uniform(S.SEED_KAPPA, sizes["history"], 0.0, 3e-4)).cuda()
    R = torch.empty(sizes["R"], dtype=torch.float64, device="cuda")
    K = torch.empty(sizes["K"], dtype=torch.float64, device="cuda")
    setup = time.time() - t0
    print(f"batch: patterns built ({setup:.1f} s)", file=sys.stderr, flush=True)
    if form >= 0:
        os.environ["NUTO_AMD_CELL_GRAM"] = str(form)
    t1 = time.time()
    if tune:
        ms_modes, best = b.autotune(u, None, kap, R, K)
    else:
        b.set_stage_bytes(0)
        ms_modes, best = [-1.0] * 7, 5
    tune_s = time.time() - t1
    print(f"batch: tuned in {tune_s:.1f} s (mode {best})", file=sys.stderr, flush=True)

    def timed(mode):
        for _ in range(warmup):
            b.assemble(u, None, kap, R, K, mode=mode)
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, z in evs:
            a.record(b.stream)
            b.assemble(u, None, kap, R, K, mode=mode)
            z.record(b.stream)
        torch.cuda.synchronize()
        return float(np.median([a.elapsed_time(z) for a, z in evs]))

    graph_ms = timed(-1 if tune else 5)
    # the plain per-mesh launches: the faster of the two kernel forms, one mesh at a time
    per_mesh_mode = 6 if 0 <= ms_modes[6] < ms_modes[4] or not tune else 4
    per_mesh_ms = timed(per_mesh_mode)
    counts = {}
    for ct, dt, *_ in meshes:
        k = {(N.TET4, N.TET4): "tet4", (N.TET4, N.TET10): "tet10", (N.HEX8, N.HEX8): "hex8",
             (N.HEX8, N.HEX20): "hex20"}[(ct, dt)]
        counts[k] = counts.get(k, 0) + 1
    names = {N.TET4: "tet4", N.TET10: "tet10", N.HEX8: "hex8", N.HEX20: "hex20"}
    forms = {0: "columns", 1: "gram, 1 row node per lane", 2: "gram, 2 row nodes per lane",
             3: "gram on the matrix cores (MFMA)", -1: "default"}
    kernels = []
    for g in range(b.group_info(0)["groups"]):
        gi = b.group_info(g)
        kernels.append({"type": names[gi["dof_type"]], "meshes": gi["meshes"], "cells": gi["cells"],
                        "cell_kernel": forms[gi["cell_kernel"]], "stage_cells": gi["stage_cells"]})
    out_bytes = 8 * (sizes["K"] + sizes["R"])
    return {"workload": f"{n_meshes} meshes, cells log-uniform in [1e3, 5e4], types {counts}, jitter 10 %, "
                        f"local isotropic damage R+K", "cells": cells, "nnz": sizes["K"],
            "output_GBps": out_bytes / (graph_ms * 1e-3) / 1e9, "groups": kernels,
            "batched_ms": graph_ms, "per_mesh_launch_ms": per_mesh_ms, "speedup": per_mesh_ms / graph_ms,
            "cells_per_s": cells / (graph_ms * 1e-3), "tuned_mode": best,
            "per_mesh_mode": per_mesh_mode,
            "autotune_ms": {"fused64": ms_modes[0], "fused128": ms_modes[1], "fused256": ms_modes[2],
                            "per_type": ms_modes[3], "per_mesh": ms_modes[4], "two_phase_per_type": ms_modes[5],
                            "two_phase_per_mesh": ms_modes[6]}, "setup_s": setup, "autotune_s": tune_s,
            "forced_form": form if form >= 0 else None}


def batched_adjoint(n_meshes, steps, warmup):
    """Gradient-damage batch: hex8 meshes with log-uniform element counts in [1e4, 2e5], per-cell
    lognormal fields; one graph replay computes R, K, lambda = R - r_obs, K^T lambda and dL/dtheta."""
    import numpy as np
    import torch

    import nuto_amd as N
    from nuto_amd import synthetic as S
    from nuto_amd import workloads as W

    t0 = time.time()
    meshes = W.gdm_batch(n_meshes)
    # device bytes per cell: K values and columns (436 nonzeros per cell), the IP state records,
    # geometry and tables; keep the batch within 80 % of the free memory (the full 256-mesh batch
    # needs about 115 GB of the MI355X's 288 GB)
    per_cell = 436 * 12 + 1152 + 400
    free, _ = torch.cuda.mem_get_info()
    dropped = 0
    while len(meshes) > 1 and per_cell * sum(m[3].shape[0] for m in meshes) > 0.8 * free:
        meshes.pop()
        dropped += 1
    b = N.Batch()
    b.set_integrand(N.gradient_damage_params())
    for ct, dt, X, C, ranks, nn in meshes:
        b.add_mesh(ct, dt, X, C, ranks, nn)
    sizes = b.build()
    cells = sum(m[3].shape[0] for m in meshes)
    b.set_material_fields(*S.material_fields(cells))
    u = torch.from_numpy(S.uniform(S.SEED_U, sizes["u"], -5e-3, 5e-3)).cuda()
    e = torch.from_numpy(S.uniform(S.SEED_E, sizes["e"], 0.0, 5e-4)).cuda()
    kap = torch.from_numpy(S.uniform(S.SEED_KAPPA, sizes["history"], 0.0, 3e-4)).cuda()
    robs = torch.from_numpy(S.uniform(0x5151, sizes["R"], -1e-2, 1e-2)).cuda()
    out = [torch.empty(sizes[k], dtype=torch.float64, device="cuda") for k in ("R", "K", "R", "theta")]
    setup = time.time() - t0
    for _ in range(warmup):
        b.residual_adjoint(u, e, kap, robs, *out)
    torch.cuda.synchronize()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for a, z in evs:
        a.record(b.stream)
        b.residual_adjoint(u, e, kap, robs, *out)
        z.record(b.stream)
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(z) for a, z in evs]))
    return {"workload": f"{len(meshes)} hex8 gradient-damage meshes, cells log-uniform in [1e4, 2e5], per-cell "
                        f"lognormal (E, f_t, G_f); R+K, lambda = R - r_obs, K^T lambda, dL/dtheta in one graph",
            "cells": cells, "nnz": sizes["K"], "ms": ms, "cells_per_s": cells / (ms * 1e-3), "setup_s": setup,
            "meshes": len(meshes), "meshes_dropped_for_memory": dropped}


def cpu_baseline(n, threads, n_one_core=24, budget_s=30.0):
    """Times the oracle (NuTo's algorithm restated: per-thread triplet lists, critical splice,
    setFromTriplets) on an n^3 sample of the same workload with `threads` OpenMP threads (R+K, and
    BuildVector and BuildMatrix alone), and on an n_one_core^3 sample with one thread (cells/s is
    intensive in the mesh size)."""
    import numpy as np

    import oracle as O
    from nuto_amd import synthetic as S

    def problem(nc):
        m = O.Mesh.bricks(nc, nc, nc)
        m.scale(float(nc))  # 1 mm spacing, as the 100 mm / 100 cells workload
        Ld = m.add_dof(O.HEX8, 3)
        Le = m.add_dof(O.HEX8, 1)
        p = O.Params(kind=1, dim=3, E=30000.0, nu=0.2, k=10.0, kappa0=1e-4, beta=100.0, alpha=0.99, c=4.0)
        P = O.Problem(m, p, 0, 2, Ld, Le)
        u = S.uniform(S.SEED_U, P.ndof[0], -5e-3, 5e-3)
        e = S.uniform(S.SEED_E, P.ndof[1], 0.0, 5e-4)
        k = S.uniform(S.SEED_KAPPA, P.ncells * 8, 0.0, 3e-4)
        return m, P, u, e, k

    def median_s(P, u, e, k, want_r, want_k, nth, runs, budget=budget_s):
        times = []
        t_start = time.time()
        while len(times) < runs and (time.time() - t_start) < budget:
            t0 = time.perf_counter()
            P.assemble(u, e, k, want_r=want_r, want_k=want_k, nthreads=nth)
            times.append(time.perf_counter() - t0)
        return float(np.median(times)), len(times)

    m, P, u, e, k = problem(n)
    P.assemble(u, e, k, want_r=True, want_k=True, nthreads=threads)  # warm-up (page faults, thread start)
    t, runs = median_s(P, u, e, k, True, True, threads, 3)
    # BuildVector and BuildMatrix separately (the reference times them as two calls)
    tv, _ = median_s(P, u, e, k, True, False, threads, 3, budget_s / 3)
    tm, _ = median_s(P, u, e, k, False, True, threads, 1, budget_s / 3)
    cells = P.ncells
    del P, m
    _, P1, u1, e1, k1 = problem(n_one_core)
    t1, _ = median_s(P1, u1, e1, k1, True, True, 1, 1)
    return {"value": cells / t, "unit": "cells/s", "cores": threads, "kind": "port",
            "sample": f"hex8 gradient damage {n}x{n}x{n} cells (1 mm spacing, {cells} cells), R+K, median of "
                      f"{runs} runs of the oracle with {threads} OpenMP threads (triplet lists, critical splice, "
                      f"setFromTriplets); one core: {n_one_core}x{n_one_core}x{n_one_core} cells",
            "build_vector_cells_per_s": cells / tv, "build_matrix_cells_per_s": cells / tm,
            "one_core_cells_per_s": n_one_core ** 3 / t1, "host": _host_cpu()}


def _host_cpu():
    """lscpu's model, sockets and cores, and the CPUs this process may use."""
    info = {}
    try:
        import subprocess

        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            key, _, val = line.partition(":")
            if key.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core", "CPU(s)"):
                info[key.strip()] = val.strip()
    except Exception:  # noqa: BLE001 - reported, never fatal
        pass
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["usable_cpus"] = os.cpu_count()
    info["OMP_NUM_THREADS"] = os.environ.get("OMP_NUM_THREADS")
    return info


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=100, help="cells per direction (100 = the headline workload)")
    ap.add_argument("--cpu-sample", type=int, default=50, help="cells per direction of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads for the CPU baseline (0: OMP_NUM_THREADS, else the usable CPUs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--pcie", action="store_true",
                    help="also time the step with host-resident inputs and outputs (pinned copies in and out)")
    ap.add_argument("--adjoint-n", type=int, default=64, help="cells per direction of the adjoint workload")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--solve-iters", type=int, default=20,
                    help="BiCGSTAB iterations of the constrained solve on the assembled system (0: skip)")
    ap.add_argument("--batch-meshes", type=int, default=256, help="meshes of the mixed-type batch (0: skip)")
    ap.add_argument("--batch-form", type=int, default=-1,
                    help="force the cell-matrix kernel of the mixed batch (0 columns, 1 / 2 Gram, 3 Gram on MFMA)")
    ap.add_argument("--batch-no-tune", action="store_true",
                    help="time the mixed batch's two-phase launch per element type without the autotune")
    ap.add_argument("--batch-gdm-meshes", type=int, default=256,
                    help="meshes of the gradient-damage residual + adjoint batch (0: skip)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (counter calibration) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU: it measures the single-GPU assembly (asked for --gpus {args.gpus}). "
              "Several GPUs would each own a range of dof nodes plus one layer of ghost cells; that sharding "
              "is not launched.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import nuto_amd as N
    from nuto_amd import synthetic as S

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(1)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    n = args.n
    t0 = time.time()
    X, C = N.structured_mesh(N.HEX8, n, n, n, (100.0 * n / 100, 100.0 * n / 100, 100.0 * n / 100))
    ranks, nn = N.dof_layer(N.HEX8, C, N.HEX8)
    ctx = N.Context()
    ctx.set_cells(N.HEX8, X, C)
    ctx.set_dof_type(0, N.HEX8, 3, 3 * nn, N.cell_dofs(ranks, 3))
    ctx.set_dof_type(1, N.HEX8, 1, nn, N.cell_dofs(ranks, 1))
    ctx.set_integrand(N.gradient_damage_params())
    ctx.set_rule(0, 2)
    nnz, n_rows = ctx.build_pattern()
    n_cells, nip = C.shape[0], 8
    u = torch.from_numpy(S.uniform(S.SEED_U, 3 * nn, -5e-3, 5e-3)).cuda()
    e = torch.from_numpy(S.uniform(S.SEED_E, nn, 0.0, 5e-4)).cuda()
    kap = torch.from_numpy(S.uniform(S.SEED_KAPPA, n_cells * nip, 0.0, 3e-4)).cuda()
    R = torch.empty(n_rows, dtype=torch.float64, device="cuda")
    K = torch.empty(nnz, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    log(f"setup {setup_s:.1f}s: {n_cells} cells, {n_rows} rows, {nnz} nonzeros")

    for _ in range(args.warmup):
        ctx.build_vector_matrix(u, e, kap, R, K)
    torch.cuda.synchronize()

    # kernel time with HIP events on the stream the kernels run on (torch's current stream)
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for a, b in evs:
        a.record()
        ctx.build_vector_matrix(u, e, kap, R, K)
        b.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    kernel_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    ms_per_step = 1000.0 * wall / args.steps
    value = n_cells * args.steps / wall

    bytes_alg = algorithmic_bytes(nnz, n_rows, nn, n_cells, nip, 8)
    achieved = bytes_alg / (kernel_ms * 1e-3) / 1e9
    rec = {
        "metric": "hex8 GDM cells/s assembled (R+K FP64), % HBM roofline; adjoint/forward ratio",
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded SplitMix64 state on a structured mesh)",
        "config": {"workload": f"hex8 gradient damage R+K, {n}x{n}x{n} cells of a {100 * n // 100} mm cube, "
                               f"2x2x2 Gauss", "cells": n_cells, "dofs": n_rows, "nnz": nnz,
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(n),
                     "kernel_ms": kernel_ms, "algorithmic_bytes": bytes_alg,
                     "measured_copy_GBps": copy_bandwidth()},
        "setup_s": setup_s,
    }
    flops = measured_counter(n, "fp64_flops_per_launch")
    if flops is not None:
        # FP64 VALU flops per step (rocprofv3 SQ_INSTS_VALU_FLOPS_FP64 x 64 lanes, both kernels)
        tflops = flops / (kernel_ms * 1e-3) / 1e12
        rec["fp64"] = {"flops_per_step": flops, "achieved_tflops": tflops, "peak_tflops": FP64_PEAK_TFLOPS,
                       "frac": tflops / FP64_PEAK_TFLOPS}
    if args.pcie:
        ctx.n_cells = n_cells
        rec["pcie_inclusive"] = pcie_inclusive(ctx, u, e, kap, R, K)
    if args.solve_iters > 0:
        rec["solve"] = solve_bench(ctx, X, C, ranks, nn, K, R, args.solve_iters, 100.0 * n / 100)
    if not args.no_adjoint:
        del R, K
        torch.cuda.empty_cache()
        rec["adjoint"] = adjoint_ratio(args.adjoint_n, args.steps, args.warmup)
        torch.cuda.empty_cache()
    if args.batch_meshes:
        rec["batched"] = batched(args.batch_meshes, min(args.steps, 10), args.warmup, args.batch_form,
                                 not args.batch_no_tune)
        torch.cuda.empty_cache()
    if args.batch_gdm_meshes:
        rec["batched_adjoint"] = batched_adjoint(args.batch_gdm_meshes, min(args.steps, 10), args.warmup)
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        # all the cores the process is given (the GPU host allots a share per GPU: OMP_NUM_THREADS)
        omp = os.environ.get("OMP_NUM_THREADS", "")
        threads = args.cpu_threads or (int(omp) if omp.isdigit() and int(omp) > 0 else _host_cpu()["usable_cpus"])
        try:
            cb = cpu_baseline(args.cpu_sample, threads)
            # context only: vs_baseline stays null (no published number exists for this metric)
            cb["gpu_over_cpu"] = value / cb["value"]
            cb["gpu_over_one_core"] = value / cb["one_core_cells_per_s"]
            rec["cpu_baseline"] = cb
        except Exception as ex:  # the baseline is reported, never the measured value
            rec["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
