"""Benchmark of the Vanadis DR-HEX8 solve path on one MI355X.

Workload: 128x128x64 HEX8 unit box (1,081,665 nodes, 1,048,576 elements),
steady advection-diffusion with DR stabilisation (V1), lambda = 1e-3, seeded
velocity, Gaussian source, Dirichlet inflow faces; data generated on the host,
resident in HBM before timing.

A step is one plain-FP64 Jacobi-BiCG solve to rtol 1e-8 from x0 = 0
(time-to-rtol).  `value` is the operator throughput in GDOF/s (nodes / time of
one FP64 matrix-free apply, owner-computes scatter), measured with HIP events
on the library's stream.  The line also carries time-to-rtol of FP64 BiCG,
FP64 BiCGSTAB and mixed precision (FP32 BiCGSTAB inner solves, FP64
refinement), the other scatter modes, the roofline of the dominant kernel
(the owner launch of every BiCG iteration: direction update p = z + beta p,
pp = zz + beta pp fused into the dual apply A p, A^T pp), that of the plain
dual apply, the 32x32x16 and 256x256x64 workloads, the batched solves (FP64
and mixed precision), the implicit-Euler learned-guess run and, in the CPU
leg, the oracle's serial and OpenMP BiCG timed on the host, the rel-L2 of every
GPU solution against the oracle's FP64 solve to 1e-12, and the legacy drop-in
kernel_dbcg_simple_ on the oracle-assembled element matrices.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--kernel-only] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "BiCG time-to-rtol 1e-8 and operator GDOF/s; mixed/learned vs plain FP64"
HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0  # dense BF16 MFMA of the MI355X


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def algorithmic_bytes(N, E, dual, fp32=False, fused=False):
    """Compulsory bytes of one matrix-free apply with this library's layout:
    input/output vectors, the uint8 Dirichlet mask and five per-element
    fields (vx, vy, vz, lambda, tau); axis coordinates are O(n^(1/3)).
    fused: the BiCG iteration launch reads z, zz and the previous directions,
    writes the new directions and A p, A^T pp, and reads and writes x
    (x += alpha p of the previous update): 10 vectors."""
    w = 4 if fp32 else 8
    vec = (10 if fused else 4 if dual else 2) * w * N
    return vec + N + 5 * w * E


def host_cpu():
    """Model, physical cores (unique core/socket pairs of `lscpu -p`), logical
    CPUs, the CPUs this process may run on and OMP_NUM_THREADS; the baseline
    threads = physical cores, capped by the affinity set and OMP_NUM_THREADS
    when the environment sets it (a host shared with other jobs)."""
    import subprocess
    info = {"model": None, "physical_cores": None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            if line.startswith("Model name:"):
                info["model"] = line.split(":", 1)[1].strip()
        pairs = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True,
                               timeout=20).stdout
        info["physical_cores"] = len({l for l in pairs.splitlines() if l and l[0] != "#"})
    except (OSError, subprocess.SubprocessError):
        pass
    t = info["physical_cores"] or info["affinity_cpus"]
    t = min(t, info["affinity_cpus"])
    if info["omp_num_threads_env"] and info["omp_num_threads_env"].isdigit():
        t = min(t, int(info["omp_num_threads_env"]))
    info["threads_used"] = max(1, t)
    return info


def cpu_baseline_and_checks(pkg, p2, sols, legacy_reps=2, serial_iters=40):
    """The CPU leg: the oracle (oracle/, the C restatement of the reference's
    dbcg_simple + matvec_0/mttvec_0, VANADIS.FOR:977-1055, 2149-2245, over
    element matrices assembled like przest3d, 713-771) timed on the host at
    the 128x128x64 workload, serial (as the reference's solver) and with the
    8-colour OpenMP products; the oracle as the reference for the GPU solutions
    (rel-L2 against its FP64 BiCG run to 1e-12: live for 32x32x16 and
    128x128x64, from the committed sketches for 256x256x64, the batches and
    the time stepping); and the legacy drop-in kernel_dbcg_simple_ on the
    GPU, fed with the oracle's element matrices as the Fortran host feeds it.
    sols: {(config, name): solution}."""
    import __graft_entry__ as ge
    po = ge.load_oracle()
    host = host_cpu()
    T = host["threads_used"]
    out = {"host": host}
    progress(f"cpu: {host['model']}, {host['physical_cores']} physical cores, {T} threads")
    # --- 128x128x64: assembly, serial sample, OpenMP solve, reference solve
    t0 = time.perf_counter()
    bb, b, d = po.assemble(p2, threads=T)
    t_asm = time.perf_counter() - t0
    b, d, x0 = po.impose_mask(p2.mask, b, d, np.zeros(p2.N))
    args = (bb, p2.iss, d, p2.nax, p2.nay, p2.naz, p2.mask)
    x = np.random.default_rng(0).standard_normal(p2.N)
    po.matvec_0(x, p2.iss, bb, p2.nax, p2.nay, p2.naz, p2.mask)
    t0 = time.perf_counter()
    for _ in range(3):
        po.matvec_0(x, p2.iss, bb, p2.nax, p2.nay, p2.naz, p2.mask)
    t_ser_apply = (time.perf_counter() - t0) / 3
    t0 = time.perf_counter()
    _, it_s, _, _ = po.dbcg_simple(b, x0, 1e-30, serial_iters, *args)
    t_ser = time.perf_counter() - t0
    with open(os.path.join(ROOT, "tests", "golden", "oracle_c2.json")) as f:
        ser_its = json.load(f)["iterations"]  # the serial oracle's count to 1e-8
    ms_it = 1e3 * t_ser / it_s
    po.matvec_0_colored(x, p2.iss, bb, p2.nax, p2.nay, p2.naz, p2.mask, T)
    t0 = time.perf_counter()
    for _ in range(5):
        po.matvec_0_colored(x, p2.iss, bb, p2.nax, p2.nay, p2.naz, p2.mask, T)
    t_omp_apply = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    x_omp, it_omp, err_omp, rc_omp = po.dbcg_simple_omp(b, x0, 1e-8, 20000, *args, T)
    t_omp = time.perf_counter() - t0
    progress(f"cpu 128x128x64: serial {ms_it:.1f} ms/iteration ({it_s} sampled), OpenMP "
             f"{it_omp} iterations in {t_omp:.2f} s on {T} threads")
    x_o, it_o, _, rc_o = po.dbcg_simple_omp(b, x0, 1e-12, 20000, *args, T)
    N = p2.N
    out.update({
        "value": round(N / t_ser_apply / 1e9, 5), "unit": "GDOF/s", "cores": 1, "kind": "port",
        "sample": f"the oracle's serial matvec_0 on stored element matrices at 128x128x64 "
                  f"({N} nodes), 3 applies; its serial dbcg_simple sampled over {it_s} "
                  f"iterations ({t_ser:.1f} s) and projected to its {ser_its} iterations to "
                  f"rtol 1e-8; assembly (OpenMP, {T} threads) {t_asm:.1f} s, reported apart",
        "ms_per_apply": round(1e3 * t_ser_apply, 3),
        "ms_per_iteration": round(ms_it, 3),
        "iterations_to_rtol": ser_its,
        "time_to_rtol_s_projected": round(ms_it * ser_its / 1e3, 3),
        "assembly_s": round(t_asm, 3),
        "openmp": {"value": round(N / t_omp_apply / 1e9, 5), "unit": "GDOF/s", "cores": T,
                   "kind": "port", "time_to_rtol_s": round(t_omp, 3), "iterations": it_omp,
                   "converged": rc_omp == 0,
                   "ms_per_iteration": round(1e3 * t_omp / max(1, it_omp), 3),
                   "sample": f"the oracle's dbcg_simple with 8-colour OpenMP products and "
                             f"fixed-order OpenMP reductions, whole solve to rtol 1e-8 from "
                             f"x0 = 0 on {T} threads, OMP_PROC_BIND "
                             f"{os.environ.get('OMP_PROC_BIND', 'unset')}, OMP_PLACES "
                             f"{os.environ.get('OMP_PLACES', 'unset')}"}})
    # --- rel-L2 of every GPU solution against the oracle
    rel = {}
    n2 = np.linalg.norm(x_o)
    for (cfg, name), xs in sols.items():
        if cfg == "c2":
            rel[f"c2/{name}"] = float(np.linalg.norm(xs - x_o) / n2)
    rel["c2/cpu_openmp_bicg"] = float(np.linalg.norm(x_omp - x_o) / n2)
    refs = {"oracle_c2": f"oracle FP64 BiCG to 1e-12 ({it_o} iterations, OpenMP), live"}
    # legacy drop-in: kernel_dbcg_simple_ on these element matrices (the
    # unchanged host's call), whole call and its phases
    leg = {}
    for sc in ("colored", "atomic"):
        times = []
        for _ in range(legacy_reps):
            t0 = time.perf_counter()
            xl, itl, errl = pkg.kernel_dbcg_simple(N, b, x0, 1e-8, 20000, bb, p2.iss, d, p2.nax,
                                                   p2.nay, p2.naz, p2.mask, scatter=sc)
            times.append((time.perf_counter() - t0, pkg._lib.legacy_last_times()))
        wall, (up, so, down) = times[-1]
        leg[sc] = {"iterations": itl, "err": errl, "whole_call_s": round(wall, 4),
                   "upload_s": round(up, 4), "solve_s": round(so, 4),
                   "download_s": round(down, 4),
                   "rel_l2_vs_oracle": float(np.linalg.norm(xl - x_o) / n2)}
        progress(f"legacy {sc}: {itl} iterations, call {wall * 1e3:.1f} ms (upload "
                 f"{up * 1e3:.1f}, solve {so * 1e3:.1f}, download {down * 1e3:.1f})")
    pkg._lib.load().vgpu_legacy_release()
    del bb
    # --- 32x32x16: serial and OpenMP to rtol, live reference solve
    p1 = pkg.problem.config_problem("c1")
    bb1, b1, d1 = po.assemble(p1, threads=T)
    b1, d1, z1 = po.impose_mask(p1.mask, b1, d1, np.zeros(p1.N))
    a1 = (bb1, p1.iss, d1, p1.nax, p1.nay, p1.naz, p1.mask)
    t0 = time.perf_counter()
    _, it1s, _, _ = po.dbcg_simple(b1, z1, 1e-8, 20000, *a1)
    t1s = time.perf_counter() - t0
    t0 = time.perf_counter()
    _, it1o, _, _ = po.dbcg_simple_omp(b1, z1, 1e-8, 20000, *a1, T)
    t1o = time.perf_counter() - t0
    x1o, _, _, _ = po.dbcg_simple_omp(b1, z1, 1e-12, 20000, *a1, T)
    out["config1"] = {"serial": {"time_to_rtol_s": round(t1s, 4), "iterations": it1s},
                      "openmp": {"time_to_rtol_s": round(t1o, 4), "iterations": it1o,
                                 "cores": T}}
    for (cfg, name), xs in sols.items():
        if cfg == "c1":
            rel[f"c1/{name}"] = float(np.linalg.norm(xs - x1o) / np.linalg.norm(x1o))
    refs["oracle_c1"] = "oracle FP64 BiCG to 1e-12 (OpenMP), live"
    # --- sketches of the oracle's solutions of the large workloads
    sk = po.load_sketches()
    for (cfg, name), xs in sols.items():
        if cfg in ("c3", "c5") and cfg in sk:
            rel[f"{cfg}/{name}"] = po.sketch_rel_l2(xs, sk[cfg])
        if cfg in ("c4a", "c4b") and cfg in sk:
            est = [po.sketch_rel_l2(xm, e) for xm, e in zip(xs, sk[cfg]["members"])]
            rel[f"{cfg}/{name}_max_over_members"] = max(est)
    refs["sketched"] = ("256x256x64, the batches and the time stepping: estimates from 32 "
                           "(batch members: 16) Gaussian projections of the oracle's FP64 "
                           "solution to 1e-12 (tests/golden/oracle_sketch.json; +-27 % / "
                           "+-41 % at 99 %)")
    return out, {"rel_l2": rel, "reference_solves": refs}, leg


def measured_traffic(tuned, kernel="iteration_launch"):
    """HBM bytes per launch of `kernel` (iteration_launch, dual_apply,
    bicgstab_direction_launch, bicgstab_s_launch) from the committed
    rocprofv3 --pmc summary of the same kernel, tile and z chunk (FETCH_SIZE
    and WRITE_SIZE, each scaled by the copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "traffic_dual_apply_fp64_128x128x64.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    key = "0/0" if tuned is None else f"{tuned['tile']}/{tuned['zchunk']}"
    e = t.get(kernel, {}).get(key)
    if not e:
        return None
    return {"bytes_per_launch": e["bytes_per_launch"],
            "source": f"profiles/traffic_dual_apply_fp64_128x128x64.json {kernel} "
                      f"tile/zchunk {key}"}


def _owner_shape(tile):
    return {0: (32, 8), 1: (16, 16), 2: (64, 4)}[int(tile)]


# owner-kernel template arguments after <double, double, ...> per kernel: the
# solver launches in their block-addressed form (mode + 10, the form the
# solvers and vgpu_time_kernel use) or, for systems whose engine block is too
# large for it, the plain one
_TRAFFIC_KERNELS = {"iteration_launch": ("true, false, false, 11,", "true, false, false, 1,"),
                    "dual_apply": ("true, false, false, 0,",),
                    "bicgstab_direction_launch": ("false, false, false, 12,",
                                                  "false, false, false, 2,"),
                    "bicgstab_s_launch": ("false, false, false, 13,",
                                          "false, false, false, 3,")}


def live_traffic(tuned, N, E, timeout=180):
    """HBM bytes per launch of the owner launches, measured in this run: two
    child processes `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` over
    `bench.py --kernel-only` at the same tile and z chunk (one counter per
    pass), each counter scaled by the factor that turns the same counter of
    the copy kernel k_copy_wide (64 N doubles read and written per launch,
    far above the last-level cache) into its known bytes.  Returns
    {kernel: {...}, "calibration": {...}} or None (profiler missing or failed)."""
    import csv
    import glob
    import shutil
    import signal
    import subprocess
    import tempfile
    if shutil.which("rocprofv3") is None:
        return None
    tile = 0 if tuned is None else int(tuned["tile"])
    zc = 0 if tuned is None else int(tuned["zchunk"])
    tx, ty = _owner_shape(tile)
    med = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="vgpu_pmc_", dir=os.environ.get("TMPDIR", "/tmp"))
        cmd = ["rocprofv3", "--pmc", counter, "--output-format", "csv", "-d", d, "-o", "k",
               "--", sys.executable, "-u", os.path.join(ROOT, "bench.py"), "--kernel-only",
               "--tile", str(tile), "--zchunk", str(zc)]
        progress(f"traffic: {counter} pass")
        pr = subprocess.Popen(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                              start_new_session=True)
        try:
            rc = pr.wait(timeout=timeout)
        except subprocess.TimeoutExpired:
            os.killpg(pr.pid, signal.SIGKILL)
            pr.wait()
            progress(f"traffic: {counter} pass timed out")
            shutil.rmtree(d, ignore_errors=True)
            return None
        files = glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True)
        if rc != 0 and not files:
            progress(f"traffic: {counter} pass failed (rc {rc})")
            shutil.rmtree(d, ignore_errors=True)
            return None
        vals = {}
        for f in files:
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    name = row.get("Kernel_Name", "")
                    if row.get("Counter_Name") != counter:
                        continue
                    key = None
                    if "k_copy_wide" in name:
                        key = "copy"
                    elif "k_owner<double, double," in name and f", {tx}, {ty}>" in name:
                        for k, sigs in _TRAFFIC_KERNELS.items():
                            if any(f"k_owner<double, double, {sig}" in name for sig in sigs):
                                key = k
                    if key:
                        vals.setdefault(key, []).append(float(row["Counter_Value"]))
        shutil.rmtree(d, ignore_errors=True)
        if "copy" not in vals:
            progress(f"traffic: no copy-kernel {counter} values")
            return None
        med[counter] = {k: float(np.median(v)) for k, v in vals.items()}
        med[counter + "_n"] = {k: len(v) for k, v in vals.items()}
    copy_bytes = 64 * 8 * N  # read, and the same written, per launch
    ff = copy_bytes / (med["FETCH_SIZE"]["copy"] * 1024.0)
    fw = copy_bytes / (med["WRITE_SIZE"]["copy"] * 1024.0)
    alg = {"iteration_launch": 81 * N + 40 * E, "dual_apply": 33 * N + 40 * E,
           "bicgstab_direction_launch": 89 * N + 40 * E, "bicgstab_s_launch": 41 * N + 40 * E}
    out = {"calibration": {"fetch_factor": round(ff, 5), "write_factor": round(fw, 5),
                           "copy_bytes_each_way": copy_bytes,
                           "copy_fetch_kib": med["FETCH_SIZE"]["copy"],
                           "copy_write_kib": med["WRITE_SIZE"]["copy"]},
           "tile": tile, "zchunk": zc}
    for k in _TRAFFIC_KERNELS:
        if k in med["FETCH_SIZE"] and k in med["WRITE_SIZE"]:
            fb = med["FETCH_SIZE"][k] * 1024.0 * ff
            wb = med["WRITE_SIZE"][k] * 1024.0 * fw
            out[k] = {"bytes_per_launch": int(round(fb + wb)), "fetch_bytes": int(round(fb)),
                      "write_bytes": int(round(wb)),
                      "fetch_kib_counter": round(med["FETCH_SIZE"][k], 1),
                      "write_kib_counter": round(med["WRITE_SIZE"][k], 1),
                      "dispatches": med["FETCH_SIZE_n"][k],
                      "ratio_to_algorithmic": round((fb + wb) / alg[k], 4)}
    return out


def config3(pkg, sols):
    """256x256x64 elements (4.29M nodes), per-element diffusivity log-uniform
    in [1e-4, 1e-2]: FP64 BiCG, FP64 BiCGSTAB and mixed precision (FP32 inner
    BiCGSTAB, rtol 1e-4, FP64 refinement) to rtol 1e-8; rel-L2 of the
    BiCGSTAB and mixed solutions vs FP64 BiCG."""
    progress("config 3: 256x256x64")
    p = pkg.problem.config_problem("c3")
    ctx = pkg.Context.from_problem(p)
    t64 = ctx.autotune(pkg.FP64, reps=5)
    b = ctx.assemble_rhs(p.q_in)
    b_dev, x_dev = ctx.dev_alloc(8 * p.N), ctx.dev_alloc(8 * p.N)
    ctx.h2d(b_dev, b)
    zeros = np.zeros(p.N)
    out = {"workload": "HEX8 unit box 256x256x64 elements, 4293185 nodes, lambda log-uniform "
                       "[1e-4, 1e-2] per element, V1 tau, rtol 1e-8 from x0=0"}
    xs = {}
    # mixed: the inner rtol 1e-4 the workload specifies, and 1e-2, the inner
    # tolerance that measured best over the configs and batch members
    # each method at its own autotuned launch shape: the FP64 solves at the
    # FP64 iteration launch's, the mixed solves at the FP32 one's (their inner
    # FP32 launches are most of their work)
    for name, prec, inner in (("fp64", pkg.FP64, 0.0), ("fp64_bicgstab", pkg.FP64, 0.0),
                              ("mixed", pkg.MIXED, 1e-4), ("mixed_inner_1e-2", pkg.MIXED, 1e-2)):
        if name == "mixed":
            t32 = ctx.autotune(pkg.MIXED, reps=5)
            out["mixed_launch_shape"] = {"tile": t32["tile"], "zchunk": t32["zchunk"],
                                         "tuned_for": "FP32 BiCGSTAB launches of the inner iterations"}
        ctx.set_method("bicgstab" if name == "fp64_bicgstab" else "bicg")
        for rep in range(2):  # first solve captures the graph
            ctx.h2d(x_dev, zeros)
            it, err, sec, rc = ctx.solve_dev(b_dev, x_dev, rtol=1e-8, itmax=50000,
                                             precision=prec, inner_rtol=inner)
        ctx.set_method("bicg")
        xs[name] = ctx.d2h(x_dev, p.N)
        sols[("c3", name)] = xs[name]
        out[name] = {"iterations": it, "seconds": round(sec, 4), "status": rc,
                     "true_rel_residual": ctx.true_residual(b, xs[name])}
        progress(f"config 3 {name}: {it} iterations, {sec * 1e3:.1f} ms, rc {rc}")
    ctx.set_options("owner", t64["zchunk"], t64["tile"])  # the FP64 shape again
    out["fp64_launch_shape"] = {"tile": t64["tile"], "zchunk": t64["zchunk"]}
    out["mixed_rel_l2_vs_fp64"] = float(np.linalg.norm(xs["mixed"] - xs["fp64"]) /
                                        np.linalg.norm(xs["fp64"]))
    out["speedup_mixed_vs_fp64"] = round(out["fp64"]["seconds"] / out["mixed"]["seconds"], 3)
    out["speedup_mixed_inner_1e-2_vs_fp64"] = round(
        out["fp64"]["seconds"] / out["mixed_inner_1e-2"]["seconds"], 3)
    out["bicgstab_rel_l2_vs_fp64"] = float(np.linalg.norm(xs["fp64_bicgstab"] - xs["fp64"]) /
                                           np.linalg.norm(xs["fp64"]))
    out["speedup_bicgstab_vs_fp64"] = round(out["fp64"]["seconds"] /
                                            out["fp64_bicgstab"]["seconds"], 3)
    out["dual_apply_ms"] = {"fp64": round(ctx.time_kernel(1, "owner", pkg.FP64, 20), 5),
                            "fp32": round(ctx.time_kernel(1, "owner", pkg.FP32, 20), 5)}
    # HBM rooflines where the working set (0.3-0.6 GB per launch) exceeds the
    # 256 MB last-level cache; same kernels and byte counts as at 128x128x64
    N, E = p.N, p.E
    roof = {"peak": HBM_PEAK_GBS, "unit": "GB/s",
            "note": "FP64 owner launches at the autotuned tile and z chunk; working set above "
                    "the 256 MB last-level cache"}
    for what, name, nbytes in ((0, "single_apply", 17 * N + 40 * E),
                               (1, "dual_apply", 33 * N + 40 * E),
                               (3, "bicg_iteration_launch", 81 * N + 40 * E),
                               (5, "bicgstab_direction_launch", 89 * N + 40 * E),
                               (6, "bicgstab_s_launch", 41 * N + 40 * E)):
        ms = float(np.median([ctx.time_kernel(what, "owner", pkg.FP64, 20) for _ in range(3)]))
        gbs = nbytes / (ms * 1e-3) / 1e9
        roof[name] = {"ms": round(ms, 5), "algorithmic_bytes_per_launch": nbytes,
                      "achieved": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    out["roofline_hbm"] = roof
    progress(f"config 3 rooflines: single {roof['single_apply']['frac']}, dual "
             f"{roof['dual_apply']['frac']}, BiCG iteration launch "
             f"{roof['bicg_iteration_launch']['frac']}, BiCGSTAB "
             f"{roof['bicgstab_direction_launch']['frac']} / {roof['bicgstab_s_launch']['frac']}")
    ctx.close()
    return out


def config1(pkg, sols):
    """32x32x16 elements (18,513 nodes): FP64 BiCG, FP64 BiCGSTAB and mixed
    precision to rtol 1e-8 from x0 = 0, time-to-rtol with the data resident
    (the second of two solves)."""
    p = pkg.problem.config_problem("c1")
    out = {"workload": "HEX8 unit box 32x32x16 elements, 18513 nodes, lambda 1e-3, V1 tau, "
                       "rtol 1e-8 from x0=0"}
    with pkg.Context.from_problem(p) as ctx:
        b = ctx.assemble_rhs(p.q_in)
        b_dev, x_dev = ctx.dev_alloc(8 * p.N), ctx.dev_alloc(8 * p.N)
        ctx.h2d(b_dev, b)
        for name, prec, method in (("fp64_bicg_owner", pkg.FP64, "bicg"),
                                   ("fp64_bicgstab", pkg.FP64, "bicgstab"),
                                   ("mixed", pkg.MIXED, "bicg")):
            ctx.set_method(method)
            for _ in range(2):
                ctx.h2d(x_dev, np.zeros(p.N))
                it, err, sec, rc = ctx.solve_dev(b_dev, x_dev, rtol=1e-8, itmax=20000,
                                                 precision=prec)
            ctx.set_method("bicg")
            sols[("c1", name)] = ctx.d2h(x_dev, p.N)
            out[name] = {"iterations": it, "seconds": round(sec, 5), "status": rc}
    progress(f"config 1: fp64 {out['fp64_bicg_owner']['seconds'] * 1e3:.2f} ms, bicgstab "
             f"{out['fp64_bicgstab']['seconds'] * 1e3:.2f} ms, mixed "
             f"{out['mixed']['seconds'] * 1e3:.2f} ms")
    return out


def conv_layers(pkg, ctx, N):
    """The learned guess's layer kernels on this node grid (BF16 MFMA, FP32
    accumulation, 16 channels, 3x3x3 taps): GPU time per launch and the two
    rooflines.  A layer does 2 x 27 x 16 x 16 flops per node; it moves 32 bytes
    per node per BF16 tensor read or written (hidden forward: X in, Y out;
    backward data: dZ and Yprev in, dZprev out; weight gradients: X and dZ in)."""
    flops = 2.0 * 27 * 16 * 16 * N
    out = {"flops_per_launch": flops, "peak_mfma_bf16_TFLOPs": MFMA_BF16_PEAK_TFLOPS,
           "peak_hbm_GBps": HBM_PEAK_GBS}
    net = pkg.Net(ctx)
    try:
        for what, name, tensors in ((0, "hidden_forward", 2), (1, "backward_data", 3),
                                    (2, "weight_gradient", 2)):
            ms = net.time_layers(what, reps=20)
            nbytes = 32.0 * tensors * N
            tf = flops / (ms * 1e-3) / 1e12
            gbs = nbytes / (ms * 1e-3) / 1e9
            t_mfma = flops / (MFMA_BF16_PEAK_TFLOPS * 1e12)
            t_hbm = nbytes / (HBM_PEAK_GBS * 1e9)
            out[name] = {"ms": round(ms, 5), "TFLOPs": round(tf, 1),
                         "frac_mfma": round(tf / MFMA_BF16_PEAK_TFLOPS, 4),
                         "algorithmic_bytes": nbytes, "GBps": round(gbs, 1),
                         "frac_hbm": round(gbs / HBM_PEAK_GBS, 4),
                         "bound": "hbm" if t_hbm > t_mfma else "mfma",
                         "frac_of_bound": round(max(t_mfma, t_hbm) / (ms * 1e-3), 4)}
        out["prediction_ms"] = round(net.time_layers(3, reps=10), 5)
    finally:
        net.close()
    progress(f"conv layers: forward {out['hidden_forward']['ms'] * 1e3:.1f} us "
             f"({out['hidden_forward']['TFLOPs']} TFLOP/s), backward data "
             f"{out['backward_data']['ms'] * 1e3:.1f} us, weight gradient "
             f"{out['weight_gradient']['ms'] * 1e3:.1f} us, prediction "
             f"{out['prediction_ms']:.3f} ms")
    return out


def learned_guess(pkg, sols):
    """Implicit Euler (dt 0.01, 200 steps, helix source of seed 0) on the
    128x128x64 box, every step to rtol 1e-8, from the previous-step guess
    with BiCG (the plain method), and with linear, quadratic and cubic
    extrapolation + Jacobi sweeps and the trained network (committed
    parameters, trained on seeds 1-10 by learned.main) + 2 Jacobi sweeps,
    each solved by BiCG and by BiCGSTAB; speedups against the plain method."""
    path = os.path.join(ROOT, "vanadis_solver-directional_residual_amd", "data",
                        "learned_guess_128x128x64.npz")
    if not os.path.exists(path):
        return {"status": "no trained parameters"}
    params = pkg.learned.load_params(path)
    modes = ("previous", "extrap+jacobi", "extrap2+jacobi", "net+jacobi", "previous@bicgstab",
             "extrap+jacobi@bicgstab", "extrap2+jacobi@bicgstab", "extrap2+jacobi4@bicgstab",
             "extrap3+jacobi8@bicgstab", "net+jacobi@bicgstab")
    progress(f"learned guess: {len(modes)} x 200 implicit-Euler steps")
    res, _ = pkg.learned.experiment(pkg, params=params, modes=modes, tune=True, keep_final=True)
    for m, xf in res.pop("finals").items():
        sols[("c5", m)] = xf
    out = {"workload": "HEX8 unit box 128x128x64, implicit Euler dt=0.01, 200 steps, "
                       "helix source seed 0, rtol 1e-8 each step; mode guess[+jacobi[N]]"
                       "[@method] (N Jacobi sweeps, default 2), default method BiCG; owner tile "
                       "and z chunk autotuned "
                       "once on the implicit-Euler system",
           "autotune": res.get("autotune")}
    for m in modes:
        out[m] = {"iterations": res[m]["iterations"], "seconds": res[m]["seconds"],
                  "fallbacks": res[m]["fallbacks"]}
        if m != "previous":
            out[m]["speedup_vs_previous"] = res[f"{m}_speedup_vs_previous"]
    out["net_speedup_vs_previous"] = res["net+jacobi_speedup_vs_previous"]
    out["extrap_speedup_vs_previous"] = res["extrap+jacobi_speedup_vs_previous"]
    out["net_bicgstab_speedup_vs_previous"] = res["net+jacobi@bicgstab_speedup_vs_previous"]
    out["extrap_bicgstab_speedup_vs_previous"] = \
        res["extrap+jacobi@bicgstab_speedup_vs_previous"]
    out["extrap2_bicgstab_speedup_vs_previous"] = \
        res["extrap2+jacobi@bicgstab_speedup_vs_previous"]
    out["extrap2_jacobi4_bicgstab_speedup_vs_previous"] = \
        res["extrap2+jacobi4@bicgstab_speedup_vs_previous"]
    out["extrap3_jacobi8_bicgstab_speedup_vs_previous"] = \
        res["extrap3+jacobi8@bicgstab_speedup_vs_previous"]
    progress(f"learned: net {out['net_speedup_vs_previous']}x, extrap "
             f"{out['extrap_speedup_vs_previous']}x, net+bicgstab "
             f"{out['net_bicgstab_speedup_vs_previous']}x, extrap+bicgstab "
             f"{out['extrap_bicgstab_speedup_vs_previous']}x, quadratic extrap+bicgstab "
             f"{out['extrap2_bicgstab_speedup_vs_previous']}x (4 sweeps "
             f"{out['extrap2_jacobi4_bicgstab_speedup_vs_previous']}x), cubic extrap + 8 sweeps "
             f"+ bicgstab {out['extrap3_jacobi8_bicgstab_speedup_vs_previous']}x vs "
             f"previous-step guess")
    return out


def batched_solves(pkg, tuned, sols):
    """Batched FP64 and mixed-precision solves to rtol 1e-8 against the same
    systems solved one by one: 32 right-hand sides (own diffusivity and
    source) on one 96x96x48 mesh, and 16 boxes of n^3 elements, n uniform in
    24..96.  The owner tile and z chunk are the autotuner's for the
    128x128x64 workload (a z chunk > 0, so that each batch member's inner
    BiCGSTAB sums like its single-context solve)."""
    res = {}
    for name, gen in (("rhs32_96x96x48", lambda: pkg.problem.batch_rhs_problems()),
                      ("subdomains16_24to96", lambda: pkg.problem.batch_subdomain_problems())):
        progress(f"batched {name}: generating")
        probs = gen()
        ctxs = [pkg.Context.from_problem(q) for q in probs]
        zc = tuned["zchunk"] if tuned and tuned["zchunk"] > 0 else 8
        for c in ctxs:
            c.set_options("owner", zc, 0 if tuned is None else tuned["tile"])
        bs = [c.assemble_rhs(q.q_in) for c, q in zip(ctxs, probs)]
        itmax = 20000
        with pkg.Batch(ctxs) as bt:
            bt.solve(bs, rtol=1e-8, itmax=itmax, raise_on_fail=False)  # warm-up, capture
            t0 = time.time()
            xs, its, errs, st, rc = bt.solve(bs, rtol=1e-8, itmax=itmax, raise_on_fail=False)
            t_batch = time.time() - t0
        for c, bb in zip(ctxs, bs):
            c.solve(bb, rtol=1e-8, itmax=itmax, raise_on_fail=False)  # warm-up
        t0 = time.time()
        same = True
        for m, (c, bb) in enumerate(zip(ctxs, bs)):
            x, it, err, _ = c.solve(bb, rtol=1e-8, itmax=itmax, raise_on_fail=False)
            same = same and it == its[m] and bool(np.array_equal(x, xs[m]))
        t_seq = time.time() - t0
        nodes = sum(q.N for q in probs)
        res[name] = {"systems": len(probs), "nodes": nodes,
                     "iterations_min_max": [int(its.min()), int(its.max())],
                     "all_converged": bool((st == 1).all()),
                     "status_counts": {str(k): int((st == k).sum()) for k in sorted(set(st))},
                     "batched_s": round(t_batch, 4), "one_by_one_s": round(t_seq, 4),
                     "speedup": round(t_seq / t_batch, 3),
                     "bitwise_equal_to_single": same,
                     "note": "host-to-device copies of b and x included in both"}
        progress(f"batched {name}: {t_batch * 1e3:.1f} ms vs {t_seq * 1e3:.1f} ms one by one")
        if not res[name]["all_converged"]:
            res[name]["with_bicgstab_fallback"] = batched_fallback(pkg, ctxs, bs)
        res[name]["mixed"] = batched_mixed(pkg, ctxs, bs)
        sols[("c4a" if name.startswith("rhs") else "c4b", "mixed")] = res[name]["mixed"].pop("xs")
        for c in ctxs:
            c.close()
    return res


def batched_mixed(pkg, ctxs, bs, itmax=20000):
    """Mixed precision (FP32 inner BiCGSTAB to 1e-4, FP64 refinement) to rtol
    1e-8: the batch (one FP64 owner launch per refinement for all members,
    the inner solves together in one pair of owner launches per iteration)
    against each member's mixed solve on its own context."""
    with pkg.Batch(ctxs) as bt:
        bt.solve(bs, rtol=1e-8, itmax=itmax, precision=pkg.MIXED, raise_on_fail=False)
        t0 = time.time()
        xs, its, errs, st, rc = bt.solve(bs, rtol=1e-8, itmax=itmax, precision=pkg.MIXED,
                                         raise_on_fail=False)
        t_batch = time.time() - t0
    for c, bb in zip(ctxs, bs):
        c.solve(bb, rtol=1e-8, itmax=itmax, precision=pkg.MIXED, raise_on_fail=False)
    t0 = time.time()
    same = True
    for m, (c, bb) in enumerate(zip(ctxs, bs)):
        x, it, err, _ = c.solve(bb, rtol=1e-8, itmax=itmax, precision=pkg.MIXED,
                                raise_on_fail=False)
        same = same and it == its[m] and bool(np.array_equal(x, xs[m]))
    t_seq = time.time() - t0
    tres = max(c.true_residual(bb, x) for c, bb, x in zip(ctxs, bs, xs))
    progress(f"  mixed: {t_batch * 1e3:.1f} ms batched vs {t_seq * 1e3:.1f} ms one by one, "
             f"all converged {bool((st == 1).all())}, bitwise {same}")
    return {"inner_iterations_min_max": [int(its.min()), int(its.max())],
            "all_converged": bool((st == 1).all()), "true_rel_residual_max": float(tres),
            "batched_s": round(t_batch, 4), "one_by_one_s": round(t_seq, 4),
            "speedup": round(t_seq / t_batch, 3), "bitwise_equal_to_single": same, "xs": xs}


def batched_fallback(pkg, ctxs, bs, cap=2000, itmax=20000):
    """The same batch with the BiCG iterations capped at `cap` and the
    members left unconverged finished by BiCGSTAB on their own contexts
    (vgpu_batch_set_fallback), against the same policy one system at a time."""
    with pkg.Batch(ctxs) as bt:
        bt.set_fallback("bicgstab", itmax=itmax)
        bt.solve(bs, rtol=1e-8, itmax=cap, raise_on_fail=False)  # warm-up, capture
        t0 = time.time()
        xs, its, errs, st, rc = bt.solve(bs, rtol=1e-8, itmax=cap, raise_on_fail=False)
        t_batch = time.time() - t0

    def one(c, bb):
        # the batch's policy on one context: BiCGSTAB after a BiCG that ends
        # unconverged or with a true residual above rtol (from zero when the
        # iterate is worse than zero)
        x, it, err, rc1 = c.solve(bb, rtol=1e-8, itmax=cap, raise_on_fail=False)
        tr = c.true_residual(bb, x)
        fb = rc1 != pkg.OK or not tr <= 1e-8
        if fb:
            if not tr <= 1.0:
                x = np.zeros_like(x)
            c.set_method("bicgstab")
            try:
                for _ in range(3):  # passes while the true residual is above rtol
                    x, it2, err, rc1 = c.solve(bb, x, rtol=1e-8, itmax=itmax,
                                               raise_on_fail=False)
                    it += it2
                    if rc1 != pkg.OK or c.true_residual(bb, x) <= 1e-8:
                        break
            finally:
                c.set_method("bicg")
        return x, it, rc1, fb

    for c, bb in zip(ctxs, bs):
        one(c, bb)  # warm-up
    t0 = time.time()
    same, fell_back = True, 0
    for m, (c, bb) in enumerate(zip(ctxs, bs)):
        x, it, _, fb = one(c, bb)
        fell_back += int(fb)
        same = same and it == its[m] and bool(np.array_equal(x, xs[m]))
    t_seq = time.time() - t0
    progress(f"  with BiCGSTAB fallback ({fell_back} members): {t_batch * 1e3:.1f} ms vs "
             f"{t_seq * 1e3:.1f} ms one by one, all converged {bool((st == 1).all())}")
    return {"bicg_cap": cap, "members_finished_by_bicgstab": fell_back,
            "iterations_min_max": [int(its.min()), int(its.max())],
            "all_converged": bool((st == 1).all()),
            "true_rel_residual_max": float(max(c.true_residual(bb, x)
                                               for c, bb, x in zip(ctxs, bs, xs))),
            "batched_s": round(t_batch, 4), "one_by_one_s": round(t_seq, 4),
            "speedup": round(t_seq / t_batch, 3), "bitwise_equal_to_single": same}


def own_zchunk_default(tuned, a):
    return tuned["zchunk"] if tuned else max(a.zchunk, 0)


def own_tile_default(tuned, a):
    return tuned["tile"] if tuned else max(a.tile, 0)


def shard_shares(pkg):
    """One GPU's share of the 256x256x64 workload split over 8 GPUs, three
    ways (z-slab 256x256x8, y-slab 256x32x64, pencil 128x64x64; the element
    sizes of the whole mesh, the per-element diffusivity field), against the
    whole mesh on one GPU: the owner launches at the autotuner's tile and z
    chunk, HIP events, the median of three timings of 20 launches.  The
    exchange of boundary planes a multi-GPU run adds is not part of it."""
    out = {"note": "owner launches of one GPU's share of 256x256x64 over 8 GPUs; "
                   "no halo exchange", "shares": {}}
    for name, size, ext in (("whole_256x256x64", (256, 256, 64), (1.0, 1.0, 1.0)),
                            ("z_slab_256x256x8", (256, 256, 8), (1.0, 1.0, 0.125)),
                            ("y_slab_256x32x64", (256, 32, 64), (1.0, 0.125, 1.0)),
                            ("pencil_128x64x64", (128, 64, 64), (0.5, 0.25, 1.0))):
        p = pkg.problem.unit_box(*size, vel_seed=0, lam_range=(1e-4, 1e-2), extent=ext)
        N, E = p.N, p.E
        with pkg.Context.from_problem(p) as ctx:
            t = ctx.autotune(pkg.FP64, reps=20)
            r = {"nodes": N, "elements": E, "tile": t["tile"], "zchunk": t["zchunk"]}
            for what, kind, nbytes in ((0, "single_apply", 17 * N + 40 * E),
                                       (1, "dual_apply", 33 * N + 40 * E),
                                       (3, "bicg_iteration_launch", 81 * N + 40 * E),
                                       (5, "bicgstab_direction_launch", 89 * N + 40 * E),
                                       (6, "bicgstab_s_launch", 41 * N + 40 * E)):
                ms = float(np.median([ctx.time_kernel(what, "owner", pkg.FP64, 20)
                                      for _ in range(3)]))
                r[kind] = {"us": round(ms * 1e3, 2),
                           "frac": round(nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            out["shares"][name] = r
            progress(f"{name}: BiCG iteration launch {r['bicg_iteration_launch']['us']} us")
    whole = out["shares"]["whole_256x256x64"]["bicg_iteration_launch"]["us"]
    for name, r in out["shares"].items():
        if name != "whole_256x256x64":
            r["iteration_launch_vs_whole_over_8"] = round(whole / 8 / r["bicg_iteration_launch"]["us"], 3)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel-only", action="store_true",
                    help="only launch the dual apply, the BiCG and BiCGSTAB owner launches "
                         "and the copy kernels (counter runs)")
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baseline, the oracle checks and the legacy drop-in leg")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-stab", action="store_true", help="skip the FP64 BiCGSTAB solves")
    ap.add_argument("--no-scatter-solves", action="store_true",
                    help="skip the BiCG solves with the colored and atomic scatters")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default owner tile (32x8) instead of the autotuner's")
    ap.add_argument("--tile", type=int, default=-1,
                    help="owner tile shape 0/1/2 (default: autotuner, or 0 with --kernel-only)")
    ap.add_argument("--zchunk", type=int, default=0,
                    help="owner z chunk with --tile (0 = balanced launch)")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the 256x256x64 FP64 / mixed-precision solves")
    ap.add_argument("--no-learned", action="store_true",
                    help="skip the implicit-Euler run with the trained initial guess")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched solves (32 right-hand sides, 16 subdomains)")
    ap.add_argument("--size", default="128,128,64")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 counter passes that measure the owner "
                         "launches' HBM bytes (then the committed summary is reported)")
    ap.add_argument("--shard-shares", action="store_true",
                    help="only time the owner launches on one GPU's share of 256x256x64 "
                         "over 8 GPUs (z-slab, y-slab, pencil) and print them as JSON")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(the slab decomposition is a design, see the sharding notes)", file=sys.stderr)
        sys.exit(2)

    # the CPU leg's OpenMP threads one per core, packed (read when the oracle's
    # OpenMP runtime starts; a caller's own settings win)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import __graft_entry__ as ge
    pkg = ge.load_package()
    pkg.load()
    if a.shard_shares:
        print(json.dumps(shard_shares(pkg)))
        return
    ex, ey, ez = (int(v) for v in a.size.split(","))
    progress(f"generating {ex}x{ey}x{ez} problem")
    p = pkg.problem.unit_box(ex, ey, ez, vel_seed=0, lam=1e-3)
    N, E = p.N, p.E
    ctx = pkg.Context.from_problem(p)
    ctx.set_options("owner", max(a.zchunk, 0), max(a.tile, 0))
    tuned = None
    if a.tile >= 0:
        tuned = {"tile": a.tile, "zchunk": max(a.zchunk, 0), "ms": None, "source": "--tile"}
    elif not a.no_tune and not a.kernel_only:
        tuned = ctx.autotune(pkg.FP64, reps=20)
        tuned["source"] = "autotuner"
        progress(f"autotuner: tile {tuned['tile']} zchunk {tuned['zchunk']} "
                 f"({tuned['ms'] * 1e3:.1f} us iteration launch)")
    b = ctx.assemble_rhs(p.q_in)
    b_dev = ctx.dev_alloc(8 * N)
    x_dev = ctx.dev_alloc(8 * N)
    ctx.h2d(b_dev, b)
    zeros = np.zeros(N)

    if a.kernel_only:
        for _ in range(3):
            ctx.time_kernel(1, "owner", pkg.FP64, reps=5)
            ctx.time_kernel(3, "owner", pkg.FP64, reps=5)
            ctx.time_kernel(5, "owner", pkg.FP64, reps=5)
            ctx.time_kernel(6, "owner", pkg.FP64, reps=5)
            ctx.time_kernel(2, "owner", pkg.FP64, reps=5)
            ctx.time_kernel(7, "owner", pkg.FP64, reps=5)
        progress("kernel-only done")
        print(json.dumps({"kernel_only": True, "N": N, "E": E}))
        return

    def solve(precision, rtol=1e-8, inner=0.0):
        ctx.h2d(x_dev, zeros)
        it, err, sec, rc = ctx.solve_dev(b_dev, x_dev, rtol=rtol, itmax=50000,
                                         precision=precision, inner_rtol=inner)
        if rc != pkg.OK:
            raise RuntimeError(f"solve failed rc={rc} it={it} err={err}")
        return it, err, sec

    sols = {}  # (config, name) -> solution, compared with the oracle at the end
    progress("warmup")
    for _ in range(a.warmup):
        solve(pkg.FP64)
    progress(f"timing {a.steps} FP64 solves")
    secs, iters = [], []
    t0 = time.time()
    for _ in range(a.steps):
        it, err, sec = solve(pkg.FP64)
        secs.append(sec)
        iters.append(it)
    wall = time.time() - t0
    x64 = ctx.d2h(x_dev, N)
    sols[("c2", "fp64_bicg_owner")] = x64
    ms_step = 1e3 * float(np.mean(secs))
    progress(f"fp64: {iters[-1]} iterations, {ms_step:.2f} ms to rtol")

    # mixed precision (reported, never fatal for the line)
    msec, mits, rel_l2_mixed, true_res_mixed, mixed_status = [], 0, None, None, "ok"
    mixed_inner2 = None
    mixed_shape = None
    if not a.no_mixed:
        try:
            # the mixed solves at the FP32 iteration launch's autotuned shape
            # (their inner FP32 launches are most of their work); the FP64
            # shape is restored after them
            if tuned is not None and tuned.get("source") == "autotuner":
                t32 = ctx.autotune(pkg.MIXED, reps=20)
                mixed_shape = {"tile": t32["tile"], "zchunk": t32["zchunk"],
                               "tuned_for": "FP32 BiCGSTAB launches of the inner iterations"}
            solve(pkg.MIXED)
            for _ in range(max(1, min(3, a.steps))):
                it, err_m, sec = solve(pkg.MIXED)
                msec.append(sec)
                mits = it
            xm = ctx.d2h(x_dev, N)
            sols[("c2", "mixed")] = xm
            rel_l2_mixed = float(np.linalg.norm(xm - x64) / np.linalg.norm(x64))
            true_res_mixed = ctx.true_residual(b, xm)
            progress(f"mixed: {mits} inner iterations, {1e3 * np.mean(msec):.2f} ms")
            # the same with inner rtol 1e-2 (the default is 1e-4, as the
            # 256x256x64 workload specifies)
            solve(pkg.MIXED, inner=1e-2)
            m2 = []
            for _ in range(max(1, min(3, a.steps))):
                it2, _, sec2 = solve(pkg.MIXED, inner=1e-2)
                m2.append(sec2)
            x2 = ctx.d2h(x_dev, N)
            sols[("c2", "mixed_inner_1e-2")] = x2
            mixed_inner2 = {"iterations": it2, "seconds": round(float(np.mean(m2)), 5),
                            "true_rel_residual": ctx.true_residual(b, x2)}
            progress(f"mixed, inner rtol 1e-2: {it2} inner iterations, "
                     f"{1e3 * np.mean(m2):.2f} ms")
        except RuntimeError as exc:
            mixed_status = f"failed: {exc}"
            progress(f"mixed {mixed_status}")
        finally:
            ctx.set_options("owner", own_zchunk_default(tuned, a), own_tile_default(tuned, a))
    else:
        mixed_status = "skipped"

    # the two scatter flavours of the reference's solver (8 colour launches;
    # FP64 atomics), FP64 BiCG to the same rtol (reported, never fatal)
    own_zchunk = own_zchunk_default(tuned, a)
    own_tile = own_tile_default(tuned, a)
    # brick shape and depth of the two brick scatters from the brick tuner
    brick_tuned = {}
    for sc in ("colored", "atomic"):
        if a.no_tune:
            brick_tuned[sc] = {"tile": 0, "zchunk": 0, "ms": None, "source": "library rule"}
        else:
            brick_tuned[sc] = ctx.autotune_bricks(sc, pkg.FP64, reps=10)
            brick_tuned[sc]["source"] = "vgpu_autotune_bricks"
            progress(f"brick tuner {sc}: shape {brick_tuned[sc]['tile']} depth "
                     f"{brick_tuned[sc]['zchunk']} ({brick_tuned[sc]['ms'] * 1e3:.1f} us dual)")
        ctx.set_options("owner", own_zchunk_default(tuned, a), own_tile_default(tuned, a))
    scatter_solves = {}
    if not a.no_scatter_solves:
        for sc in ("colored", "atomic"):
            ctx.set_options(sc, brick_tuned[sc]["zchunk"], brick_tuned[sc]["tile"])
            try:
                solve(pkg.FP64)
                csec = []
                for _ in range(2):
                    it_c, _, sec = solve(pkg.FP64)
                    csec.append(sec)
                xc = ctx.d2h(x_dev, N)
                sols[("c2", f"fp64_bicg_{sc}")] = xc
                scatter_solves[sc] = {
                    "iterations": it_c, "ms": round(1e3 * float(np.mean(csec)), 3),
                    "speedup_vs_owner": round(ms_step / (1e3 * float(np.mean(csec))), 3),
                    "rel_l2_vs_owner": float(np.linalg.norm(xc - x64) / np.linalg.norm(x64)),
                    "true_rel_residual": ctx.true_residual(b, xc)}
                progress(f"{sc} scatter: {it_c} iterations, {1e3 * np.mean(csec):.2f} ms")
            except RuntimeError as exc:
                scatter_solves[sc] = {"status": f"failed: {exc}"}
                progress(f"{sc} scatter solve failed: {exc}")
            finally:
                ctx.set_options("owner", own_zchunk, own_tile)

    # FP64 BiCGSTAB (same preconditioner, stopping rule and rtol; reported,
    # never fatal for the line)
    ssec, sits, rel_l2_stab, true_res_stab, stab_status = [], 0, None, None, "ok"
    if not a.no_stab:
        try:
            ctx.set_method("bicgstab")
            solve(pkg.FP64)
            for _ in range(max(1, min(3, a.steps))):
                it, err_s, sec = solve(pkg.FP64)
                ssec.append(sec)
                sits = it
            xs = ctx.d2h(x_dev, N)
            sols[("c2", "fp64_bicgstab")] = xs
            rel_l2_stab = float(np.linalg.norm(xs - x64) / np.linalg.norm(x64))
            true_res_stab = ctx.true_residual(b, xs)
            progress(f"bicgstab: {sits} iterations, {1e3 * np.mean(ssec):.2f} ms")
        except RuntimeError as exc:
            stab_status = f"failed: {exc}"
            progress(f"bicgstab {stab_status}")
        finally:
            ctx.set_method("bicg")
    else:
        stab_status = "skipped"

    # kernels, HIP events on the library stream: each figure the median of
    # three timings of `reps` back-to-back launches (one timing can catch a
    # clock transient)
    reps = 50

    def timed(what, n_launch):
        return float(np.median([ctx.time_kernel(what, "owner", pkg.FP64, n_launch)
                                for _ in range(3)]))

    t_single = timed(0, reps)
    t_dual = timed(1, reps)
    t_iter = timed(3, reps)
    t_copy = timed(2, reps)
    t_copy8 = timed(4, 20)
    t_copyh = timed(7, 10)
    t_sdir = timed(5, reps)
    t_ss = timed(6, reps)
    try:
        conv = conv_layers(pkg, ctx, N)
    except Exception as exc:  # reported, never fatal for the line
        conv = {"status": f"failed: {exc}"}
        progress(f"conv layers {conv['status']}")
    modes = {}
    for sc in ("owner", "colored", "atomic"):
        # brick scatters: the brick tuner's shape and depth
        if sc != "owner":
            ctx.set_options(sc, brick_tuned[sc]["zchunk"], brick_tuned[sc]["tile"])
        for prec, pn in ((pkg.FP64, "fp64"), (pkg.FP32, "fp32")):
            ms = ctx.time_kernel(0, sc, prec, 20)
            modes[f"{sc}_{pn}"] = {"ms": round(ms, 5), "GDOF/s": round(N / ms / 1e6, 2)}
        ctx.set_options("owner", own_zchunk, own_tile)
    gdofs = N / t_single / 1e6
    dual_bytes = algorithmic_bytes(N, E, dual=True)
    dual_gbs = dual_bytes / (t_dual * 1e-3) / 1e9
    iter_bytes = algorithmic_bytes(N, E, dual=True, fused=True)
    achieved = iter_bytes / (t_iter * 1e-3) / 1e9
    copy_gbs = 16.0 * N / (t_copy * 1e-3) / 1e9
    copy8_gbs = 128.0 * N / (t_copy8 * 1e-3) / 1e9
    copyh_gbs = 1024.0 * N / (t_copyh * 1e-3) / 1e9  # 64 N doubles in and out
    # BiCGSTAB owner launches: direction (R, V, P, T in; R, P, V out; x in
    # and out, d and RT at the owned nodes; mask; records) and S (R, V in;
    # T out; d and RT at the owned nodes; mask; records)
    sdir_bytes = 89 * N + 40 * E
    ss_bytes = 41 * N + 40 * E
    stab_roof = {
        "direction_launch": {"ms": round(t_sdir, 5), "algorithmic_bytes_per_launch": sdir_bytes,
                             "achieved": round(sdir_bytes / (t_sdir * 1e-3) / 1e9, 1),
                             "peak": HBM_PEAK_GBS, "unit": "GB/s",
                             "frac": round(sdir_bytes / (t_sdir * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                             "kernel": "k_owner<double,double,single,kStabDir>"},
        "s_launch": {"ms": round(t_ss, 5), "algorithmic_bytes_per_launch": ss_bytes,
                     "achieved": round(ss_bytes / (t_ss * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ss_bytes / (t_ss * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "kernel": "k_owner<double,double,single,kStabS>"}}
    progress(f"apply {t_single * 1e3:.1f} us ({gdofs:.1f} GDOF/s), dual {t_dual * 1e3:.1f} us "
             f"({dual_gbs:.0f} GB/s), iteration launch {t_iter * 1e3:.1f} us "
             f"({achieved:.0f} GB/s), copy {copy_gbs:.0f} GB/s (8 N: {copy8_gbs:.0f} GB/s), "
             f"BiCGSTAB launches {t_sdir * 1e3:.1f} / {t_ss * 1e3:.1f} us")

    c1 = None
    try:
        c1 = config1(pkg, sols)
    except Exception as exc:  # reported, never fatal for the line
        c1 = {"status": f"failed: {exc}"}
        progress(f"c1 {c1['status']}")

    batch = None
    if not a.no_batch:
        try:
            batch = batched_solves(pkg, tuned, sols)
        except Exception as exc:  # reported, never fatal for the line
            batch = {"status": f"failed: {exc}"}
            progress(f"batched {batch['status']}")

    c3 = None
    if not a.no_c3:
        try:
            c3 = config3(pkg, sols)
        except Exception as exc:  # reported, never fatal for the line
            c3 = {"status": f"failed: {exc}"}
            progress(f"c3 {c3['status']}")

    learned = None
    if not a.no_learned:
        try:
            learned = learned_guess(pkg, sols)
        except Exception as exc:  # reported, never fatal for the line
            learned = {"status": f"failed: {exc}"}
            progress(f"learned {learned['status']}")

    cpu, checks, legacy = None, None, None
    if not a.no_cpu:
        progress("cpu baseline, oracle checks and the legacy drop-in")
        cpu, checks, legacy = cpu_baseline_and_checks(pkg, p, sols)

    out = {
        "metric": METRIC, "value": round(gdofs, 3), "unit": "GDOF/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(ms_step, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded velocity/source fields, host-generated)",
        "config": {"workload": f"HEX8 unit box {ex}x{ey}x{ez} elements, {N} nodes, steady "
                               "DR advection-diffusion lambda=1e-3, V1 tau, FP64 Jacobi-BiCG "
                               "to rtol 1e-8 from x0=0, owner-computes scatter",
                   "elements": E, "nodes": N},
        "time_to_rtol_s": {"fp64": round(float(np.mean(secs)), 5),
                           "mixed": round(float(np.mean(msec)), 5) if msec else None,
                           "fp64_bicgstab": round(float(np.mean(ssec)), 5) if ssec else None},
        "iterations": {"fp64": iters[-1], "mixed_inner_total": mits, "fp64_bicgstab": sits},
        "speedup_bicgstab_vs_fp64": (round(float(np.mean(secs)) / float(np.mean(ssec)), 3)
                                     if ssec else None),
        "bicgstab_status": stab_status,
        "bicgstab_rel_l2_vs_fp64": rel_l2_stab, "bicgstab_true_rel_residual": true_res_stab,
        "speedup_mixed_vs_fp64": (round(float(np.mean(secs)) / float(np.mean(msec)), 3)
                                  if msec else None),
        "mixed_status": mixed_status, "mixed_launch_shape": mixed_shape,
        "mixed_rel_l2_vs_fp64": rel_l2_mixed, "mixed_true_rel_residual": true_res_mixed,
        "mixed_inner_1e-2": (dict(mixed_inner2, speedup_vs_fp64=round(
            float(np.mean(secs)) / mixed_inner2["seconds"], 3)) if mixed_inner2 else None),
        "wall_s_timed_steps": round(wall, 3),
        "apply_ms": {"single_fp64": round(t_single, 5), "dual_fp64": round(t_dual, 5),
                     "iteration_launch_fp64": round(t_iter, 5)},
        "ms_per_iteration_fp64": round(ms_step / iters[-1], 5),
        "scatter_modes": modes,
        "scatter_solves_fp64": scatter_solves,
        "brick_autotune": brick_tuned,
        "copy_GBps": round(copy_gbs, 1),
        "copy_8N_GBps": round(copy8_gbs, 1),
        "copy_hbm_GBps": round(copyh_gbs, 1),
        "practical_hbm_roof": {"GBps": round(copyh_gbs, 1), "frac_of_peak":
                               round(copyh_gbs / HBM_PEAK_GBS, 4),
                               "note": "16-byte copy of 64 N doubles (0.55 GB each way), far "
                                       "above the 256 MB last-level cache"},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None, "kernel": "k_owner<double,double,dual,direction>",
                     "algorithmic_bytes_per_launch": iter_bytes},
        "roofline_bicgstab": stab_roof,
        "conv_layers_bf16": conv,
        "roofline_dual_apply": {"bound": "hbm", "achieved": round(dual_gbs, 1),
                                "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                "frac": round(dual_gbs / HBM_PEAK_GBS, 4),
                                "kernel": "k_owner<double,double,dual>",
                                "algorithmic_bytes_per_launch": dual_bytes},
        # the same launches priced at the bytes a general unstructured layout
        # moves (x, y, node coordinates, mask; 8-node connectivity and
        # velocity + diffusivity per element: 41 N + 64 E per single apply,
        # 193 N + 64 E per fused BiCG iteration); this structured kernel reads
        # no coordinates or connectivity, so its own compulsory bytes (the
        # rooflines above) are the lower, stricter count
        "general_layout_bytes": {
            "single_apply": {"bytes": 41 * N + 64 * E, "ms": round(t_single, 5),
                             "GBps": round((41 * N + 64 * E) / (t_single * 1e-3) / 1e9, 1),
                             "frac": round((41 * N + 64 * E) / (t_single * 1e-3) / 1e9
                                           / HBM_PEAK_GBS, 4)},
            "bicg_iteration": {"bytes": 193 * N + 64 * E, "ms": round(ms_step / iters[-1], 5),
                               "GBps": round((193 * N + 64 * E) / (ms_step / iters[-1] * 1e-3)
                                             / 1e9, 1),
                               "frac": round((193 * N + 64 * E) / (ms_step / iters[-1] * 1e-3)
                                             / 1e9 / HBM_PEAK_GBS, 4)}},
        "cpu_baseline": cpu,
        "rel_l2_vs_oracle": checks,
        "legacy_drop_in": legacy,
        "config1": c1,
        "autotune": tuned,
        "batched": batch,
        "learned_guess": learned,
        "config3": c3,
    }
    live = None if a.no_traffic else live_traffic(tuned, N, E)
    if live and "iteration_launch" in live:
        out["traffic_live"] = live
        src = (f"measured in this run: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE over bench.py "
               f"--kernel-only --tile {live['tile']} --zchunk {live['zchunk']}, scaled by "
               f"the copy kernel's factors {live['calibration']['fetch_factor']} / "
               f"{live['calibration']['write_factor']}")
        out["roofline"]["traffic"] = live["iteration_launch"]["bytes_per_launch"]
        out["roofline"]["traffic_source"] = src
        out["roofline_dual_apply"]["traffic"] = live.get("dual_apply", {}).get("bytes_per_launch")
        for key, kname in (("direction_launch", "bicgstab_direction_launch"),
                           ("s_launch", "bicgstab_s_launch")):
            stab_roof[key]["traffic"] = live.get(kname, {}).get("bytes_per_launch")
    else:
        # the profiler is unavailable or was skipped: the committed summary
        if not a.no_traffic:
            progress("traffic: no live counter values for the iteration launch; the committed "
                     "summary is reported")
        traffic = measured_traffic(tuned)
        if traffic:
            out["roofline"]["traffic"] = traffic["bytes_per_launch"]
            out["roofline"]["traffic_source"] = traffic["source"] + " (committed, not this run)"
        for key, kname in (("direction_launch", "bicgstab_direction_launch"),
                           ("s_launch", "bicgstab_s_launch")):
            tr = measured_traffic(tuned, kname)
            stab_roof[key]["traffic"] = tr["bytes_per_launch"] if tr else None
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
