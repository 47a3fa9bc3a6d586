#!/usr/bin/env python3
"""Headline benchmark: Jacobi-PCG time to relative residual 1e-10 on a 1000x1000
Q4 plane-strain cantilever (2,004,002 dofs, 36,024,004 nnz, per-element E
log-uniform in [1e-3, 1] from a fixed seed), FP32-value mixed precision vs FP64,
plus the SpMV's achieved HBM bandwidth.

A step is one full PCG solve on the resident, assembled and constrained system
(x0 = 0, stop when the recurrence residual satisfies ||r||/||b|| <= 1e-10, the
reference's rule with a relative threshold). Prints one JSON line. The CPU baseline times the compiled reference (oracle/_ref) on a bounded
number of PCG iterations of the same system on this host's cores and scales the
per-iteration time to the iteration count of the full solve.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG time-to-1e-10 (ms) + SpMV HBM GB/s, 1000² Q4; FP32-mixed vs FP64"
HBM_PEAK_GBS = 8000.0


def _sync():
    lib = ctypes.CDLL("libamdhip64.so")
    lib.hipDeviceSynchronize()


KERNELS = {1: "pcg_spmv_kernel<4>", 2: "pcg_spmv_kernel<8>", 3: "pcg_spmv_kernel<16>",
           4: "pcg_spmv_stream_kernel<24, 64>", 5: "pcg_spmv_stream_kernel<24, 32>",
           6: "pcg_spmv_slices_kernel<9>", 7: "pcg_spmv_slices_dir_kernel<9, 0>",
           8: "pcg_spmv_stencil_kernel<4, true>", 9: "pcg_spmv_stencil_kernel<9, false>",
           10: "pcg_fused_stencil_kernel<4, 0>", 11: "pcg_fused_grid_kernel<0, 4>",
           12: "pcg_fused_slices_kernel<9, 0>", 13: "pcg_slices_multiply_kernel<9, 0>",
           14: "pcg_fold_slices_kernel<9, 0>", 15: "pcg_sym_slices_multiply_kernel<4, 5, 0>"}


def kernel_name(variant: int, layout: dict) -> str:
    """The kernel a launch of `variant` runs on this layout (variant 8 takes the 2D tiles of
    the fused grid iteration where the stencil is row-structured)."""
    if variant == 8 and layout.get("grid_row_len"):
        return "pcg_spmv_grid_kernel<4>"
    return KERNELS[variant]


def spmv_bytes(n: int, nnz: int, value_bytes: int, variant: int, layout: dict) -> int:
    """Algorithmic bytes of one SpMV launch: the values, the column indices of the
    layout, its row (or slice) offsets, x read once and y written once. CSR variants
    carry one int32 column per value; the node-slice layout one per 2x2 block of four
    values and one offset per 64 nodes (padding not counted). Variant 7 forms the
    search direction p = M^-1 r + beta p_old on the fly: it reads r, the Jacobi
    inverse and p_old once and writes q and the new p. The node-stencil layouts store
    no columns: one 2x2 block per node and plane (the symmetric form: the diagonal
    plane and one plane per positive offset, its FP32 copy with an FP64 diagonal)."""
    n_nodes = n // 2
    if variant in (10, 11):
        # the whole Jacobi iteration: the symmetric planes, and per node r, q, p and the
        # Jacobi inverse read (4 x 16 B), r, p, q written (3 x 16 B), and x updated every
        # second launch for two iterations (x and the older p read, x written: 48 B per
        # two launches); per launch on average 136 B
        planes = layout["stencil_sym_planes"]
        matrix = n_nodes * (planes * 32 if value_bytes == 8 else (planes - 1) * 16 + 32)
        return matrix + n_nodes * 136
    if variant == 8:
        planes = layout["stencil_sym_planes"]
        matrix = n_nodes * (planes * 32 if value_bytes == 8 else (planes - 1) * 16 + 32)
        return matrix + 16 * n
    if variant == 9:
        return n_nodes * layout["stencil_exact_planes"] * 32 + 16 * n
    if variant in (6, 7, 12, 13, 14):
        n_slices = (n // 2 + 63) // 64
        matrix = nnz * value_bytes + 4 * (nnz // 4) + 4 * (n_slices + 1)
        if variant in (12, 14):
            # the whole Jacobi iteration as in 10 and 11 (the column nodes' r, q, p are
            # gathered from the XCD's L2; each counted once): 136 B per node
            return matrix + n_nodes * 136
        if variant == 13:
            # the vector launch (r, q, p, M^-1 read, r', p' written, x every second launch:
            # 120 B per node) and the multiply (p' gathered, r', M^-1 read, q' written: 64 B)
            return matrix + n_nodes * 184
        return matrix + (40 * n if variant == 7 else 16 * n)
    if variant == 15:
        # symmetric node slices: each node's blocks from the diagonal on hold values (32 B;
        # FP32 copy: 16 B, the diagonal entry zero, plus the compensated FP64 diagonal block,
        # 32 B per node) and a column, the blocks before it a column and the position of the
        # mirrored block (read from the L2, not counted), per slice two offsets; the vector
        # and multiply launches of variant 13 (184 B per node)
        n_slices = (n // 2 + 63) // 64
        blocks = nnz // 4
        upper = (blocks + n_nodes) // 2
        lower = blocks - upper
        values = upper * 36 if value_bytes == 8 else upper * 20 + n_nodes * 32
        return values + lower * 8 + 8 * (n_slices + 1) + n_nodes * 184
    return nnz * (value_bytes + 4) + 4 * (n + 1) + 16 * n


def measured_traffic(kernel: str, n_mesh: int):
    """HBM bytes per launch of `kernel` at this mesh size from a committed FETCH_SIZE /
    WRITE_SIZE counter summary (rocprofv3 --pmc, calibrated on a 1 GiB copy), or None."""
    prof = os.path.join(ROOT, "profiles")
    if not os.path.isdir(prof):
        return None
    key = kernel.replace(" ", "")
    for name in sorted(os.listdir(prof)):
        if not name.endswith(f"_{n_mesh}x{n_mesh}_traffic.md"):
            continue
        with open(os.path.join(prof, name)) as handle:
            for line in handle:
                parts = line.split()
                if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == key:
                    return int(parts[2])
    return None


def host_threads() -> int:
    """This host's CPU share: OMP_NUM_THREADS where the job sets it, else every core."""
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def parse_cli_line(stdout: str, tag: str) -> dict:
    """The numeric key=value fields of one `TAG: k=v, k=v` line of the CLI's stdout."""
    out = {}
    for line in stdout.splitlines():
        if line.startswith(tag + ": "):
            for kv in line[len(tag) + 2:].split(", "):
                key, _, value = kv.partition("=")
                try:
                    out[key] = float(value)
                except ValueError:
                    pass
            break
    return out


def cpu_protocol(runs: int = 7) -> dict:
    """The reference's measurement protocol (README.md:115-138: CLI `solve`, the TIMING
    line's element_stiffness_ms, assembly_ms and pcg_ms, median of `runs` runs in one
    batch) on C1 (50x50, serial, Jacobi) and C2 (200x200, RCM, coloring, Jacobi): the
    compiled reference CLI at threads = this host's share and threads = 1, and this
    repository's CLI (GPU backend) at the host's share."""
    from lightfem_amd import meshgen
    ref_cli = os.path.join(ROOT, "oracle", "_ref", "lightfem_ref")
    our_cli = os.path.join(ROOT, "lightfem_amd", "bin", "lightfem")
    if not (os.path.exists(ref_cli) and os.path.exists(our_cli)):
        return None
    keys = ("element_stiffness_ms", "assembly_ms", "pcg_ms", "matvec_ms", "process_ms")
    nproc = host_threads()
    out = {"runs": runs, "statistic": "median", "host_threads": nproc}
    with tempfile.TemporaryDirectory() as work:
        for name, make in (("c1_50x50_serial_jacobi", meshgen.cantilever_50),
                           ("c2_200x200_rcm_coloring_jacobi", meshgen.cantilever_200_rcm)):
            entry = {}
            for who, cli, threads in (("reference", ref_cli, nproc), ("reference", ref_cli, 1),
                                      ("gpu_cli", our_cli, nproc)):
                model = make()
                model.settings["threads"] = threads
                fem = os.path.join(work, f"{name}_{threads}.fem")
                model.to_fem(fem)
                samples, iterations = [], None
                for _ in range(runs):
                    res = subprocess.run([cli, "solve", fem, os.path.join(work, "out.vtk")], capture_output=True,
                                         text=True, timeout=600)
                    if res.returncode != 0:
                        break
                    samples.append(parse_cli_line(res.stdout, "TIMING"))
                    iterations = int(parse_cli_line(res.stdout, "SOLVED")["pcg_iterations"])
                if not samples:
                    entry[f"{who}_threads_{threads}"] = {"error": res.stderr.strip()[-200:]}
                    continue
                med = {k: float(np.median([smp[k] for smp in samples])) for k in keys}
                entry[f"{who}_threads_{threads}"] = dict(med, pcg_iterations=iterations)
                print(f"bench: cpu protocol {name} {who} threads {threads} {json.dumps(entry[f'{who}_threads_{threads}'])}",
                      file=sys.stderr, flush=True)
            out[name] = entry
    return out


def cpu_baseline(n_mesh: int, young: np.ndarray, iterations_full: int, sample_iters: int) -> dict:
    ref = os.path.join(ROOT, "oracle", "_ref", "lfref")
    threads = host_threads()
    if os.path.exists(ref):
        with tempfile.NamedTemporaryFile(suffix=".bin", delete=False) as handle:
            young.astype("<f8").tofile(handle)
            path = handle.name
        try:
            out = subprocess.run([ref, "bench", str(n_mesh), str(n_mesh), path, "plane_strain", "0.3",
                                  str(sample_iters), str(threads), "jacobi"], capture_output=True, text=True,
                                 timeout=600, check=True)
        finally:
            os.unlink(path)
        rec = json.loads(out.stdout.strip().splitlines()[-1])
        per_iter = rec["pcg_ms"] / max(1, rec["iterations"])
        return {"value": per_iter * iterations_full, "unit": "ms", "cores": rec["threads_used"],
                "kind": "reference",
                "sample": f"compiled reference pcg_solve, {rec['iterations']} Jacobi-PCG iterations of the same "
                          f"1000x1000 system ({rec['pcg_ms']:.0f} ms, matvec {rec['matvec_ms']:.0f} ms) scaled to "
                          f"{iterations_full} iterations, the reference's own count to ||r|| <= 1e-10 ||b|| on "
                          f"this system (its complete run is recorded in tests/golden/c3_oracle_1000.npz)",
                "ms_per_iteration": per_iter, "element_stiffness_ms": rec["element_stiffness_ms"],
                "assembly_ms": rec["assembly_ms"]}
    return None


def load_c3_oracle(n_mesh: int):
    """The FP64 oracle fixture of the 1000x1000 system (tests/golden/c3_oracle_1000.npz,
    written by tests/golden/gen_c3_oracle.py): a fixed subsample of u, R and the
    element-centre von Mises stress of the compiled reference's Jacobi-PCG solution and
    of a SuperLU direct solve."""
    path = os.path.join(ROOT, "tests", "golden", f"c3_oracle_{n_mesh}.npz")
    if not os.path.exists(path):
        return None
    with np.load(path) as data:
        out = {k: data[k] for k in data.files}
    out["meta"] = json.loads(out["meta"].tobytes().decode())
    return out


def oracle_errors(dev, u: np.ndarray, oracle: dict) -> dict:
    """Relative (norm-wise) differences of u, the reactions at the constrained dofs and the
    sampled element-centre von Mises stress of the device's last solve from both FP64
    oracles, on the fixture's subsample."""
    reactions = dev.reactions()[oracle["cdofs"]]
    centre, _, _ = dev.postprocess()
    vm = centre[oracle["elem_idx"], 6]
    us = u[oracle["dof_idx"]]
    out = {}
    for name in ("ref", "direct"):
        for key, mine in (("u", us), ("R", reactions), ("vm", vm)):
            theirs = oracle[f"{key}_{name}"]
            out[f"{key}_vs_{'reference_pcg' if name == 'ref' else 'direct'}"] = float(
                np.linalg.norm(mine - theirs) / np.linalg.norm(theirs))
    return out


def run_batch(dev, n_problems: int, tol: float, forms=("fp64", "mixed"), max_iterations: int = 1_000_000) -> dict:
    """The uneven batch: n_problems Q4 cantilevers (nx, ny ~ U{32..256}, 4 load cases each)
    solved together by the batched Jacobi-PCG to ||r|| <= tol ||b|| per system, with FP64
    values and mixed (each problem moves to compensated FP32 values at 1e-3 ||b||)."""
    from lightfem_amd import meshgen, pipeline
    t0 = time.perf_counter()
    specs = meshgen.batch_problems(n_problems)
    models = [meshgen.batch_problem_model(sp) for sp in specs]
    loads = [meshgen.batch_case_loads(m, sp) for m, sp in zip(models, specs)]
    bs = pipeline.batch_setup(models, loads, device=dev)
    setup_s = time.perf_counter() - t0
    out = {"workload": f"uneven-batch-{n_problems}-problems-x4-cases-seed{meshgen.SEED_BATCH}", "problems": n_problems,
           "systems": n_problems * 4, "dofs": int(sum(2 * ix.n_nodes for ix in bs.indexed)),
           "dofs_padded": int(dev.n), "nnz": int(dev.nnz), "setup_s": round(setup_s, 2),
           "stop_rule": f"recurrence ||r||/||b|| <= {tol:g} per system"}
    print(f"bench: batch {n_problems} problems, {dev.n} dofs, {dev.nnz} nnz, setup {setup_s:.1f} s", file=sys.stderr,
          flush=True)
    for form in forms:
        t = time.perf_counter()
        _, rep, summary = dev.batch_pcg(form, "relative", tol, max_iterations, raise_on_error=False)
        wall = (time.perf_counter() - t) * 1e3
        true_rel = rep["true_residual"] / np.maximum(rep["initial_residual"], 1e-300)
        out[form] = {"ms": round(wall, 2), "device_ms": round(summary["solve_ms"], 2),
                     "systems_per_s": round(n_problems * 4 / (summary["solve_ms"] * 1e-3), 1),
                     "systems_converged": summary["systems_converged"],
                     "iterations_max": int(rep["iterations"].max()),
                     "iterations_median": float(np.median(rep["iterations"])),
                     "iterations_total": int(rep["iterations"].sum()),
                     "true_relative_residual_max": float(true_rel.max()),
                     "problems_switched_to_fp32": summary["problems_switched"],
                     "graph_launches": summary["graph_launches"],
                     "launch_shape": {"block_threads": summary["block_threads"],
                                      "slices_per_wave": summary["slices_per_wave"],
                                      "tune_ms": round(summary["tune_ms"], 2)}}
        # the dominant kernel: the batch multiply (and its per-problem fold) with every problem
        # running, timed by the launch-shape tuner. Algorithmic bytes: the node-slice values
        # (FP64, or FP32 plus the FP64 compensated diagonal, 32 B per node), one column per 2x2
        # block, the slice offsets, and per node the 4 cases' p read and q written (128 B)
        nodes = dev.n // 2
        n_slices = (nodes + 63) // 64
        if form == "fp64" and summary.get("sym_slices"):
            # symmetric node slices: per node its blocks from the diagonal on (32 B of values
            # and a column each) and the blocks before it (a column and the mirrored block's
            # position: the mirrored values come from the L2), two offset arrays
            blocks = dev.nnz // 4
            upper = (blocks + nodes) // 2
            moved = upper * 36 + (blocks - upper) * 8 + 8 * (n_slices + 1) + 128 * nodes
        else:
            values = 8 * dev.nnz if form == "fp64" else 4 * dev.nnz + 32 * nodes
            moved = values + 4 * (dev.nnz // 4) + 4 * (n_slices + 1) + 128 * nodes
        ms = summary["multiply_ms"]
        out[form]["sym_slices"] = bool(summary.get("sym_slices"))
        out[form]["roofline"] = {"kernel": "batch_spmv_kernel + batch_spmv_finish_kernel", "bound": "hbm",
                                 "bytes_per_launch": int(moved), "avg_launch_ms": round(ms, 4),
                                 "achieved": round(moved / (ms * 1e-3) / 1e9, 1) if ms > 0 else None,
                                 "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                 "frac": round(moved / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms > 0 else None}
        print(f"bench: batch {form} {json.dumps(out[form])}", file=sys.stderr, flush=True)
    return out


PRECOND_CASES = (("jacobi", "jacobi", "csr"), ("block_1_node", "block", "csr"), ("block_rcb_64", "subdomain", "csr"),
                 ("ic0", "ic0", "csr"), ("jacobi_bsr", "jacobi", "bsr"))


def run_preconditioners_200(dev, tol: float, names=None, max_iterations: int = 100000) -> dict:
    """The 200x200 cantilever in RCM order with coloring assembly, every preconditioner and
    storage of the reference: Jacobi, 2x2 nodal block, the default block (rcb, 64 nodes per
    subdomain), IC(0), and Jacobi with block storage. Device solve time (hipEvents) to the
    reference stop rule ||r|| <= tol max(1, ||r0||)."""
    from lightfem_amd import meshgen, native, pipeline
    sys_ = pipeline.setup(meshgen.cantilever_200_rcm(), device=dev, assembly="coloring", reordering="rcm")
    nodes, count = native.subdomain_split(sys_.indexed.xy, "rcb", 64)
    dev.set_subdomains(nodes, count)
    out = {"workload": "q4-cantilever-200x200-plane-stress-rcm-coloring", "dofs": int(dev.n), "nnz": int(dev.nnz),
           "stop_rule": f"reference ||r|| <= {tol:g} max(1, ||r0||)"}
    for name, precond, storage in PRECOND_CASES:
        if names and name not in names:
            continue
        dev.pcg(precond, "fp64", "reference", tol, max_iterations, storage=storage,
                raise_on_error=False)  # set-up and warm-up
        _, rep = dev.pcg(precond, "fp64", "reference", tol, max_iterations, storage=storage, raise_on_error=False)
        out[name] = {"device_solve_ms": round(rep["solve_ms"], 3), "iterations": rep["iterations"],
                     "converged": bool(rep["converged"]), "setup_ms": round(rep["setup_ms"], 3),
                     "us_per_iteration": round(1e3 * rep["solve_ms"] / max(1, rep["iterations"]), 2)}
        if precond == "ic0":
            out[name]["sweep_levels"] = rep["levels"]  # forward + backward
        print(f"bench: preconditioners 200x200 {name} {json.dumps(out[name])}", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1000, help="mesh size (n x n elements)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-iters", type=int, default=2000)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-cpu-protocol", action="store_true",
                    help="skip the C1 / C2 CLI timings (reference protocol, median of --protocol-runs)")
    ap.add_argument("--protocol-runs", type=int, default=7)
    ap.add_argument("--variant", type=int, default=0)
    ap.add_argument("--skip-batch", action="store_true", help="skip the 256-problem batch workload")
    ap.add_argument("--batch-problems", type=int, default=256)
    ap.add_argument("--batch-only", action="store_true", help="only the batch workload (profiling)")
    ap.add_argument("--batch-forms", default="fp64,mixed", help="comma-separated batch precisions to run")
    ap.add_argument("--skip-precond", action="store_true", help="skip the 200x200 preconditioner comparison")
    ap.add_argument("--precond-only", action="store_true", help="only the 200x200 preconditioner comparison (profiling)")
    ap.add_argument("--precond-cases", default="", help="comma-separated subset of the 200x200 preconditioner cases")
    ap.add_argument("--assembly-only", action="store_true", help="only the assembly strategies (profiling)")
    ap.add_argument("--check-every", type=int, default=0)
    ap.add_argument("--max-iterations", type=int, default=0,
                    help="with --precond-only / --batch-only: cap every solve at this many iterations and end "
                         "with two 1 GiB copies (a short run for kernel traces and traffic counters)")
    ap.add_argument("--profile", type=int, default=0,
                    help="profiling run: this many FP64 SpMV launches and FP64 PCG iterations, nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the path is a single-device solve and --gpus {args.gpus} is not "
              f"supported (multi-GPU sharding is design-only).", file=sys.stderr)
        return 2

    from lightfem_amd import meshgen, native, pipeline

    if args.assembly_only:
        # the three assembly strategies of the 1000x1000 cantilever, five times each (profiling)
        dev = native.Device(0)
        pipeline.setup(meshgen.loguniform_cantilever(args.n), device=dev, assembly="coloring")
        out = {}
        for strategy in ("coloring", "atomic", "serial"):
            reps = [dev.assemble(strategy) for _ in range(5)]
            out[strategy] = {"element_stiffness_ms": float(np.median([r["element_stiffness_ms"] for r in reps])),
                             "assembly_ms": float(np.median([r["assembly_ms"] for r in reps]))}
        print(json.dumps({"assembly": out}))
        dev.close()
        return 0
    if args.precond_only:
        dev = native.Device(0)
        names = [c for c in args.precond_cases.split(",") if c]
        cap = args.max_iterations or 100000
        out = {"preconditioners_200x200": run_preconditioners_200(dev, args.tol, names, cap)}
        if args.max_iterations:
            out["copy_bytes"], out["copy_ms"] = 1 << 30, dev.bench_copy(1 << 30, 2)
        print(json.dumps(out))
        dev.close()
        return 0
    if args.batch_only:
        dev = native.Device(0)
        cap = args.max_iterations or 1_000_000
        out = {"batch": run_batch(dev, args.batch_problems, args.tol, args.batch_forms.split(","), cap)}
        if args.max_iterations:
            out["copy_bytes"], out["copy_ms"] = 1 << 30, dev.bench_copy(1 << 30, 2)
        print(json.dumps(out))
        dev.close()
        return 0
    n_mesh = args.n
    t0 = time.perf_counter()
    model = meshgen.loguniform_cantilever(n_mesh)
    dev = native.Device(0)
    print(f"bench: {n_mesh}x{n_mesh} plane-strain log-uniform E, setup ...", file=sys.stderr, flush=True)
    sys_ = pipeline.setup(model, device=dev, assembly="coloring")
    setup_s = time.perf_counter() - t0
    n, nnz = dev.n, dev.nnz
    # assembly strategies (device times from hipEvents; element matrices are formed inside
    # the assembly kernels). Algorithmic bytes: serial and coloring run the owner-computes
    # row kernel, which reads each node's (element, corner) list (16 B per element), the
    # element's nodes and material (20 B per element), each node's coordinates once, the
    # 64-entry map (256 B per element) and the row offsets, and writes every value once; the
    # atomic scatter reads the element data and the map and adds into each value (8 B read +
    # 8 B written per map entry) after zeroing K.
    n_el, n_nd = model.enodes.shape[0], n // 2
    rows_bytes = (16 + 20 + 256) * n_el + 16 * n_nd + 4 * (n + 1) + 8 * nnz
    scatter_bytes = 20 * n_el + 16 * n_nd + 64 * n_el * (4 + 16) + 8 * nnz
    assembly = {}
    for strategy in ("coloring", "atomic", "serial"):
        dev.assemble(strategy)  # warm-up
        rep = dev.assemble(strategy)
        moved = scatter_bytes if strategy == "atomic" else rows_bytes
        assembly[strategy] = {"assembly_ms": rep["assembly_ms"], "bytes": moved,
                              "GBs": round(moved / (rep["assembly_ms"] * 1e-3) / 1e9, 1)}
    dev.set_load(sys_.indexed.load)
    dev.dirichlet(sys_.indexed.cdofs, sys_.indexed.cvals)
    layout = dev.layout_info()
    # the library's default Jacobi iteration: fused stencil, stencil multiply, node slices
    default_variant = 11 if layout["grid_row_len"] else (
        10 if layout["fused_window_pairs"] else (8 if layout["stencil_sym_planes"] else (
            14 if layout["n_slices"] <= 1024 else 13)))
    if args.profile:
        # Profiling run: the FP64 SpMV of the headline variant alone (args.profile launches),
        # then FP64 Jacobi-PCG iterations (a whole number of captured graphs), no other kernels.
        variant = args.variant or default_variant  # the multiply the default PCG iteration runs
        ms = dev.bench_spmv("fp64", variant, args.profile)
        b = spmv_bytes(n, nnz, 8, variant, layout)
        prof = {"kernel": kernel_name(variant, layout), "launches": args.profile, "avg_launch_ms": ms, "bytes": b,
                "GBs": b / (ms * 1e-3) / 1e9}
        print(f"bench: profile spmv {json.dumps(prof)}", file=sys.stderr, flush=True)
        per_graph = args.check_every or 32
        iters = -(-args.profile // per_graph) * per_graph
        _, rep = dev.pcg("jacobi", "fp64", "relative", args.tol, iters, args.check_every, spmv_variant=variant,
                         raise_on_error=False)
        prof["pcg_iterations"] = rep["iterations"]
        prof["pcg_ms_per_iteration"] = rep["solve_ms"] / max(1, rep["iterations"])
        copy_bytes = 1 << 30  # calibration dispatches for the traffic counters: 1 GiB read + 1 GiB written
        prof["copy_bytes"] = copy_bytes
        prof["copy_ms"] = dev.bench_copy(copy_bytes, 2)
        print(json.dumps(prof))
        dev.close()
        return 0
    # SpMV kernels: average launch duration over 50 launches (hipEvents on the library
    # stream). The solves use the library's default, variant 8 (node stencil, symmetric
    # form) where the pattern admits it, unless --variant says otherwise.
    solve_variant = args.variant or default_variant
    spmv, sweep, iteration_kernel = {}, {}, {}
    for precision, vbytes in (("fp64", 8), ("mixed", 4)):
        sweep[precision] = {}
        for variant in (1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15):
            if (variant in (6, 7, 12, 13, 14) and not layout["node_blocked"]) or (
                    variant == 15 and not layout["sym_slices"]) or (
                    variant == 14 and layout["n_slices"] > 1024) or (
                    variant == 8 and not layout["stencil_sym_planes"]) or (variant == 9 and (
                    precision == "mixed" or not layout["stencil_exact_planes"])) or (
                    variant == 10 and not layout["fused_window_pairs"]) or (
                    variant == 11 and not layout["grid_row_len"]):
                continue
            ms = dev.bench_spmv(precision, variant, 50)
            b = spmv_bytes(n, nnz, vbytes, variant, layout)
            sweep[precision][variant] = {"kernel": kernel_name(variant, layout), "ms": ms, "bytes": b,
                                         "GBs": b / (ms * 1e-3) / 1e9}
        # the multiply alone (the stand-alone kernel of the solve's layout) and the kernel the
        # PCG iteration spends its time in (the same, or the fused iteration)
        multiply = 8 if solve_variant in (10, 11) else (6 if solve_variant in (12, 13, 14) else solve_variant)
        spmv[precision] = dict(sweep[precision][multiply], variant=multiply)
        iteration_kernel[precision] = dict(sweep[precision][solve_variant], variant=solve_variant)
    print(f"bench: spmv {json.dumps(spmv)}", file=sys.stderr, flush=True)
    print(f"bench: iteration kernel {json.dumps(iteration_kernel)}", file=sys.stderr, flush=True)
    copy_bytes = 1 << 30
    copy_ms = dev.bench_copy(copy_bytes, 20)
    copy = {"bytes_per_copy": copy_bytes, "ms": copy_ms, "GBs": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9}
    print(f"bench: device copy {json.dumps(copy)}", file=sys.stderr, flush=True)
    # streaming calibration at the dominant kernel's launch size: four buffers read and one
    # written, 16-B pairs, the same read / write mix and bytes as one fused iteration
    stream_bytes = iteration_kernel["fp64"]["bytes"] // 5 // 16 * 16
    stream_ms = dev.bench_stream(stream_bytes, 50)
    stream = {"bytes_per_launch": 5 * stream_bytes, "ms": stream_ms,
              "GBs": 5 * stream_bytes / (stream_ms * 1e-3) / 1e9}
    print(f"bench: stream calibration {json.dumps(stream)}", file=sys.stderr, flush=True)
    forms = {}
    cap = 10 ** 6
    oracle = load_c3_oracle(n_mesh)
    for form in ("fp64", "mixed"):
        variant = solve_variant
        steps, warmup = args.steps, args.warmup  # both forms timed alike
        for _ in range(warmup):
            dev.pcg("jacobi", form, "relative", args.tol, cap, args.check_every, spmv_variant=variant,
                    raise_on_error=False)
            print(f"bench: {form} warm-up solve done", file=sys.stderr, flush=True)
        _sync()
        t = time.perf_counter()
        reps = []
        for _ in range(steps):
            u, rep = dev.pcg("jacobi", form, "relative", args.tol, cap, args.check_every, spmv_variant=variant,
                             raise_on_error=False)
            reps.append(rep)
            print(f"bench: {form} solve {rep['iterations']} iterations, converged={rep['converged']}", file=sys.stderr,
                  flush=True)
        _sync()
        wall = (time.perf_counter() - t) * 1e3 / steps
        if form == "fp64":
            cap = 2 * reps[-1]["iterations"]  # the mixed forms get at most twice the FP64 count
        forms[form] = {"ms_per_solve": wall, "device_solve_ms": float(np.median([r["solve_ms"] for r in reps])),
                       "iterations": reps[-1]["iterations"], "converged": bool(reps[-1]["converged"]),
                       "refine_passes": reps[-1]["refine_passes"],
                       "true_relative_residual": reps[-1]["true_residual"] / reps[-1]["initial_residual"],
                       "spmv_variant": variant, "steps": steps, "warmup": warmup, "u": u}
        if oracle is not None:
            forms[form]["error_vs_fp64_oracles"] = oracle_errors(dev, u, oracle)
            print(f"bench: {form} against the FP64 oracles {json.dumps(forms[form]['error_vs_fp64_oracles'])}",
                  file=sys.stderr, flush=True)
        print(f"bench: {form} {wall:.1f} ms/solve, {reps[-1]['iterations']} iterations", file=sys.stderr, flush=True)
    # FP32 values from the first iteration (residual replacement only), capped at 2x the FP64 count
    t = time.perf_counter()
    u32, rep32 = dev.pcg("jacobi", "mixed", "relative", args.tol, cap, args.check_every,
                         spmv_variant=solve_variant, mixed_switch=-1.0, raise_on_error=False)
    forms["fp32_values_throughout"] = {
        "ms_per_solve": (time.perf_counter() - t) * 1e3, "device_solve_ms": rep32["solve_ms"],
        "iterations": rep32["iterations"], "converged": bool(rep32["converged"]),
        "refine_passes": rep32["refine_passes"],
        "true_relative_residual": rep32["true_residual"] / rep32["initial_residual"], "iteration_cap": cap}
    if oracle is not None:
        forms["fp32_values_throughout"]["error_vs_fp64_oracles"] = oracle_errors(dev, u32, oracle)
    print(f"bench: fp32-throughout {forms['fp32_values_throughout']}", file=sys.stderr, flush=True)
    u64, umx = forms["fp64"].pop("u"), forms["mixed"].pop("u")
    forms["mixed"]["rel_u_diff_vs_fp64"] = float(np.linalg.norm(umx - u64) / np.linalg.norm(u64))
    forms["fp32_values_throughout"]["rel_u_diff_vs_fp64"] = float(np.linalg.norm(u32 - u64) / np.linalg.norm(u64))
    if solve_variant != 6 and layout["node_blocked"]:
        # FP64 with the node-slice multiply, whose rows sum exactly as the reference's CSR loop
        t = time.perf_counter()
        u6, rep6 = dev.pcg("jacobi", "fp64", "relative", args.tol, 10 ** 6, args.check_every, spmv_variant=6,
                           raise_on_error=False)
        forms["fp64_node_slices"] = {
            "ms_per_solve": (time.perf_counter() - t) * 1e3, "device_solve_ms": rep6["solve_ms"],
            "iterations": rep6["iterations"], "converged": bool(rep6["converged"]), "spmv_variant": 6,
            "true_relative_residual": rep6["true_residual"] / rep6["initial_residual"],
            "rel_u_diff_vs_fp64": float(np.linalg.norm(u6 - u64) / np.linalg.norm(u64))}
        print(f"bench: fp64 node slices {forms['fp64_node_slices']}", file=sys.stderr, flush=True)
        # the same layout with the iteration fused into two launches (variant 13: the default
        # for node-blocked patterns without fixed offsets at this size, e.g. RCM numbering)
        t = time.perf_counter()
        u13, rep13 = dev.pcg("jacobi", "fp64", "relative", args.tol, 10 ** 6, args.check_every, spmv_variant=13,
                             raise_on_error=False)
        forms["fp64_node_slices_fused"] = {
            "ms_per_solve": (time.perf_counter() - t) * 1e3, "device_solve_ms": rep13["solve_ms"],
            "iterations": rep13["iterations"], "converged": bool(rep13["converged"]), "spmv_variant": 13,
            "true_relative_residual": rep13["true_residual"] / rep13["initial_residual"],
            "rel_u_diff_vs_fp64": float(np.linalg.norm(u13 - u64) / np.linalg.norm(u64))}
        print(f"bench: fp64 node slices fused {forms['fp64_node_slices_fused']}", file=sys.stderr, flush=True)
    if solve_variant != 15 and layout["sym_slices"]:
        # the symmetric node slices (variant 15: the FP64 default for large node-blocked
        # patterns without fixed offsets, e.g. RCM numbering): lower blocks read as the
        # transposes of the column nodes' blocks
        t = time.perf_counter()
        u15, rep15 = dev.pcg("jacobi", "fp64", "relative", args.tol, 10 ** 6, args.check_every, spmv_variant=15,
                             raise_on_error=False)
        forms["fp64_sym_slices"] = {
            "ms_per_solve": (time.perf_counter() - t) * 1e3, "device_solve_ms": rep15["solve_ms"],
            "iterations": rep15["iterations"], "converged": bool(rep15["converged"]), "spmv_variant": 15,
            "true_relative_residual": rep15["true_residual"] / rep15["initial_residual"],
            "rel_u_diff_vs_fp64": float(np.linalg.norm(u15 - u64) / np.linalg.norm(u64))}
        print(f"bench: fp64 symmetric node slices {forms['fp64_sym_slices']}", file=sys.stderr, flush=True)
        # mixed precision on the same layout (its FP32 copy with the compensated diagonal),
        # the default mixed iteration of such patterns
        t = time.perf_counter()
        um15, repm15 = dev.pcg("jacobi", "mixed", "relative", args.tol, 10 ** 6, args.check_every, spmv_variant=15,
                               raise_on_error=False)
        forms["mixed_sym_slices"] = {
            "ms_per_solve": (time.perf_counter() - t) * 1e3, "device_solve_ms": repm15["solve_ms"],
            "iterations": repm15["iterations"], "converged": bool(repm15["converged"]), "spmv_variant": 15,
            "refine_passes": repm15["refine_passes"],
            "true_relative_residual": repm15["true_residual"] / repm15["initial_residual"],
            "rel_u_diff_vs_fp64": float(np.linalg.norm(um15 - u64) / np.linalg.norm(u64)),
            "speedup_vs_fp64_sym_slices": rep15["solve_ms"] / max(repm15["solve_ms"], 1e-9)}
        print(f"bench: mixed symmetric node slices {forms['mixed_sym_slices']}", file=sys.stderr, flush=True)
    per_iter_ms = {k: v["device_solve_ms"] / max(1, v["iterations"]) for k, v in forms.items()}
    cpu, protocol = None, None
    if not args.skip_cpu:
        ref_iterations = forms["fp64"]["iterations"]
        if oracle is not None and oracle["meta"].get("reference_pcg", {}).get("iterations"):
            ref_iterations = int(oracle["meta"]["reference_pcg"]["iterations"])
        try:
            cpu = cpu_baseline(n_mesh, sys_.indexed.materials[sys_.indexed.elem_mat][:, 0], ref_iterations,
                               args.cpu_iters)
            if oracle is not None:
                cpu["reference_complete_run_recorded"] = {
                    k: oracle["meta"]["reference_pcg"][k] for k in ("iterations", "pcg_ms", "matvec_ms",
                                                                    "threads_used")}
                cpu["reference_complete_run_recorded"]["where"] = "the build container's cores, not this host"
        except Exception as exc:  # reported, not fatal
            print(f"bench: cpu baseline failed: {exc}", file=sys.stderr)
        if not args.skip_cpu_protocol:
            try:
                protocol = cpu_protocol(args.protocol_runs)
            except Exception as exc:  # reported, not fatal
                print(f"bench: cpu protocol failed: {exc}", file=sys.stderr)
    batch = None
    if not args.skip_batch:
        batch = run_batch(dev, args.batch_problems, args.tol, args.batch_forms.split(","))
    precond200 = None if args.skip_precond else run_preconditioners_200(dev, args.tol)
    # Headline: the FP64 form. The mixed form (FP32 values with a compensated FP64
    # diagonal once ||r|| <= 1e-2 ||r0||, FP64 vectors and sums) stops on the same
    # recurrence rule and is reported beside it with its displacement difference.
    value = forms["fp64"]["ms_per_solve"]
    dominant = iteration_kernel["fp64"]
    mixed_ok = forms["mixed"]["converged"]
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(value, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"q4-cantilever-{n_mesh}x{n_mesh}-plane-strain-loguniformE-1e-3..1-seed"
                               f"{meshgen.SEED_LOGE}", "dofs": n, "nnz": nnz, "preconditioner": "jacobi",
                   "stop_rule": f"recurrence relative residual ||r||/||b|| <= {args.tol:g}", "assembly": "coloring",
                   "form": "fp64"},
        "fp64_ms": round(forms["fp64"]["ms_per_solve"], 3),
        "mixed_ms": round(forms["mixed"]["ms_per_solve"], 3) if mixed_ok else None,
        "mixed_converged": mixed_ok,
        "mixed_speedup_vs_fp64": round(forms["fp64"]["ms_per_solve"] / forms["mixed"]["ms_per_solve"], 4)
        if mixed_ok else None,
        "forms": forms, "per_iteration_ms": per_iter_ms,
        "spmv": spmv, "iteration_kernel": iteration_kernel,
        "spmv_variants": {p: {str(k): round(v["GBs"], 1) for k, v in d.items()} for p, d in sweep.items()},
        "device_copy": copy, "stream_calibration": stream, "assembly": assembly, "setup_s": round(setup_s, 2),
        "roofline": {"bound": "hbm", "kernel": f"{dominant['kernel']} (FP64 values)",
                     "achieved": round(dominant["GBs"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(dominant["GBs"] / HBM_PEAK_GBS, 4),
                     "traffic": measured_traffic(dominant["kernel"], n_mesh),
                     "bytes_per_launch": dominant["bytes"], "avg_launch_ms": dominant["ms"],
                     "spmv_kernel": spmv["fp64"]["kernel"], "spmv_GBs": round(spmv["fp64"]["GBs"], 1),
                     "spmv_frac": round(spmv["fp64"]["GBs"] / HBM_PEAK_GBS, 4),
                     "fp32_values_spmv_GBs": round(spmv["mixed"]["GBs"], 1),
                     "fp32_values_iteration_GBs": round(iteration_kernel["mixed"]["GBs"], 1),
                     "stream_same_bytes_GBs": round(stream["GBs"], 1),
                     "frac_of_stream": round(dominant["GBs"] / stream["GBs"], 4),
                     "runtime_copy_GBs": round(copy["GBs"], 1)},
        "cpu_baseline": cpu,
        "cpu_protocol_c1_c2": protocol,
        "batch": batch,
        "preconditioners_200x200": precond200,
    }
    print(json.dumps(line))
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
