"""Benchmark of the Born operator pair on one MI355X.

A step is one forward + adjoint pair of the operator on the config-1
workload (128-element array, 512x512 FOV at pitch/4, 257 frequencies,
11 plane waves in one launch, identity receive mixing, one right-hand
side).  The tables are built on the device before the timed region; inputs
are resident in HBM.  Prints one JSON line (see the keys below).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision split_bf16]
                    [--no-compare] [--no-cpu]

`value` is pairs/s of the selected form (split BF16 by default); the plain
FP32 form is timed beside it on the same plan and input, and the ratio and
the relative L2 difference between the two forms' outputs are reported.
`err_vs_fp64` gives each form's relative L2 and max-abs error against the
FP64 oracle on a dense input at 3 of the frequencies (outside the timed
region).  `cpu_baseline` times the oracle's C FP64 loop nest (Phi
materialized, as the reference's CPU path) on a sample of the config-1 pair,
with the fused loop beside it; `config0_full` times config 0 in full on both
the CPU loops and the device.
"""
import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "fwd+adj operator pairs/s vs plain FP32 at ≤1e-4 rel L2 err vs FP64 oracle"
PEAK = {"fp32": (157.3, "TFLOP/s"), "split_bf16": (2500.0, "TFLOP/s"), "split_fp16": (2500.0, "TFLOP/s")}


def algorithmic_flops(spec, n_rhs: int) -> float:
    """F_fwd = 8 N_meas N_f N_e N_pts N_rhs + 6 N_meas N_f N_pts N_rhs
    (+ 8 N_meas N_f N_mix N_act N_rhs for non-identity mixing); F_adj = F_fwd."""
    n_meas = spec.N_measurements
    n_f = len(spec.frequencies)
    n_e = spec.array.N_elements
    n_pts = spec.fov.N_points
    f = 8.0 * n_meas * n_f * n_e * n_pts * n_rhs + 6.0 * n_meas * n_f * n_pts * n_rhs
    mixing_identity = all(len(mx.indices_active) == 1 for mm in spec.mixes for mx in mm)
    if not mixing_identity:
        f += sum(8.0 * n_f * len(mx.indices_active) * n_rhs for mm in spec.mixes for mx in mm)
    return f


def algorithmic_bytes(spec, n_rhs: int) -> float:
    """B = 8 N_f N_pts (table) + 8 N_meas N_f N_pts (p_inc) + |gamma| + |u|
    in FP32 complex."""
    n_f = len(spec.frequencies)
    n_pts = spec.fov.N_points
    return 8.0 * n_f * n_pts + 8.0 * spec.N_measurements * n_f * n_pts + 8.0 * n_pts * n_rhs + \
        8.0 * spec.N_observations() * n_rhs


def _cpu_tables(op, spec, n_mixes: int):
    """FP64 copies of the device tables of measurement 1 and its first
    `n_mixes` mixes, in the oracle's table layout."""
    from oracle import born_oracle as O

    class _T:
        pass

    class _M:
        pass

    T = _T()
    T.h_ref = op.table("h_ref").astype(np.complex128)
    T.shift = O.shift_lateral(spec)
    ms = _M()
    ms.idx_f_to_global = spec.measurement_frequencies(0) + 1
    ms.p_incident = op.table("p_incident", 0).astype(np.complex128)
    fidx = spec.measurement_frequencies(0)
    pref_u = O.prefactors_unique(spec)
    ms.mixes = []
    for mix in spec.mixes[0][:n_mixes]:
        mt = _M()
        mt.idx_f_to_meas = np.arange(len(fidx)) + 1
        mt.idx_active = np.asarray(mix.indices_active) + 1
        mt.prefactors = pref_u[fidx][:, None] * np.asarray(mix.weights)[None, :]
        ms.mixes.append(mt)
    T.meas = [ms]
    return T


def _cpu_pair_seconds(T, g, fused: bool, threads: int) -> float:
    from oracle import ref_loop as R
    P = R.Prepared(T, 0)
    g = np.ascontiguousarray(g, dtype=np.complex128)
    t0 = time.perf_counter()
    u = R.apply_measurement(P, 1, g, fused=fused, threads=threads)
    R.apply_measurement(P, 2, u, fused=fused, threads=threads)
    return time.perf_counter() - t0


def host_cpu() -> dict:
    """CPU model and logical CPU count of the host that runs the CPU legs."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count()}


def cpu_baseline(op, spec, n_mixes: int, threads: int):
    """The oracle's C restatement of forward_quick_cpu / adjoint_quick_cpu
    (Phi materialized per active element, then a GEMV) on measurement 1 and
    the first `n_mixes` receive channels, scaled to the full pair; beside it
    the oracle's fused loop (no Phi) on the same sample."""
    T = _cpu_tables(op, spec, n_mixes)
    g = spec_gamma(spec)
    dt = _cpu_pair_seconds(T, g, False, threads)
    dt_f = _cpu_pair_seconds(T, g, True, threads)
    n_units = sum(len(mm) for mm in spec.mixes)          # (measurement, mix) pairs in a full pair
    scale = n_units / n_mixes
    return {
        "value": 1.0 / (dt * scale), "unit": "pairs/s", "cores": threads, "kind": "port",
        "sample": f"C FP64 loop nest (operator_born.m:332-596, Phi materialized) on measurement 1, "
                  f"{n_mixes} of {len(spec.mixes[0])} receive channels, forward+adjoint in {dt:.1f} s, "
                  f"scaled x{scale:.0f} to the full pair",
        "fused": {"value": 1.0 / (dt_f * scale), "unit": "pairs/s",
                  "sample": f"the oracle's fused FP64 loop (no Phi) on the same sample in {dt_f:.1f} s"},
        "host": host_cpu(),
    }


def host_entry_rate(op, form, x0, pairs: int = 3):
    """Pairs/s through the host-buffer entry points (born_forward /
    born_adjoint: FP64 host arrays, conversions and PCIe transfers
    included), timed after the device legs, outside their timed region."""
    op.set_precision(form)
    xh = np.asarray(x0, dtype=np.complex128).reshape(-1, 1)
    op.adjoint(op.forward(xh))   # warm-up
    t0 = time.perf_counter()
    for _ in range(pairs):
        op.adjoint(op.forward(xh))
    dt = (time.perf_counter() - t0) / pairs
    return {"precision": form, "pairs": pairs, "ms_per_pair": dt * 1e3, "pairs_per_s": 1.0 / dt,
            "note": "host FP64 buffers: conversions and PCIe transfers of gamma and u included"}


def err_vs_fp64(op, spec, forms, f_sel=(0, 128, 256)):
    """The checker: error of each numeric form against the FP64 oracle on a
    dense input over the whole FOV, at 3 of the frequencies (first, middle,
    last) -- the forward of the speckle phantom at those frequencies'
    observations and the adjoint of random u supported on them -- through
    the same device entry points and kernels as the timed pairs.  Runs
    outside the timed region."""
    import torch
    from oracle import born_oracle as O
    t0 = time.perf_counter()
    f_sel = [i for i in f_sel if i < len(spec.frequencies)]
    T, rows = O.frequency_tables(spec, np.asarray(f_sel))
    g = spec_gamma(spec)
    rng = np.random.Generator(np.random.PCG64(spec.seed + 200))
    u_sub = rng.standard_normal(len(rows)) + 1j * rng.standard_normal(len(rows))
    u_ref = O.forward_by_frequency(T, g)
    g_ref = O.adjoint_by_frequency(T, u_sub)
    t_oracle = time.perf_counter() - t0
    x = torch.tensor(g.astype(np.complex64), device="cuda")
    uu = torch.zeros(op.N_observations, dtype=torch.complex64, device="cuda")
    uu[torch.from_numpy(rows).cuda()] = torch.from_numpy(u_sub.astype(np.complex64)).cuda()
    out = {"sample": f"dense speckle phantom over all {spec.fov.N_points} points; frequencies {list(f_sel)} "
                     f"({len(rows)} observations); FP64 oracle in {t_oracle:.1f} s"}
    for form in forms:
        op.set_precision(form)
        u = op.forward_dev(x).cpu().numpy()[rows]
        gh = op.adjoint_dev(uu).cpu().numpy()
        out[form] = {"forward": O.rel_l2(u, u_ref), "adjoint": O.rel_l2(gh, g_ref),
                     "max_abs": {"forward": O.max_abs(u, u_ref), "adjoint": O.max_abs(gh, g_ref)},
                     "max_abs_rel": {"forward": O.max_abs(u, u_ref) / float(np.max(np.abs(u_ref))),
                                     "adjoint": O.max_abs(gh, g_ref) / float(np.max(np.abs(g_ref)))}}
    return out


def config0_full(threads: int, steps: int):
    """Config 0 in full (64 elements, 256x256 FOV, 65 bins, one wave, 30
    occupied points): the C FP64 loop nest, Phi materialized and fused, and
    the device pair in both forms, all on the same input."""
    import torch
    from comp_ui_toolbox_amd import OperatorBorn
    from comp_ui_toolbox_amd import configs as C
    spec = C.config0()
    op = OperatorBorn.from_sequence(spec, precision="fp32")
    g0 = C.phantom_points(spec)
    T = _cpu_tables(op, spec, len(spec.mixes[0]))
    dt = _cpu_pair_seconds(T, g0, False, threads)
    dt_f = _cpu_pair_seconds(T, g0, True, threads)
    x = torch.tensor(g0.astype(np.complex64)[None, :], device="cuda")
    u = torch.empty((1, op.N_observations), dtype=torch.complex64, device="cuda")
    g = torch.empty((1, op.N_points), dtype=torch.complex64, device="cuda")
    gpu = {}
    for form in ("split_bf16", "fp32"):
        el, _ = time_form(op, form, x, u, g, steps, 3)
        gpu[form] = steps / el
    op.close()
    return {"workload": "config0: 64 el, 256x256 FOV, 65 f, 1 wave, 4160 obs, 30 occupied points, 1 rhs",
            "cpu_pairs_per_s": 1.0 / dt, "cpu_fused_pairs_per_s": 1.0 / dt_f, "cores": threads,
            "gpu_pairs_per_s": gpu}


_GAMMA = {}


def spec_gamma(spec):
    """The workload's own phantom (configs.phantom): config 0, 30 occupied
    points; configs 1 and 2, speckle with 6 disks and 20 bright points;
    config 4, speckle with 4 anechoic spheres."""
    from comp_ui_toolbox_amd import configs as C
    if spec.name not in _GAMMA:
        _GAMMA[spec.name] = C.phantom(spec)
    return _GAMMA[spec.name]


def time_form(op, precision, x, u, g, steps, warmup):
    """Warm up, then time `steps` forward+adjoint pairs of one numeric form.
    x (N_rhs, N_pts), u (N_rhs, N_obs), g (N_rhs, N_pts).  Returns wall
    seconds and the per-launch kernel milliseconds."""
    import torch
    op.set_precision(precision)
    for _ in range(warmup):
        op.forward_dev(x, out=u)
        op.adjoint_dev(u, out=g)
    torch.cuda.synchronize()
    op.get_timing()
    op.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        op.forward_dev(x, out=u)
        op.adjoint_dev(u, out=g)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    op.set_timing(False)
    timing = op.get_timing()
    kern = {name: timing[name][0] / max(timing[name][1], 1) for name in ("forward", "adjoint")}
    print(f"[bench] {precision}: {steps} pairs in {elapsed * 1e3:.1f} ms, kernels "
          f"fwd {kern['forward']:.3f} ms adj {kern['adjoint']:.3f} ms", file=sys.stderr, flush=True)
    return elapsed, kern


def roofline(F, kern, precision, config):
    dom = max(kern, key=kern.get)
    peak, unit = PEAK[precision]
    achieved = F / (kern[dom] * 1e-3) / 1e12
    traffic = None
    tfile = ROOT / "profiles" / f"{config}_hbm_traffic.md"
    if tfile.exists():
        # table rows: | precision | kernel | launch | fetch | write | bytes_per_launch | ratio |
        for line in tfile.read_text().splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) == 7 and cells[0] == precision and cells[1] == dom:
                try:
                    traffic = float(cells[5])
                except ValueError:
                    traffic = None
    r = {"bound": "mfma", "kernel": f"{dom} contraction ({precision})", "achieved": achieved, "peak": peak,
         "unit": unit, "frac": achieved / peak, "traffic": traffic}
    if precision != "fp32":
        # hi*hi + hi*lo + lo*hi per real product: 12 MFMAs per complex
        # (tile, K-block) in the adjoints and the 3-D slice forward (3x the
        # algorithmic flops), 9 in the 2-D forward's three-product form (2.25x)
        factor = 2.25 if dom == "forward" and config != "config4" else 3.0
        r["mfma_work_frac"] = factor * achieved / peak
    return r


def rel_l2(a, b):
    import torch
    a = a.to(torch.complex128)
    b = b.to(torch.complex128)
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))


def copy_probe(mib: int):
    """A plain copy kernel of known size: its FETCH_SIZE / WRITE_SIZE
    counters give the factor that turns counter readings into bytes."""
    import torch
    n = mib * 2**20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(f"[bench] copy probe: 5 copies of {n} bytes", file=sys.stderr, flush=True)


def bench_batched(args):
    """Config 3: 24 operator instances of uneven size (FOV side 128-512 at
    pitch/4, 33-257 bins, 1-11 plane waves, config-1 array) applied as one
    batch: a few persistent contraction launches per direction (one per group
    of instances that take the same kernel form), replayed from a hipGraph,
    tile shapes chosen by the autotuner.  A step is the forward and
    the adjoint of every instance."""
    import torch
    from comp_ui_toolbox_amd import OperatorBorn, OperatorBatch
    from comp_ui_toolbox_amd import configs as C

    specs = C.config3_instances()
    prec = "fp32" if args.precision == "split_bf16" and not args.batched_split else args.precision
    t0 = time.perf_counter()
    ops = [OperatorBorn.from_sequence(s, precision=prec) for s in specs]
    torch.cuda.synchronize()
    t_tables = time.perf_counter() - t0
    B = OperatorBatch(ops)
    tf, ta = B.autotune()
    print(f"[bench] config3: {len(ops)} instances, tables {t_tables:.2f} s, autotuned contraction "
          f"fwd {tf:.3f} ms adj {ta:.3f} ms", file=sys.stderr, flush=True)
    rng = np.random.Generator(np.random.PCG64(4))
    xs = [torch.tensor(C.phantom_points(s, seed=int(rng.integers(1 << 30))).astype(np.complex64), device="cuda")
          for s in specs]
    us = [torch.empty(op.N_observations, dtype=torch.complex64, device="cuda") for op in ops]
    gs = [torch.empty(op.N_points, dtype=torch.complex64, device="cuda") for op in ops]
    B.set_graph(True)
    for _ in range(args.warmup):
        B.forward_dev(xs, outs=us)
        B.adjoint_dev(us, outs=gs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        B.forward_dev(xs, outs=us)
        B.adjoint_dev(us, outs=gs)
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    # kernel times of the batched contractions (eager launches, event timing)
    B.set_timing(True)
    for _ in range(3):
        B.forward_dev(xs, outs=us)
        B.adjoint_dev(us, outs=gs)
    torch.cuda.synchronize()
    tm = B.get_timing()
    B.set_timing(False)
    kern = {k: v[0] / max(v[1], 1) for k, v in tm.items()}
    F = sum(algorithmic_flops(sp, 1) for sp in specs)
    Bb = sum(algorithmic_bytes(sp, 1) for sp in specs)
    per_instance = None
    if args.per_instance:
        # every instance alone (its own launches, event timing), beside the batch
        per_instance = []
        for sp, op, x in zip(specs, ops, xs):
            u1 = torch.empty(op.N_observations, dtype=torch.complex64, device="cuda")
            g1 = torch.empty(op.N_points, dtype=torch.complex64, device="cuda")
            for _ in range(2):
                op.forward_dev(x, out=u1)
                op.adjoint_dev(u1, out=g1)
            torch.cuda.synchronize()
            op.get_timing()
            op.set_timing(True)
            for _ in range(3):
                op.forward_dev(x, out=u1)
                op.adjoint_dev(u1, out=g1)
            torch.cuda.synchronize()
            op.set_timing(False)
            t1 = op.get_timing()
            f_ms = t1["forward"][0] / max(t1["forward"][1], 1)
            a_ms = t1["adjoint"][0] / max(t1["adjoint"][1], 1)
            Fi = algorithmic_flops(sp, 1)
            per_instance.append({"name": sp.name, "fov": sp.fov.N_points_axis[0], "n_f": len(sp.frequencies),
                                 "waves": sp.N_measurements, "gflop": Fi / 1e9, "forward_ms": f_ms,
                                 "adjoint_ms": a_ms, "forward_tflops": Fi / f_ms / 1e9,
                                 "adjoint_tflops": Fi / a_ms / 1e9})
    result = {
        "metric": "batched operator pairs/s (24 uneven instances, forward+adjoint of all per pair)",
        "value": args.steps / elapsed,
        "unit": "batch pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": prec,
        "data": "synthetic (config-3 instances, 30-point phantoms; tables built on the device)",
        "config": {
            "workload": "config3: 24 instances, FOV side {128,256,384,512} at pitch/4, 33-257 f, 1-11 waves, "
                        "128 el, 1 rhs each; persistent launches per kernel form and direction, hipGraph replay, autotuned",
            "precision": prec,
            "instance_pairs_per_s": len(ops) * args.steps / elapsed,
            "table_build_s": round(t_tables, 3),
            "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
            "flops_per_pass": F,
            "bytes_per_pass": Bb,
        },
        "roofline": roofline(F, kern, prec, "config3"),
        "cpu_baseline": None,
    }
    if per_instance is not None:
        result["per_instance"] = per_instance
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="split_bf16", choices=list(PEAK),
                    help="the form whose pairs/s is reported as value")
    ap.add_argument("--no-compare", action="store_true",
                    help="do not time the plain FP32 form beside the selected one")
    ap.add_argument("--config", default="config1")
    ap.add_argument("--autotune", action="store_true",
                    help="run the library's autotuner for each form before timing it (its candidate "
                         "launches then share kernel names with the timed ones in a profile)")
    ap.add_argument("--rhs", type=int, default=0,
                    help="gamma right-hand sides per application (default 4 for config4, else 1)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-mixes", type=int, default=128)
    ap.add_argument("--batched-split", action="store_true",
                    help="config3: run the batch in split BF16 (default: plain FP32)")
    ap.add_argument("--per-instance", action="store_true",
                    help="config3: also time every instance alone and report its kernel times")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates HBM byte counters) and exit")
    args = ap.parse_args()
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              f"(the operator pair runs on a single MI355X).", file=sys.stderr, flush=True)
        sys.exit(2)

    if args.config == "config3":
        bench_batched(args)
        return

    import torch
    from comp_ui_toolbox_amd import OperatorBorn
    from comp_ui_toolbox_amd import configs as C

    spec = C.CONFIGS[args.config]()
    t0 = time.perf_counter()
    op = OperatorBorn.from_sequence(spec, precision="fp32")
    torch.cuda.synchronize()
    t_tables = time.perf_counter() - t0
    print(f"[bench] tables built on device in {t_tables:.2f} s, path {op.path}", file=sys.stderr, flush=True)

    n_rhs = args.rhs if args.rhs > 0 else (4 if args.config == "config4" else 1)
    x0 = spec_gamma(spec)
    rng = np.random.Generator(np.random.PCG64(spec.seed + 100))
    # column 0: the phantom; further columns: independent N(0,1) speckle
    xs = [x0] + [rng.standard_normal(x0.shape) for _ in range(n_rhs - 1)]
    x = torch.tensor(np.stack(xs).astype(np.complex64), device="cuda")
    u = torch.empty((n_rhs, op.N_observations), dtype=torch.complex64, device="cuda")
    g = torch.empty((n_rhs, op.N_points), dtype=torch.complex64, device="cuda")

    forms = [args.precision]
    if not args.no_compare and args.precision != "fp32":
        forms.append("fp32")
    res, outs = {}, {}
    tuned = {}
    for form in forms:
        if args.autotune:
            op.set_precision(form)
            tuned[form] = op.autotune()
            print(f"[bench] {form}: autotuned contraction fwd {tuned[form][0]:.3f} ms adj {tuned[form][1]:.3f} ms",
                  file=sys.stderr, flush=True)
        res[form] = time_form(op, form, x, u, g, args.steps, args.warmup)
        # one pair from the same input, kept for the cross-form comparison
        uf = torch.empty_like(u)
        gf = torch.empty_like(g)
        op.forward_dev(x, out=uf)
        op.adjoint_dev(uf, out=gf)
        outs[form] = (uf, gf)
    torch.cuda.synchronize()

    F = algorithmic_flops(spec, n_rhs)
    B = algorithmic_bytes(spec, n_rhs)
    elapsed, kern = res[args.precision]
    pairs_per_s = args.steps / elapsed
    result = {
        "metric": METRIC,
        "value": pairs_per_s,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else args.precision,
        "data": f"synthetic ({args.config} speckle phantom; tables built on the device from the {args.config} geometry)",
        "config": {
            "workload": f"{args.config}: {spec.array.N_elements} el, {spec.fov.N_points_axis[0]}x"
                        f"{spec.fov.N_points_axis[-1]} FOV, {len(spec.frequencies)} f, "
                        f"{spec.N_measurements} waves, {spec.N_observations()} obs, {n_rhs} rhs",
            "precision": args.precision,
            "path": op.path,
            "autotuned": bool(args.autotune),
            "table_build_s": round(t_tables, 3),
            "kernel_ms": {k: round(v, 4) for k, v in kern.items()},
            "flops_per_pass": F,
            "bytes_per_pass": B,
        },
        "roofline": roofline(F, kern, args.precision, args.config),
    }
    if "fp32" in res and args.precision != "fp32":
        e32, k32 = res["fp32"]
        result["config"]["fp32_pairs_per_s"] = args.steps / e32
        result["config"]["fp32_kernel_ms"] = {k: round(v, 4) for k, v in k32.items()}
        result["config"]["ratio_vs_fp32"] = e32 / elapsed
        result["config"]["rel_l2_vs_fp32"] = {
            "forward": rel_l2(outs[args.precision][0], outs["fp32"][0]),
            "adjoint": rel_l2(outs[args.precision][1], outs["fp32"][1]),
        }
        result["roofline_fp32"] = roofline(F, k32, "fp32", args.config)
    if n_rhs == 1:
        try:
            result["config"]["host_entry"] = host_entry_rate(op, args.precision, x0)
        except Exception as e:  # report, do not hide
            result["config"]["host_entry"] = {"error": repr(e)}
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        if spec.fov.N_dimensions == 2 and n_rhs == 1:
            try:
                result["err_vs_fp64"] = err_vs_fp64(op, spec, forms)
            except Exception as e:  # report, do not hide
                result["err_vs_fp64"] = {"error": repr(e)}
        try:
            result["cpu_baseline"] = cpu_baseline(op, spec, args.cpu_mixes, threads)
        except Exception as e:  # report, do not hide
            result["cpu_baseline"] = {"value": None, "error": repr(e)}
        if args.config == "config1":
            op.close()
            try:
                result["config0_full"] = config0_full(threads, args.steps)
            except Exception as e:
                result["config0_full"] = {"error": repr(e)}
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
