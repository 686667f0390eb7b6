"""Benchmark of the D3Q19 TRT step on one MI355X.

Workload (one "step" = one lattice update of every fluid cell):
  512^3 channel (walls at y = 0 and y = 511, periodic x and z) with 2000
  seeded spheres (radius U[4,12]), aa-soa, omega 1.25 (tau 0.8),
  Lambda 3/16, body force g = 1e-6, start f_i = w_i (1 + eps),
  eps ~ U(-1e-3, 1e-3) from seed 0.  Inputs are built on the device; the
  timed region holds only step launches (hipEvents on the library's stream).

Printed: one JSON line.  value = FP64 MLUPS (fluid cells x steps / s);
roofline = 304 B per fluid-cell update over the measured step time against
8 TB/s; fp16 = the same workload with FP16 shifted-deviation storage
(76 B per update) and its error against the FP64 run; cpu_baseline = the
reference's own aa-soa (oracle/_ref/lbmbench-ref, serial as in the
reference) on a 256^3 channel on this host.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lbm-bench_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "MLUPS + %HBM roofline, D3Q19 TRT 512\u00b3 FP64 vs FP16-store; adjoint MLUPS"
BYTES_PER_LUP = {"native": 304, "fp32": 152, "fp16": 76, "bf16": 76}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def traffic_from_profiles(kernel_tag):
    """HBM bytes per launch of the dominant kernel, from rocprofv3 --pmc
    FETCH_SIZE / WRITE_SIZE runs summarised in profiles/traffic.json (not
    measured by this run: a profiler cannot run inside it) -> (bytes, source)"""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        data = json.load(f)
    entry = data.get(kernel_tag)
    if entry is None:
        return None, None
    return entry["bytes"], entry["source"]


# the copy shapes of the achievable-bandwidth sweep: 16-byte vectors per lane,
# non-temporal or cached, one workgroup per tile or a persistent grid
COPY_SHAPES = [(v, nt, pg) for v in (1, 2, 4, 8) for nt in (True, False) for pg in (0, 2048)]


def copy_sweep(lib, nbytes=4 << 30, reps=8):
    """GB/s of every copy shape; the best one is the achievable line"""
    rows = []
    for v, nt, pg in COPY_SHAPES:
        gbs = lib.copy_bandwidth_shape(nbytes, reps, v, nt, pg)
        rows.append({"vectors_per_lane": v, "non_temporal": nt, "persistent_groups": pg,
                     "gbs": round(gbs, 1)})
    return rows


def cpu_baseline():
    """LBM-Bench's own aa-soa (serial in the reference) on a 256^3 channel:
    44 steps, about 12 s of stepping at the ~60 MLUPS one host core gives"""
    exe = os.path.join(ROOT, "oracle", "_ref", "lbmbench-ref")
    if not os.path.exists(exe):
        return None
    iters = 44
    cmd = [exe, "-d", "256x256x256", "-g", "channel", "-z", "-k", "aa-soa", "-i", str(iters),
           "-o", "1.25", "-f", "1e-6"]
    env = dict(os.environ, OMP_NUM_THREADS="1")
    log("cpu baseline: " + " ".join(cmd[1:]))
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600).stdout
    m = re.search(r"^P:\s+([0-9.]+) MFLUP/s\s+d: ([0-9.]+) s", out, flags=re.M)
    if not m:
        return None
    return {"value": float(m.group(1)), "unit": "MLUPS", "cores": 1, "kind": "reference",
            "sample": f"LBM-Bench aa-soa (serial in the reference), 256^3 channel -z, {iters} "
                      f"steps, {float(m.group(2)):.1f} s; the reference cannot index 512^3"}


def adjoint_run(lib, lbmhip):
    """differentiation form: 128^3 porous channel, T = 1024, sqrt(T)
    checkpoints; FP64 and FP32 (FP32 forward and adjoint)"""
    dims, T = (128, 128, 128), 1024
    s = lbmhip.porous_field(lib, dims, 0, 3.0)
    out = {"workload": "Brinkman TRT 128^3 porous channel (s: U[0,1] noise, Gaussian sigma 3, "
                       "logistic 4z-2), body force 1e-5, rest start, T 1024, J = outlet-plane "
                       "kinetic energy, sqrt(T) = 32-step checkpoints"}
    for form in ("native", "fp32"):
        A = lbmhip.Adjoint(lib, "channel", dims, (0, 0, 1), omega=1.25, magic=3.0 / 16.0,
                           xforce=1e-5, storage=form)
        nf = int(A.fluid_mask().sum())
        A.set_field(s)
        A.init("rest", 0, 1.0, 0.0)
        A.step_times(3)
        st = A.step_times(20)
        J, g, stats = A.gradient(T, [32])
        total = stats["forward_s"] + stats["recompute_s"] + stats["adjoint_s"]
        key = "fp64" if form == "native" else "fp32"
        # whole gradient: forward sweep, recomputation and adjoint sweep
        out[key] = {
            "J": J,
            "forward_step_ms": round(st["forward_s"] * 1e3, 4),
            "adjoint_step_ms": round(st["adjoint_s"] * 1e3, 4),
            "adjoint_to_forward_step": round(st["adjoint_s"] / st["forward_s"], 3),
            "mlups_forward": round(nf * T / stats["forward_s"] / 1e6, 1),
            "mlups_adjoint_sweep": round(nf * T / stats["adjoint_s"] / 1e6, 1),
            "mlups_gradient": round(nf * T / total / 1e6, 1),
            "gradient_to_forward_time": round(total / stats["forward_s"], 3),
            "device_gib": round(stats["device_bytes"] / 2 ** 30, 2),
        }
        if key == "fp64":
            g64 = g
        else:
            out[key]["gradient_l2_rel_err_vs_fp64"] = float(np.linalg.norm(g - g64) / np.linalg.norm(g64))
        log(f"adjoint {key}: step ratio {out[key]['adjoint_to_forward_step']}, gradient "
            f"{out[key]['mlups_gradient']} MLUPS, {out[key]['device_gib']} GiB")
        A.close()
    return out


def dominant_name(kernel, kind):
    """the HIP kernel behind a step kind of the bench's default settings"""
    if kernel == "aa-soa":
        return "kRowGroup<AA odd, FP64>" if kind == "odd" else "kStepRow<AA even, FP64>"
    return f"{kernel} step (FP64)"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_reference_openmp():
    """LBM-Bench's own pull-soa, which the reference parallelises with
    OpenMP (collapse(2) over x, y), on up to 16 host threads, 256^3 channel"""
    exe = os.path.join(ROOT, "oracle", "_ref", "lbmbench-ref")
    if not os.path.exists(exe):
        return None
    threads, iters = min(16, os.cpu_count() or 1), 60
    cmd = [exe, "-d", "256x256x256", "-g", "channel", "-z", "-k", "pull-soa", "-i", str(iters),
           "-o", "1.25", "-f", "1e-6"]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
               OMP_PROC_BIND="close")
    log(f"cpu reference, {threads} threads: " + " ".join(cmd[1:]))
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=600).stdout
    m = re.search(r"^P:\s+([0-9.]+) MFLUP/s\s+d: ([0-9.]+) s", out, flags=re.M)
    if not m:
        return None
    return {"value": float(m.group(1)), "unit": "MLUPS", "threads": threads, "kind": "reference",
            "kernel": "pull-soa", "nproc": os.cpu_count(), "cpu_model": cpu_model(),
            "sample": f"LBM-Bench pull-soa (OpenMP), 256^3 channel -z, {iters} steps, "
                      f"{float(m.group(2)):.1f} s"}


def ref_cpu_mlups(kernel, n, iters, threads):
    """LBM-Bench's own binary (oracle/_ref/lbmbench-ref) for one variant on
    an n^3 channel: its printed MFLUP/s, or None where it is not built"""
    exe = os.path.join(ROOT, "oracle", "_ref", "lbmbench-ref")
    if not os.path.exists(exe):
        return None
    cmd = [exe, "-d", f"{n}x{n}x{n}", "-g", "channel", "-z", "-k", kernel, "-i", str(iters),
           "-o", "1.25", "-f", "1e-6"]
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PLACES="cores",
               OMP_PROC_BIND="close")
    out = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=900).stdout
    m = re.search(r"^P:\s+([0-9.]+) MFLUP/s", out, flags=re.M)
    return float(m.group(1)) if m else None


def study_variants(lib, lbmhip, n, steps, cpu=True):
    """the seven variants on an n^3 channel (walls at y, periodic x and z),
    f = w(1+eps) from seed 0, FP64: MLUPS and GB/s each, beside LBM-Bench's
    own CPU kernels on the host (OpenMP where the reference has it: push-soa
    and pull-soa; the other five are serial there)"""
    dims = (n, n, n)
    kw = dict(omega=1.25, xforce=1e-6, body_force=True, magic=3.0 / 16.0)
    rows = {}
    for kernel in lbmhip.KERNELS:
        L = lbmhip.Lattice(lib, kernel, "channel", dims, (0, 0, 1), **kw)
        L.init("random", 0, 1.0, 1e-3)
        L.run(20)
        dur = L.run(steps)
        ml = L.n_fluid * steps / dur / 1e6
        gbs = BYTES_PER_LUP["native"] * ml / 1e3
        L.close()
        threads = min(16, os.cpu_count() or 1) if kernel in ("push-soa", "pull-soa") else 1
        host = ref_cpu_mlups(kernel, n, 40 if threads > 1 else 8, threads) if cpu else None
        rows[kernel] = {"mlups": round(ml, 1), "gb_s": round(gbs, 1),
                        "frac_of_8tb": round(gbs / HBM_PEAK_GBS, 4),
                        "cpu_reference_mlups": host, "cpu_threads": threads}
        log(f"{kernel}: {ml:.0f} MLUPS, {gbs:.0f} GB/s; LBM-Bench on {threads} host "
            f"thread(s): {host if cpu else 'skipped'} MLUPS")
    return {"study": "the seven variants", "workload": f"{n}^3 channel -z, omega 1.25, Lambda "
            f"3/16, body force 1e-6, f = w(1+eps), {steps} steps, FP64", "cpu_model": cpu_model(),
            "variants": rows}


def study_ceiling(lib, n):
    """the step's loads and stores without arithmetic (lbmPatternBandwidth) on
    n^3 cells: GB/s by plane count, in place or to a second array, and by
    element size; the copy kernel beside them"""
    cells = n ** 3
    rows = []
    for planes in (1, 2, 4, 8, 19):
        for in_place in (True, False):
            rows.append((planes, 8, in_place))
    rows += [(19, 4, True), (19, 2, True)]
    out = []
    for planes, eb, in_place in rows:
        gbs = lib.pattern_bandwidth(planes, eb, in_place, cells, 6)
        ms = 2.0 * planes * cells * eb / (gbs * 1e9) * 1e3
        out.append({"planes": planes, "element_bytes": eb, "in_place": in_place,
                    "gb_s": round(gbs, 1), "frac_of_8tb": round(gbs / HBM_PEAK_GBS, 4),
                    "ms_per_launch": round(ms, 3)})
        log(f"{planes:2d} planes x {eb} B, {'in place' if in_place else 'two arrays'}: "
            f"{gbs:.0f} GB/s, {ms:.3f} ms")
    copy_gbs = lib.copy_bandwidth(4 << 30, 10)
    log(f"copy (4 GiB): {copy_gbs:.0f} GB/s")
    return {"study": "access-pattern ceiling", "workload": f"{n}^3 cells per plane, non-temporal "
            "loads and stores, one element per lane, 512-lane blocks, 6 launches each",
            "patterns": out, "copy_bandwidth_gbs": round(copy_gbs, 1)}


def study_precision(lib, lbmhip, n, spheres, total, chunk):
    """The 512^3 spheres workload in FP64, FP32, FP16 and BF16 storage, run
    side by side for `total` steps; after every `chunk` steps the velocity
    error and mass drift of each form against the FP64 run"""
    dims = (n, n, n)
    kw = dict(omega=1.25, xforce=1e-6, body_force=True, magic=3.0 / 16.0, seed=0)
    geo = f"spheres-{spheres}"
    lats = {form: lbmhip.Lattice(lib, "aa-soa", geo, dims, (0, 0, 1), storage=form, **kw)
            for form in ("native", "fp32", "fp16", "bf16")}
    for L in lats.values():
        L.init("random", 0, 1.0, 1e-3)
    m0 = {form: L.mass_momentum() for form, L in lats.items()}
    nfluid = lats["native"].n_fluid
    secs = {form: 0.0 for form in lats}
    rows = []
    for done in range(chunk, total + 1, chunk):
        for form, L in lats.items():
            secs[form] += L.run(chunk)
        row = {"steps": done}
        for form, L in lats.items():
            m = L.mass_momentum()
            entry = {"mass_drift_rel": (m[0] - m0[form][0]) / m0[form][0]}
            if form != "native":
                cmp = L.compare_momentum(lats["native"])
                entry["velocity_l2_rel_err_vs_fp64"] = (cmp[0] / cmp[1]) ** 0.5
                entry["velocity_max_abs_err_vs_fp64"] = cmp[2]
            row["fp64" if form == "native" else form] = entry
        rows.append(row)
        log(f"precision study: {done} steps, u L2 err fp32 "
            f"{row['fp32']['velocity_l2_rel_err_vs_fp64']:.3e} fp16 "
            f"{row['fp16']['velocity_l2_rel_err_vs_fp64']:.3e} bf16 "
            f"{row['bf16']['velocity_l2_rel_err_vs_fp64']:.3e}")
    for L in lats.values():
        L.close()
    return {"study": "precision forms against FP64", "workload": f"aa-soa {n}^3 channel with "
            f"{spheres} seeded spheres, omega 1.25, Lambda 3/16, body force 1e-6, "
            f"f = w(1+eps)", "fluid_cells": nfluid,
            "mlups": {("fp64" if f == "native" else f): round(nfluid * total / secs[f] / 1e6, 1)
                      for f in secs},
            "series": rows}


def study_adjoint(lib, lbmhip, n, T, intervals, sigma):
    """The differentiation form at scale: an n^3 porous channel (s from
    U[0,1] noise smoothed with sigma and a logistic), FP16 forward storage
    with an FP32 adjoint, multi-level checkpoints; J also from an FP64
    forward run of the same case"""
    dims = (n, n, n)
    s = lbmhip.porous_field(lib, dims, 0, sigma)
    kw = dict(omega=1.25, magic=3.0 / 16.0, xforce=1e-5)
    A = lbmhip.Adjoint(lib, "channel", dims, (0, 0, 1), storage="fp16", **kw)
    nf = int(A.fluid_mask().sum())
    A.set_field(s)
    A.init("rest", 0, 1.0, 0.0)
    st = A.step_times(10)
    log(f"adjoint study: {n}^3, T {T}, checkpoints {intervals}: forward step "
        f"{st['forward_s'] * 1e3:.3f} ms, adjoint step {st['adjoint_s'] * 1e3:.3f} ms")
    J, g, stats = A.gradient(T, intervals)
    A.close()
    total = stats["forward_s"] + stats["recompute_s"] + stats["adjoint_s"]
    log(f"adjoint study: gradient in {total:.1f} s, peak {stats['device_bytes'] / 2**30:.1f} GiB")
    B = lbmhip.Adjoint(lib, "channel", dims, (0, 0, 1), storage="native", **kw)
    B.set_field(s)
    B.init("rest", 0, 1.0, 0.0)
    J64, t64 = B.forward(T)
    B.close()
    return {"study": "adjoint at scale", "workload": f"Brinkman TRT {n}^3 porous channel "
            f"(s: U[0,1] noise, Gaussian sigma {sigma}, logistic 4z-2), body force 1e-5, rest "
            f"start, T {T}, J = outlet-plane kinetic energy", "fluid_cells": nf,
            "storage": "FP16 forward, FP32 adjoint", "checkpoint_intervals": intervals,
            "J": J, "J_fp64_forward": J64, "J_rel_diff": abs(J - J64) / abs(J64),
            "forward_step_ms": round(st["forward_s"] * 1e3, 4),
            "adjoint_step_ms": round(st["adjoint_s"] * 1e3, 4),
            "adjoint_to_forward_step": round(st["adjoint_s"] / st["forward_s"], 3),
            "forward_s": round(stats["forward_s"], 3), "recompute_s": round(stats["recompute_s"], 3),
            "adjoint_s": round(stats["adjoint_s"], 3),
            "mlups_gradient": round(nf * T / total / 1e6, 1),
            "gradient_to_forward_time": round(total / stats["forward_s"], 3),
            "peak_device_gib": round(stats["device_bytes"] / 2 ** 30, 2),
            "gradient_l2_norm": float(np.linalg.norm(g))}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--spheres", type=int, default=2000)
    ap.add_argument("--kernel", default="aa-soa")
    ap.add_argument("--geometry", default=None,
                    help="geometry name instead of spheres-<--spheres> (e.g. channel)")
    ap.add_argument("--no-sweep", action="store_true",
                    help="skip the copy-shape sweep and the 19-plane pattern (profiling runs)")
    ap.add_argument("--no-fp16", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--study-steps", type=int, default=None,
                    help="steps per variant of --study variants256 (default 500; fewer for "
                         "counter runs)")
    ap.add_argument("--study", choices=["precision5000", "adjoint384", "variants256", "ceiling"],
                    help="instead of the bench line, one JSON line of a longer study: the four "
                         "storage forms over 5000 steps, the 384^3 FP16/FP32 adjoint with "
                         "T = 4096 and two checkpoint levels, the seven variants at 256^3, or "
                         "the step's memory traffic without arithmetic at --size^3")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the halo-exchange path across GPUs is not built)", file=sys.stderr)
        sys.exit(2)

    import lbmhip
    lib = lbmhip.Library("dp")
    dev, cus, mem = lib.device_info()
    log(f"device {dev}, {cus} CUs, {mem / 2**30:.0f} GiB")

    if args.study == "precision5000":
        print(json.dumps(study_precision(lib, lbmhip, args.size, args.spheres, 5000, 1000)),
              flush=True)
        return
    if args.study == "variants256":
        print(json.dumps(study_variants(lib, lbmhip, 256, args.study_steps or 500,
                                        cpu=not args.no_cpu)), flush=True)
        return
    if args.study == "ceiling":
        print(json.dumps(study_ceiling(lib, args.size)), flush=True)
        return
    if args.study == "adjoint384":
        print(json.dumps(study_adjoint(lib, lbmhip, 384, 4096, [64, 8], 5.0)), flush=True)
        return

    n = args.size
    dims = (n, n, n)
    geo = args.geometry or f"spheres-{args.spheres}"
    kw = dict(omega=1.25, xforce=1e-6, body_force=True, magic=3.0 / 16.0, seed=0)

    copy_gbs = lib.copy_bandwidth(4 << 30, 10)
    log(f"device copy bandwidth {copy_gbs:.0f} GB/s")
    sweep, best_copy, pattern_gbs = None, None, None
    if not args.no_sweep:
        sweep = copy_sweep(lib)
        best_copy = max(sweep, key=lambda r: r["gbs"])
        log(f"best copy shape {best_copy}")
        # the AA step's loads and stores without arithmetic, on this lattice's size
        pattern_gbs = lib.pattern_bandwidth(19, 8, True, n ** 3, 6)
        log(f"19-plane in-place pattern bandwidth {pattern_gbs:.0f} GB/s")

    result = {}
    L64 = lbmhip.Lattice(lib, args.kernel, geo, dims, (0, 0, 1), **kw)
    placement = L64.placement()
    log(f"placement probe: {placement[0]} candidate allocation(s), kept one at "
        f"{placement[1]:.0f} GB/s")
    nfluid = L64.n_fluid
    solid_fraction = 1.0 - nfluid / float(n ** 3)
    log(f"lattice {dims}, {nfluid} fluid cells, solid fraction {solid_fraction:.4f}, "
        f"{L64.device_bytes() / 2**30:.1f} GiB of PDFs")
    L64.init("random", 0, 1.0, 1e-3)
    m0 = L64.mass_momentum()
    L64.run(args.warmup)
    lib.synchronize()
    t0 = time.perf_counter()
    dur = L64.run(args.steps)
    lib.synchronize()
    wall = time.perf_counter() - t0
    m1 = L64.mass_momentum()
    mlups = nfluid * args.steps / dur / 1e6

    step_s = dur / args.steps
    achieved = BYTES_PER_LUP["native"] * nfluid / step_s / 1e9
    log(f"fp64 {args.kernel}: {mlups:.0f} MLUPS, {step_s * 1e3:.3f} ms/step, "
        f"{achieved:.0f} GB/s ({achieved / HBM_PEAK_GBS * 100:.1f} % of 8 TB/s), wall {wall:.2f} s")

    forms = {}
    for form in ([] if args.no_fp16 else ["fp16", "fp32", "bf16"]):
        Lf = lbmhip.Lattice(lib, args.kernel, geo, dims, (0, 0, 1), storage=form, **kw)
        Lf.init("random", 0, 1.0, 1e-3)
        Lf.run(args.warmup)
        lib.synchronize()
        df = Lf.run(args.steps)
        mf = Lf.mass_momentum()
        cmp = Lf.compare_momentum(L64)
        mlf = nfluid * args.steps / df / 1e6
        achf = BYTES_PER_LUP[form] * nfluid / (df / args.steps) / 1e9
        forms[form] = {
            "mlups": round(mlf, 1), "speedup_vs_fp64": round(mlf / mlups, 3),
            "ms_per_step": round(df / args.steps * 1e3, 4),
            "roofline": {"bound": "hbm", "achieved": round(achf, 1), "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": round(achf / HBM_PEAK_GBS, 4),
                         "bytes_per_lup": BYTES_PER_LUP[form]},
            "velocity_l2_rel_err_vs_fp64": (cmp[0] / cmp[1]) ** 0.5 if cmp[1] > 0 else None,
            "velocity_max_abs_err_vs_fp64": cmp[2],
            "mass_drift_rel": (mf[0] - m0[0]) / m0[0],
            "steps_compared": args.warmup + args.steps}
        log(f"{form} {args.kernel}: {mlf:.0f} MLUPS ({mlf / mlups:.2f}x fp64), "
            f"u L2 err {forms[form]['velocity_l2_rel_err_vs_fp64']:.3e}")
        Lf.close()
    # per-kernel durations: the same steps launched one by one with a hipEvent
    # after each launch on the library's stream (aa-soa: even and odd steps)
    _, kt = L64.run_timed(args.steps)
    kernels = {}
    for kind, (secs, count) in kt.items():
        if count:
            ms_k = secs / count * 1e3
            gbs_k = BYTES_PER_LUP["native"] * nfluid / (secs / count) / 1e9
            kernels[kind] = {"launches": count, "avg_ms": round(ms_k, 6),
                             "achieved_gbs": round(gbs_k, 1),
                             "frac": round(gbs_k / HBM_PEAK_GBS, 4)}
    dominant = max(kernels, key=lambda k: kernels[k]["avg_ms"])
    dom = kernels[dominant]
    log(f"kernels {kernels}; dominant {dominant}")
    L64.close()

    adjoint = None if args.no_adjoint else adjoint_run(lib, lbmhip)
    traffic = traffic_from_profiles(f"{args.kernel}-fp64-{n}-{geo}-{dominant}")

    cpu = None if args.no_cpu else cpu_baseline()
    cpu_omp = None if args.no_cpu else cpu_reference_openmp()

    out = {
        "metric": METRIC,
        "value": round(mlups, 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(step_s * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded sphere packing and seeded f_i = w_i(1+eps), built on the device",
        "config": {"workload": f"D3Q19 TRT {args.kernel}, {n}^3 " + (
                               f"'{args.geometry}' geometry" if args.geometry else
                               f"channel (walls at y) with {args.spheres} seeded spheres r~U[4,12]")
                               + ", omega 1.25, Lambda 3/16, body force 1e-6",
                   "kernel": args.kernel, "lattice": list(dims),
                   "fluid_cells": nfluid, "solid_fraction": round(solid_fraction, 5),
                   "storage": "fp64"},
        "roofline": {"bound": "hbm", "achieved": dom["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": dom["frac"],
                     "kernel": dominant_name(args.kernel, dominant),
                     "kernel_avg_ms": dom["avg_ms"],
                     "bytes_per_launch": BYTES_PER_LUP["native"] * nfluid,
                     "traffic": traffic[0], "traffic_source": traffic[1],
                     "step_pair_achieved": round(achieved, 1),
                     "step_pair_frac": round(achieved / HBM_PEAK_GBS, 4),
                     "achievable_gbs": best_copy["gbs"] if best_copy else None,
                     "achievable_shape": {k: best_copy[k] for k in ("vectors_per_lane",
                                                                    "non_temporal",
                                                                    "persistent_groups")}
                     if best_copy else None,
                     "frac_of_achievable": round(dom["achieved_gbs"] / best_copy["gbs"], 4)
                     if best_copy else None},
        "kernels": kernels,
        "copy_sweep_gbs": sweep,
        "cpu_baseline": cpu,
        "cpu_reference_openmp": cpu_omp,
        "fp16": forms.get("fp16"),
        "fp32": forms.get("fp32"),
        "bf16": forms.get("bf16"),
        "adjoint_mlups": adjoint["fp64"]["mlups_gradient"] if adjoint else None,
        "adjoint_sweep_mlups": adjoint["fp64"]["mlups_adjoint_sweep"] if adjoint else None,
        "adjoint": adjoint,
        "mass_drift_rel": (m1[0] - m0[0]) / m0[0],
        "placement_probe": {"candidates": placement[0], "kept_gbs": round(placement[1], 1)},
        "copy_bandwidth_gbs": round(copy_gbs, 1),
        "pattern_bandwidth_gbs": round(pattern_gbs, 1) if pattern_gbs else None,
        "device": dev,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
