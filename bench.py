"""RK3 steps/s of the shear-periodic box on one MI355X.

Workload: 384x385x384 (mgalx x my x mgalz), Lx:Ly:Lz = 2:1:1, S = 1,
Re_z = 8000, CFL 0.5 adaptive dt, explicit viscous terms, FP64, seeded
k^4 exp(-2 (k/k0)^2) solenoidal field (synthetic).  A step is one full RK3
step (3 substeps) of the whole box with the state resident in HBM.

Prints one JSON line (see the "roofline" and "cpu_baseline" members for the
dominant kernel's HBM roofline and the CPU oracle timed on a bounded sample).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--grid", type=str, default="384,385,384")
    ap.add_argument("--Re", type=float, default=8000.0)
    ap.add_argument("--explicit", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed precision leg")
    ap.add_argument("--batch", type=int, default=128,
                    help="boxes of the ensemble leg (0: skip it)")
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-only", action="store_true",
                    help="only the ensemble's box-table legs (FP64 and mixed), for a kernel "
                         "trace of the ensemble; prints the ensemble object alone")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="precision of the main timed leg (mixed: no separate mixed leg)")
    ap.add_argument("--grid2", type=str, default="512,513,512",
                    help="second grid of the mixed-precision configuration ('' to skip)")
    ap.add_argument("--grid2-steps", type=int, default=20,
                    help="timed steps on the second grid (FP64 and mixed)")
    ap.add_argument("--cmp-steps", type=int, default=2000,
                    help="steps of the FP64-vs-mixed statistics comparison (2000: the "
                         "length the mixed-precision workload is specified over)")
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="oracle processes of the CPU baseline (capped at the host's CPUs)")
    ap.add_argument("--cpu-steps", type=int, default=1,
                    help="RK3 steps of each CPU baseline process")
    ap.add_argument("--cpu-grid", type=str, default="",
                    help="grid of the CPU baseline sample (default: the workload grid)")
    ap.add_argument("--copy-probe", type=float, default=0.0,
                    help="only run the streaming-copy probe on this many GiB and exit")
    return ap.parse_args()


def algorithmic_bytes(kxn, mz, my, mgalz, precision="fp64"):
    """compulsory HBM bytes per launch of each kernel class.  Nc = kxn*mz*my
    spectral coefficients, Nh = kxn*mgalz*my hybrid values.  FP64: every
    field 16 B a value.  Mixed: hybrid planes and the six inter-stage fields
    (v, dv/dy, d(omega_y)/dy; rdome, rhg, rhv) 8 B, the state and the RK
    accumulators 16 B."""
    nc = kxn * mz * my
    nh = kxn * mgalz * my
    hb = 16.0 if precision == "fp64" else 8.0   # hybrid planes
    ib = 16.0 if precision == "fp64" else 8.0   # inter-stage fields
    return {
        # A: reads vor, phi, v, dv/dy, d(omega_y)/dy; writes 6 hybrid fields
        "xz_A": 16.0 * 2 * nc + ib * 3 * nc + hb * 6 * nh,
        # B: reads 6 hybrid fields, writes 3
        "xz_B": hb * 9 * nh,
        # C: reads 3 hybrid fields and v, writes 3 spectral fields
        "xz_C": hb * 3 * nh + ib * 4 * nc,
        # fused y stage, averaged over the 3 substeps: reads rdome, rhv, rhg,
        # vor, phi (+ vorwk, phiwk at rk 2,3); writes vor, phi, vorwk,
        # phiwk (not at rk 3), v, dv/dy, d(omega_y)/dy
        "y_lines": nc * (6 * ib + 4 * 16.0 + 4 * 16.0 * 2.0 / 3.0),
    }


def step_bytes(nc, precision="fp64"):
    """compulsory bytes of one RK3 step with the physical planes on chip:
    B64 = 992 Nc (FP64), B_mix = 704 Nc (mixed: FP32 inter-stage fields)"""
    return (992.0 if precision == "fp64" else 704.0) * nc


def measured_traffic(cls, grid, explicit, precision="fp64"):
    """HBM bytes per launch of a kernel class, from the committed rocprofv3
    counter summary of the same workload (FETCH_SIZE and WRITE_SIZE per
    dispatch, scaled by the copy-kernel calibration); None if there is none."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "rk3_step_%dx%dx%d_%s.traffic.json" % (grid + (precision,)))
    if not os.path.exists(path):
        return None, None
    with open(path) as f:
        d = json.load(f)
    if d.get("explicit") != int(explicit) or cls not in d.get("classes", {}):
        return None, None
    return d["classes"][cls]["bytes_per_launch"], os.path.relpath(path, os.path.dirname(path) + "/..")


def _cpu_sample(args):
    """one process: nstep RK3 steps of the numpy oracle at the sample grid;
    returns (seconds per step, start and end wall-clock times)"""
    grid, sample_grid, Re, explicit, seed, nstep = args
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.grid import Grid
    from oracle.cross import Cross
    from shear_nnf_amd.ic import make_ic
    mgalx, my, mgalz = sample_grid
    g = Grid(mgalx, my, mgalz, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0, s=1.0, CFL=0.5)
    vor, phi, u00, w00 = make_ic(mgalx, my, mgalz, g.alp, g.gam, g.Ly, seed=seed)
    c = Cross(g, explicit=explicit, nhist=10 ** 9)
    c.set_state(vor, phi, u00, w00)
    t0 = time.perf_counter()
    for _ in range(nstep):
        c.run(1)
        print(f"cpu baseline: one {mgalx}x{my}x{mgalz} step of the oracle done", file=sys.stderr,
              flush=True)
    t1 = time.perf_counter()
    return (t1 - t0) / nstep, t0, t1


def host_cpu():
    """CPU model and logical core count of this host"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def cpu_baseline(grid, sample_grid, Re, explicit, seed, procs, nstep=2):
    """The numpy oracle on the host cores: procs processes at once, each
    running nstep RK3 steps of the same seeded state (one numpy thread each;
    the oracle has no threading of its own), so every core works on the
    reference algorithm.  By default the sample is the workload itself, one
    step per process (about 37 s a step at 384x385x384 with 16 processes,
    10 GB each); a smaller sample grid is scaled to the workload grid by
    the number of spectral coefficients (the oracle's cost is linear in Nc
    up to the log factor of the FFTs).  Runs before anything initialises
    the GPU (the workers are fresh processes)."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    with ctx.Pool(procs) as pool:
        res = pool.map(_cpu_sample, [(grid, sample_grid, Re, explicit, seed, nstep)] * procs)
    per_step = [r[0] for r in res]
    span = max(r[2] for r in res) - min(r[1] for r in res)
    rate = procs * nstep / span
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from oracle.grid import Grid
    G = Grid(grid[0], grid[1], grid[2], Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0)
    g = Grid(*sample_grid, Re=Re, alp=math.pi, gam=2 * math.pi, Ly=1.0)
    scale = g.nc / G.nc
    model, ncpu = host_cpu()
    return rate * scale, (f"{procs} processes at once, each {nstep} RK3 step(s) of the numpy oracle "
                          f"at {sample_grid[0]}x{sample_grid[1]}x{sample_grid[2]} "
                          f"({min(per_step):.2f}-{max(per_step):.2f} s per step per process, "
                          f"{span:.1f} s for all); aggregate {rate:.3f} steps/s scaled to "
                          f"{grid[0]}x{grid[1]}x{grid[2]} by the Nc ratio {scale:.4f}; host: "
                          f"{model}, {ncpu} logical CPUs")


def uner_stats(hist):
    """energy 0.5 (u'^2 + v'^2 + w'^2), Reynolds stress <uv>, dissipation of
    every UNER record (time, Deltat, uner(10), ...; cross.f90:2067-2119)"""
    import numpy as np
    h = np.asarray(hist)
    return {"energy": 0.5 * np.sum(h[:, 2:5] ** 2, axis=1), "reynolds_stress": h[:, 5],
            "dissipation": h[:, 11]}


def timed_leg(snf, ic, grid, Re, explicit, steps, warmup, precision):
    """warm-up, then exactly `steps` steps in one advance call (which
    synchronises the device before it returns), with per-kernel HIP events"""
    box = snf.ShearBox(*grid, Re, explicit=explicit, nhist=10 ** 9, precision=precision)
    box.set_state(*ic)
    box.advance_async_status(warmup)
    box.set_timing(True)
    t0 = time.perf_counter()
    st = box.advance_async_status(steps)
    t1 = time.perf_counter()
    timing = box.get_timing()
    box.set_timing(False)
    kxn, mz, nc = box.kxn, box.mz, box.nc
    box.close()
    return st, t1 - t0, timing, (kxn, mz, nc)


_FP64_RUNS = {}


def mixed_error(snf, ic, grid, Re, explicit, nstep):
    """FP64 and mixed runs of nstep steps from the same state with a UNER
    record every step: the largest relative difference of each statistic,
    and the relative L2 difference of the final fields (the FP64 run of a
    workload is kept for the next call)"""
    import numpy as np
    recs, fields = {}, {}
    for prec in ("fp64", "mixed"):
        key = (tuple(grid), Re, explicit, nstep, id(ic))
        if prec == "fp64" and key in _FP64_RUNS:
            recs[prec], fields[prec] = _FP64_RUNS[key]
            continue
        box = snf.ShearBox(*grid, Re, explicit=explicit, nhist=1, precision=prec)
        box.set_state(*ic)
        box.cross(nstep)
        recs[prec] = uner_stats(box.history)
        fields[prec] = box.get_state()
        box.close()
        if prec == "fp64":
            _FP64_RUNS[key] = (recs[prec], fields[prec])
    n = min(len(recs["fp64"]["energy"]), len(recs["mixed"]["energy"]))
    err = {k: float(np.max(np.abs(recs["mixed"][k][:n] - recs["fp64"][k][:n])
                           / np.abs(recs["fp64"][k][:n]))) for k in recs["fp64"]}
    names = ("vor", "phi", "u00", "w00")
    err["fields_rel_l2"] = {nm: float(np.linalg.norm(a - b) / np.linalg.norm(b))
                            for nm, a, b in zip(names, fields["mixed"], fields["fp64"])}
    err["steps"] = n
    return err


def grid2_leg(snf, make_ic, grid, Re, explicit, steps, warmup, cmp_steps, seed):
    """The mixed-precision configuration's second grid (512x513x512 by
    default): FP64 and mixed steps/s over `steps` timed steps each, the
    per-launch kernel times, and the mixed mode's statistics error against
    FP64 over cmp_steps steps from the same seeded state."""
    ic = make_ic(*grid, math.pi, 2 * math.pi, 1.0, seed=seed)
    res = {"grid": list(grid), "steps": steps}
    for prec in ("fp64", "mixed"):
        st, wall, timing, (_, _, nc) = timed_leg(snf, ic, grid, Re, explicit, steps, warmup, prec)
        sps = st.steps_done / wall
        res[prec] = {"value": sps, "unit": "steps/s",
                     "kernel_ms_per_launch": {k: ms / cnt for k, (ms, cnt) in timing.items() if cnt},
                     "step_roofline_frac": step_bytes(nc, prec) * sps / 1e9 / HBM_PEAK_GBS}
    res["mixed"]["speedup_vs_fp64"] = res["mixed"]["value"] / res["fp64"]["value"]
    res["mixed"]["rel_err_vs_fp64"] = (mixed_error(snf, ic, grid, Re, explicit, cmp_steps)
                                       if cmp_steps > 0 else None)
    return res


def batch_leg(snf, make_ic, nbox, steps, seed,
              modes=("sequential", "batched", "batched_mixed", "batched_streams")):
    """Ensemble of independent boxes (grids n^3 and n x (n+1) x n, n cycling
    over 32, 48, 64, 96; Re_z uniform in [500, 3000]; seeded fields): box-steps
    per second of snf_advance_batch (each size class advanced by one launch
    per kernel from box tables), against the same boxes advanced one after
    another and against one stream per box."""
    import numpy as np
    rng = np.random.default_rng(seed)
    ns = (32, 48, 64, 96)
    specs = []
    for q in range(nbox):
        n = ns[q % 4]
        my = n + 1 if (q // 4) % 2 == 0 else n
        specs.append(((n, my, n), float(rng.uniform(500.0, 3000.0)), seed + 1 + q))
    ics = [make_ic(*g, math.pi, 2 * math.pi, 1.0, seed=sd) for g, _, sd in specs]

    def boxes(precision):
        out = []
        for (g, Re, _), ic in zip(specs, ics):
            b = snf.ShearBox(*g, Re, nhist=10 ** 9, precision=precision)
            b.set_state(*ic)
            out.append(b)
        return out

    rates = {}
    for mode, prec in (("sequential", "fp64"), ("batched", "fp64"), ("batched_mixed", "mixed"),
                       ("batched_streams", "fp64")):
        if mode not in modes:
            continue
        # batched: one launch per kernel for each size class (box tables);
        # batched_streams: every box on its own stream, for comparison
        os.environ["SNF_BATCH_TABLES"] = "0" if mode == "batched_streams" else "1"
        bs = boxes(prec)
        snf.advance_batch(bs, 2)                        # warm-up
        t0 = time.perf_counter()
        if mode.startswith("batched"):
            st = snf.advance_batch(bs, steps)
        else:
            st = [b.advance_async_status(steps) for b in bs]
        t1 = time.perf_counter()
        rates[mode] = sum(s.steps_done for s in st) / (t1 - t0)
        for b in bs:
            b.close()
    if "sequential" not in rates:
        return {"boxes": nbox, "steps": steps,
                "grids": sorted({"%dx%dx%d" % g for g, _, _ in specs}), "unit": "box-steps/s",
                **rates}
    return {"boxes": nbox, "steps": steps, "grids": sorted({"%dx%dx%d" % g for g, _, _ in specs}),
            "unit": "box-steps/s", "dtype": "f64 (batched_mixed: the mixed precision mode)",
            "batched": rates["batched"], "sequential": rates["sequential"],
            "speedup": rates["batched"] / rates["sequential"],
            "batched_streams": rates["batched_streams"],
            "tables_vs_streams": rates["batched"] / rates["batched_streams"],
            "batched_mixed": rates["batched_mixed"],
            "mixed_vs_fp64_batched": rates["batched_mixed"] / rates["batched"]}


def main():
    a = parse()
    # GPU_MAX_HW_QUEUES stays at HIP's default (4 hardware queues): the
    # box-table ensemble, one stream per size class, runs 21% slower with 16
    # (8760 against 6946 box-steps/s), which only the one-stream-per-box
    # comparison leg would gain from.
    if a.gpus != 1:
        print("bench.py: one GPU only; the sharded multi-GPU path is designed, not built "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    mgalx, my, mgalz = (int(x) for x in a.grid.split(","))
    seed = 20260001
    cpu = None
    if not a.no_cpu and a.copy_probe <= 0 and not a.batch_only:
        # before anything touches the GPU: the workers are fresh processes
        cg = tuple(int(x) for x in a.cpu_grid.split(",")) if a.cpu_grid else (mgalx, my, mgalz)
        procs = max(1, min(a.cpu_procs, os.cpu_count() or 1))
        v, sample = cpu_baseline((mgalx, my, mgalz), cg, a.Re, a.explicit, seed, procs,
                                 a.cpu_steps)
        cpu = {"value": v, "unit": "steps/s", "cores": procs, "kind": "port",
               "sample": sample + " (numpy FFT and line solves)"}
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    import shear_nnf_amd as snf
    from shear_nnf_amd.ic import make_ic
    if a.copy_probe > 0:
        nbytes = int(a.copy_probe * 2 ** 30)
        ms = snf.copy_probe(nbytes, iters=20)
        print(json.dumps({"copy_probe_bytes": nbytes, "ms": ms,
                          "GBs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}), flush=True)
        return
    if a.batch_only:
        print(json.dumps({"ensemble": batch_leg(snf, make_ic, a.batch, a.batch_steps, seed,
                                                ("batched", "batched_mixed"))}), flush=True)
        return
    ic = make_ic(mgalx, my, mgalz, math.pi, 2 * math.pi, 1.0, seed=seed)
    grid = (mgalx, my, mgalz)
    # timed region: exactly K steps in one advance call, which synchronises the
    # device before it returns; per-kernel HIP events on the library's stream
    st2, wall, timing, (kxn, mz, nc) = timed_leg(snf, ic, grid, a.Re, a.explicit, a.steps,
                                                 a.warmup, a.precision)
    st = st2
    steps_done = st2.steps_done
    sps = steps_done / wall
    ab = algorithmic_bytes(kxn, mz, my, mgalz, a.precision)
    per_cls = {}
    for k, (ms, cnt) in timing.items():
        if cnt:
            per_cls[k] = ms / cnt
    dom = max((k for k in per_cls if k in ab), key=lambda k: timing[k][0])
    achieved = ab[dom] / (per_cls[dom] * 1e-3) / 1e9
    traffic, traffic_src = measured_traffic(dom, (mgalx, my, mgalz), a.explicit, a.precision)
    out = {
        "metric": "RK3 steps/s at 384\u00b3 (FP64 vs mixed/batched, rel err vs FP64); % HBM peak",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / max(steps_done, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.precision == "fp64" else "f32/f64 mixed",
        "data": "synthetic (seeded k^4 exp(-2(k/k0)^2) solenoidal field, k0=4)",
        "config": {"workload": f"shear-periodic box DNS {mgalx}x{my}x{mgalz}, Lx:Ly:Lz=2:1:1, "
                               f"S=1, Re_z={a.Re:g}, CFL 0.5 adaptive dt, "
                               f"{'explicit' if a.explicit else 'implicit'} viscous, "
                               f"{'FP64' if a.precision == 'fp64' else 'mixed precision'}",
                   "grid": [mgalx, my, mgalz], "Nc": nc},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": traffic_src,
                     "algorithmic_bytes_per_launch": ab[dom],
                     "avg_launch_ms": per_cls[dom]},
        "step_roofline": {"bytes_per_step": step_bytes(nc, a.precision),
                          "achieved_GBs": step_bytes(nc, a.precision) * sps / 1e9,
                          "frac": step_bytes(nc, a.precision) * sps / 1e9 / HBM_PEAK_GBS},
        "kernel_ms_per_launch": per_cls,
        "kernel_ms_per_step": {k: timing[k][0] / max(st.steps_done, 1) for k in timing},
        "sim_time_advanced": st2.time,
        "Deltat": st2.Deltat,
    }
    if not a.no_mixed and a.precision == "fp64":
        stm, wallm, timm, _ = timed_leg(snf, ic, grid, a.Re, a.explicit, a.steps, a.warmup, "mixed")
        spsm = stm.steps_done / wallm
        out["mixed"] = {
            "value": spsm, "unit": "steps/s", "speedup_vs_fp64": spsm / sps,
            "ms_per_step": 1e3 * wallm / max(stm.steps_done, 1),
            "dtype": "f32 x-z transforms, hybrid planes and u x omega; f64 state, "
                     "KMM reconstruction, y operators, RK update, statistics sums",
            "kernel_ms_per_launch": {k: ms / cnt for k, (ms, cnt) in timm.items() if cnt},
            "step_roofline_frac": step_bytes(nc, "mixed") * spsm / 1e9 / HBM_PEAK_GBS,
            "step_bytes_B_mix": step_bytes(nc, "mixed"),
            "rel_err_vs_fp64": mixed_error(snf, ic, grid, a.Re, a.explicit, a.cmp_steps)
            if a.cmp_steps > 0 else None,
        }
        # the opt-in variant with FP32 line transforms in the explicit y substeps
        prev = os.environ.get("SNF_Y32")
        os.environ["SNF_Y32"] = "1"
        try:
            sty, wally, timy, _ = timed_leg(snf, ic, grid, a.Re, a.explicit, a.steps, a.warmup, "mixed")
            spsy = sty.steps_done / wally
            out["mixed"]["fp32_y_transforms"] = {
                "value": spsy, "unit": "steps/s", "speedup_vs_fp64": spsy / sps,
                "switch": "SNF_Y32=1",
                "kernel_ms_per_launch": {k: ms / cnt for k, (ms, cnt) in timy.items() if cnt},
                "rel_err_vs_fp64": mixed_error(snf, ic, grid, a.Re, a.explicit, a.cmp_steps)
                if a.cmp_steps > 0 else None,
            }
        finally:
            if prev is None:
                del os.environ["SNF_Y32"]
            else:
                os.environ["SNF_Y32"] = prev
    if a.grid2 and not a.no_mixed and a.precision == "fp64":
        out["grid2"] = grid2_leg(snf, make_ic, tuple(int(x) for x in a.grid2.split(",")), a.Re,
                                 a.explicit, a.grid2_steps, a.warmup, a.cmp_steps, seed)
    if a.batch > 0 and a.precision == "fp64":
        out["ensemble"] = batch_leg(snf, make_ic, a.batch, a.batch_steps, seed)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
