"""Throughput of the TRiSK shallow-water RK4 step on one MI355X.

Workload: create_hexagon_mesh(9; centroidal=true) (2,621,442 cells), TRiSK PV flux,
cell kinetic energy, Earth parameters on the unit sphere, seeded synthetic state,
RK4 dt = 45 s.  A "step" is one RK4 step of the whole mesh (four tendencies and the
stage updates, 12 kernel launches) with the state resident in HBM.  The FP64 path and
the FP32-bulk path (FP32 stencils, FP64 state/heights/accumulation) are timed in the
same run; ``value`` is the FP32-bulk throughput in cell-steps/s.

Beside the headline the line carries the batched ensemble, the adjoint, the ICON triangle
model (ICONShallowWater at the cell count of ICON's R02B08) and the MPAS-Gassmann stage
operators (hexagon 7 x 30 levels) measurements.

Prints one JSON line (rank 0).  The CPU baseline is the oracle (FP64 restatement of the
reference kernels, OpenMP over the host cores) on a bounded sample of the same
workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK = 8.0e12                      # B/s, MI355X HBM3E
MODEL_BYTES = {"fp64": 4064.0, "fp64_pair": 4064.0, "fp32_bulk": 2704.0}  # model bytes per cell-step
METRIC = "cell-steps/s at SCVT lvl9, FP32-bulk vs FP64 (same GPU) + % of HBM peak"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def per_s(x, ms):
    """x per second over ``ms`` milliseconds; None when the time is missing or not positive."""
    return None if ms is None or not ms > 0 else x / (ms / 1e3)


def frac_of_peak(x, ms):
    r = per_s(x, ms)
    return None if r is None else r / HBM_PEAK


def roofline_of(stats, traffic=None, key=None):
    """The roofline entry of the dominant kernel of ``stats`` (largest total time, or
    ``key``): its algorithmic bytes per launch over its mean launch time, against the HBM
    peak; ``traffic(name)`` gives the measured HBM bytes per launch, if known."""
    dom = max(stats, key=key or (lambda k: k["avg_ms"] * k["launches"]))
    bw = per_s(dom["bytes"], dom["avg_ms"])
    t = traffic(dom["name"]) if traffic else None
    return dict(bound="hbm", kernel=dom["name"], achieved=None if bw is None else bw / 1e9, peak=HBM_PEAK / 1e9,
                unit="GB/s", frac=None if bw is None else bw / HBM_PEAK, traffic=t,
                traffic_frac=None if t is None else frac_of_peak(t, dom["avg_ms"]),
                avg_ms=dom["avg_ms"], algorithmic_bytes=dom["bytes"])


def spherical_mesh(level, scvt, cache_dir="", log_fn=None):
    """create_spherical_mesh(level; centroidal=scvt) with Lloyd on the GPU, kept in
    ``cache_dir`` (an .npz of the vertices and triangles) so that later runs in the same
    place (profiler passes) reuse the same mesh instead of repeating Lloyd."""
    from dmplex_amd import create_spherical_mesh, plex_from_polygons

    f = Path(cache_dir) / f"{'scvt' if scvt else 'bisection'}_level{level}.npz" if cache_dir else None
    if f is not None and f.exists():
        d = np.load(f)
        x, tri = d["x"], d["tri"]
        if len(x) == 10 * 4 ** level + 2 and len(tri) == 20 * 4 ** level:
            return plex_from_polygons((np.arange(0, 3 * len(tri) + 1, 3, dtype=np.int64), tri.reshape(-1)), x)
    dm = create_spherical_mesh(level, centroidal=scvt, device=0, log=log_fn)
    if f is not None:
        f.parent.mkdir(parents=True, exist_ok=True)
        np.savez(f, x=dm.coordinates, tri=dm.cell_vertex_table())
    return dm


def hexagon_mesh(level, scvt, cache_dir="", log_fn=None):
    """create_hexagon_mesh(level; centroidal=scvt): the Voronoi dual of spherical_mesh."""
    from dmplex_amd import create_hexagon_mesh

    return create_hexagon_mesh(level, primal=spherical_mesh(level, scvt, cache_dir, log_fn))


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--level", type=int, default=9)
    p.add_argument("--dt", type=float, default=45.0)
    p.add_argument("--precision", choices=["both", "fp64", "fp32_bulk"], default="both",
                   help="both: FP64 (bit-exact tables), FP64 on the cell-pair path, and FP32-bulk")
    p.add_argument("--cpu-steps", type=int, default=100, help="RK4 steps of the CPU baseline sample (0: skip)")
    p.add_argument("--no-kernel-events", action="store_true")
    p.add_argument("--no-graphs", action="store_true", help="launch the kernels directly instead of graph replay")
    p.add_argument("--path", choices=["auto", "tables", "pair"], default="auto",
                   help="kernel path (auto: bit-exact tables for FP64, cell pairs for FP32-bulk)")
    p.add_argument("--grid-file", default="",
                   help="run on the mesh of an MPAS or ICON grid file (read_grid) instead of generating one")
    p.add_argument("--grid-dual", action="store_true",
                   help="with --grid-file: its dual mesh (the hexagons of an ICON file)")
    p.add_argument("--no-scvt", action="store_true",
                   help="skip Lloyd (same topology and kernels; for short tests of the script only)")
    p.add_argument("--mesh-cache", default="", metavar="DIR",
                   help="keep the generated meshes in DIR and reuse them (profiler passes over the SCVT mesh)")
    p.add_argument("--accuracy", type=int, default=0, metavar="DAYS",
                   help="instead of timing: run FP64 (bit-exact), FP64 cell-pair and FP32-bulk for DAYS "
                        "simulated days from the same state and report their errors against FP64 after "
                        "day 1 and day DAYS")
    p.add_argument("--oracle-steps", type=int, default=10,
                   help="accuracy mode: RK4 steps compared bitwise with the CPU oracle first")
    p.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                   help="run a plain device copy of MIB MiB first (HBM counter calibration)")
    p.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    p.add_argument("--ensemble-only", action="store_true", help="only the batched-ensemble measurement")
    p.add_argument("--ens-members", type=int, default=16, help="ensemble members per level (levels 4-7)")
    p.add_argument("--ens-steps", type=int, default=96, help="timed RK4 steps of the ensemble")
    p.add_argument("--ens-schedule", default="",
                   help="JSON list of three kernel schedules to use instead of autotuning")
    p.add_argument("--no-member-loop", action="store_true",
                   help="skip timing the members model by model")
    p.add_argument("--pv-flux", choices=["trisk", "gassmann", "gassmann_mean"], default="trisk")
    p.add_argument("--kinetic-energy", choices=["cell", "mixed", "vertex"], default="cell")
    p.add_argument("--no-adjoint", action="store_true", help="skip the adjoint measurement")
    p.add_argument("--adjoint-only", action="store_true", help="only the adjoint measurement")
    p.add_argument("--no-icon", action="store_true", help="skip the ICON triangle-model measurement")
    p.add_argument("--icon-only", action="store_true", help="only the ICON triangle-model measurement")
    p.add_argument("--icon-level", type=int, default=9, help="SCVT triangle level of the ICON measurement")
    p.add_argument("--icon-steps", type=int, default=100, help="timed RK4 steps of the ICON measurement")
    p.add_argument("--icon-dt", type=float, default=20.0, help="RK4 dt (s) of the ICON measurement")
    p.add_argument("--icon-accuracy", type=int, default=0, metavar="DAYS",
                   help="instead of timing: FP32-bulk ICON against FP64 after 1 and DAYS simulated days")
    p.add_argument("--no-mpas", action="store_true", help="skip the MPAS-Gassmann operator measurement")
    p.add_argument("--mpas-only", action="store_true", help="only the MPAS-Gassmann operator measurement")
    p.add_argument("--mpas-level", type=int, default=7, help="hexagon level of the MPAS-Gassmann measurement")
    p.add_argument("--mpas-calls", type=int, default=50, help="timed slow-tendency calls")
    return p.parse_args()


def kernel_stats(m):
    import ctypes as C

    from dmplex_amd import _device as D

    names = (C.c_char_p * 8)()
    avg = np.zeros(8)
    nl = np.zeros(8, dtype=np.int64)
    cnt = C.c_int32()
    D.check(D.lib().trisk_kernel_times(m.handle, 8, names, avg.ctypes.data_as(C.POINTER(C.c_double)),
                                       nl.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(cnt)))
    by = np.zeros(8)
    bnames = (C.c_char_p * 8)()
    cnt2 = C.c_int32()
    D.check(D.lib().trisk_kernel_bytes(m.handle, 8, bnames, by.ctypes.data_as(C.POINTER(C.c_double)),
                                       C.byref(cnt2)))
    bytes_of = {bnames[i].decode(): float(by[i]) for i in range(cnt2.value)}
    return [dict(name=names[i].decode(), avg_ms=float(avg[i]), launches=int(nl[i]),
                 bytes=bytes_of.get(names[i].decode(), 0.0)) for i in range(cnt.value)]


def load_traffic(kernel: str, precision: str):
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if present."""
    f = ROOT / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        d = json.loads(f.read_text())
        return d.get(precision, {}).get(kernel)
    except (ValueError, OSError):
        return None


def run_gpu(grid, y0, precision, args, params, path=None, trisk_cell=True, profiled=False):
    """One precision on the GPU.  ``trisk_cell``: TRiSK PV flux with cell kinetic energy, the
    workload the fixed byte model counts; ``profiled``: the level-9 workload whose measured
    HBM traffic per kernel is committed (profiles/traffic.json)."""
    import torch

    from dmplex_amd import HIPBackend, TRiSKShallowWater, diagnostics, rk4_steps_
    from dmplex_amd import _device as D

    path = path or args.path
    log(f"[bench] building {precision} model ({path} path)")
    m = TRiSKShallowWater(grid, backend=HIPBackend(0, precision), path=path, **params)
    log(f"[bench] {precision} path: {m.path_info()[1]}; int32 (wide) cells/vertices/edges {m.wide_counts()}")
    y = m.upload_state(y0)
    mass0, energy0, _ = diagnostics(m, y)
    stream = torch.cuda.current_stream()

    def timed(steps):
        t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        wall = time.perf_counter()
        t0.record(stream)
        rk4_steps_(y, args.dt, steps, m)
        t1.record(stream)
        torch.cuda.synchronize()
        return t0.elapsed_time(t1), time.perf_counter() - wall

    # timed run 1: direct launches with HIP events around every kernel (per-kernel
    # averages for the roofline); timed run 2: the same K steps replayed as HIP graphs
    rk4_steps_(y, args.dt, args.warmup, m)
    stats = []
    ms_direct = None
    if not args.no_kernel_events:
        D.check(D.lib().trisk_set_timing(m.handle, 1))
        ms_direct, _ = timed(args.steps)
        D.check(D.lib().trisk_set_timing(m.handle, 0))
        stats = kernel_stats(m)
    if not args.no_graphs:
        m.use_graphs(True)
        rk4_steps_(y, args.dt, 8, m)  # capture
    ms, wall = timed(args.steps)
    mass, energy, _ = diagnostics(m, y)
    nc = grid.num_cells()
    step_ms = ms / args.steps
    value = per_s(nc * args.steps, ms)
    # the kernels' own algorithmic bytes (trisk_kernel_bytes) per step, and the HBM bytes the
    # counters measured for the same launches (profiles/traffic.json, level-9 workload only)
    kbytes = sum(k["bytes"] * k["launches"] for k in stats) / args.steps if stats else None
    tkey = "fp64_pair" if precision == "fp64" and path == "pair" else precision
    traffic = (lambda name: load_traffic(name, tkey)) if profiled else None
    mbytes = None
    if stats and traffic and all(traffic(k["name"]) for k in stats):
        mbytes = sum(traffic(k["name"]) * k["launches"] for k in stats) / args.steps
    out = dict(value=value, ms_per_step=step_ms, wall_s=wall, path=m.path_info()[1],
               kernel_bytes_per_cell_step=None if kbytes is None else kbytes / nc,
               hbm_frac=None if kbytes is None else frac_of_peak(kbytes, step_ms),
               measured_bytes_per_cell_step=None if mbytes is None else mbytes / nc,
               hbm_frac_measured=None if mbytes is None else frac_of_peak(mbytes, step_ms),
               reference_layout_bytes_per_cell_step=MODEL_BYTES[precision] if trisk_cell else None,
               hbm_frac_reference_layout_effective=MODEL_BYTES[precision] * value / HBM_PEAK
               if trisk_cell and value else None,
               graphs=not args.no_graphs,
               ms_per_step_direct_with_events=None if ms_direct is None else ms_direct / args.steps,
               path_name=path, mass_drift=(mass - mass0) / mass0, energy_drift=(energy - energy0) / energy0,
               kernels=stats)
    if stats:
        out["roofline"] = roofline_of(stats, traffic)
    frac, mfrac = out["hbm_frac"], out["hbm_frac_measured"]
    log(f"[bench] {precision}: {step_ms:.3f} ms/step, {value:.3e} cell-steps/s, HBM fraction: kernel bytes "
        f"{'n/a' if frac is None else f'{frac:.3f}'}, measured {'n/a' if mfrac is None else f'{mfrac:.3f}'}")
    del m
    torch.cuda.empty_cache()
    return out


ICON_REFERENCE_MS = 7.8  # ICONShallowWater, TC2, R02B08 (5.2M cells), ms per RK4 step on one A100 (docs/STATUS.md:102-104)


def run_icon(args):
    """ICONShallowWater on the SCVT triangle mesh of level 9: 5,242,880 cells, the cell count
    of ICON's R02B08, where the reference publishes 7.8 ms per RK4 step (FP64) on one A100.
    Timed in FP64 and in FP32-bulk, each without divergence damping (the TC2 setting) and
    with ICON's fourth-order damping (divdamp_fac 0.0025, nu = 0.0025 mean-area^2 / dt)."""
    import torch

    from dmplex_amd import HIPBackend, ICONShallowWater, diagnostics, rk4_steps_
    from dmplex_amd import _device as D
    from dmplex_amd.synthetic import earth_parameters, seeded_state

    t = time.perf_counter()
    dm = spherical_mesh(args.icon_level, not args.no_scvt, args.mesh_cache)
    log(f"[bench] icon: SCVT triangle mesh level {args.icon_level} ({dm.nc} cells) in {time.perf_counter() - t:.1f} s")
    dt, steps = args.icon_dt, args.icon_steps
    nu = 0.0025 * (4 * np.pi / dm.nc) ** 2 / dt
    out = {"workload": f"ICONShallowWater on SCVT triangles level {args.icon_level} ({dm.nc} cells, "
                       f"{dm.ne} edges, {dm.nv} vertices), FP64 and FP32-bulk, RK4 dt={dt} s, seeded state, "
                       "Earth parameters",
           "reference_a100_ms_per_step": ICON_REFERENCE_MS}
    stream = torch.cuda.current_stream()
    grid = None
    runs = (("no_damping", 0.0, "fp64"), ("damping", nu, "fp64"), ("fp32_bulk_no_damping", 0.0, "fp32_bulk"),
            ("fp32_bulk_damping", nu, "fp32_bulk"))
    for key, damping, precision in runs:
        t = time.perf_counter()
        m = ICONShallowWater(grid if grid is not None else dm, divergence_damping=damping,
                             backend=HIPBackend(0, precision), **earth_parameters())
        grid = m.grid
        log(f"[bench] icon {key}: tables and upload in {time.perf_counter() - t:.1f} s")
        y = m.upload_state(seeded_state(grid))
        mass0, energy0, _ = diagnostics(m, y)
        rk4_steps_(y, dt, args.warmup, m)

        def timed(n):
            t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            torch.cuda.synchronize()
            t0.record(stream)
            rk4_steps_(y, dt, n, m)
            t1.record(stream)
            torch.cuda.synchronize()
            return t0.elapsed_time(t1)

        D.check(D.lib().trisk_set_timing(m.handle, 1))
        ms_direct = timed(steps)
        D.check(D.lib().trisk_set_timing(m.handle, 0))
        stats = kernel_stats(m)
        m.use_graphs(True)
        rk4_steps_(y, dt, 8, m)
        ms = timed(steps)
        mass, energy, _ = diagnostics(m, y)
        kbytes = sum(k["bytes"] * k["launches"] for k in stats) / steps
        tkey = ("icon" if precision == "fp64" else "icon_fp32_bulk") + ("_damping" if damping else "")
        traffic = (lambda name: load_traffic(name, tkey)) if args.icon_level == 9 and not args.no_scvt else None
        step_ms = ms / steps
        r = dict(precision=precision, value=per_s(dm.nc * steps, ms), unit="cell-steps/s", ms_per_step=step_ms,
                 ms_per_step_direct_with_events=ms_direct / steps,
                 speedup_vs_reference_a100=ICON_REFERENCE_MS / step_ms if step_ms > 0 else None,
                 kernel_bytes_per_cell_step=kbytes / dm.nc, hbm_frac=frac_of_peak(kbytes, step_ms),
                 mass_drift=(mass - mass0) / mass0, energy_drift=(energy - energy0) / energy0,
                 roofline=roofline_of(stats, traffic), kernels=stats)
        if damping:
            r["divergence_damping_nu"] = nu
        out[key] = r
        log(f"[bench] icon {key}: {r['ms_per_step']:.3f} ms/step ({r['speedup_vs_reference_a100']:.2f}x the "
            f"reference's A100), HBM fraction (kernel bytes) {r['hbm_frac']:.3f}")
        del m
        torch.cuda.empty_cache()
    return out


MPAS_REFERENCE_VN_MS = 4.5  # vn_tendency per step (two calls), hexagon 7, 30 layers, one A100 (docs/EXAMPLES.md:238-249)


def run_mpas(args):
    """The explicit stage tendencies of MPAS-Gassmann (slow_tendencies!, pressure 0 as in the
    ICON-IAP step) on the SCVT hexagon mesh of level 7 with 30 layers to 30 km, FP64.  The
    reference publishes 4.5 ms of vn_tendency per step (two calls) on one A100 there."""
    import torch

    from dmplex_amd import TRiSKGrid
    from dmplex_amd.mpas_ops import MPASOperators, vertical_grid

    t = time.perf_counter()
    dm = hexagon_mesh(args.mpas_level, not args.no_scvt, args.mesh_cache)
    g = TRiSKGrid(dm, centers=dm.generators)
    v = vertical_grid(30, 30000.0)
    m = MPASOperators(g, v)
    log(f"[bench] mpas: hexagon {args.mpas_level} x 30 levels, tables and upload in {time.perf_counter() - t:.1f} s")
    rng = np.random.default_rng(5)
    nl, nc, ne = 30, g.num_cells(), g.num_edges()
    rho = np.exp(-v.z_full / 8000.0)[:, None] * (1 + 0.05 * rng.random((nl, nc)))
    rt = rho * 300.0 * np.exp(v.z_full / 20000.0)[:, None]
    w = 0.5 * rng.standard_normal((nl + 1, nc))
    w[0] = w[-1] = 0
    dev = m.upload_state(20 * rng.standard_normal((nl, ne)), w, rho, rt)
    for _ in range(3):
        m.slow_tendencies(*dev, pressure=0.0)
    m.set_timing(True)
    for _ in range(args.mpas_calls):
        m.slow_tendencies(*dev, pressure=0.0)
    stats = m.kernel_stats()
    m.set_timing(False)
    stream = torch.cuda.current_stream()
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record(stream)
    for _ in range(args.mpas_calls):
        m.slow_tendencies(*dev, pressure=0.0)
    t1.record(stream)
    torch.cuda.synchronize()
    ms = t0.elapsed_time(t1) / args.mpas_calls
    vn_ms = next(s["avg_ms"] for s in stats if s["name"] == "mpas_vn_tendency_kernel")
    kbytes = sum(s["bytes"] for s in stats)
    traffic = (lambda name: load_traffic(name, "mpas")) if args.mpas_level == 7 and not args.no_scvt else None
    out = dict(workload=f"MPAS-Gassmann slow_tendencies! (ICON-IAP PV, mass-flux vertical coupling, five-band "
                        f"solver's K_w order), SCVT hexagon {args.mpas_level} ({nc} cells, {ne} edges) x 30 levels, "
                        f"FP64, flat ground", ms_per_call=ms, unit="ms per slow_tendencies! call",
               vn_tendency_ms_per_call=vn_ms, vn_tendency_ms_per_iap_step=2 * vn_ms,
               reference_a100_vn_tendency_ms_per_step=MPAS_REFERENCE_VN_MS,
               vn_tendency_speedup_vs_reference_a100=MPAS_REFERENCE_VN_MS / (2 * vn_ms) if vn_ms > 0 else None,
               hbm_frac=frac_of_peak(kbytes, ms),
               roofline=roofline_of(stats, traffic, key=lambda k: k["avg_ms"]), kernels=stats)
    log(f"[bench] mpas: {ms:.3f} ms per call, vn_tendency {vn_ms:.3f} ms ({out['vn_tendency_speedup_vs_reference_a100']:.1f}x "
        f"the reference's A100 per step)")
    del m
    torch.cuda.empty_cache()
    return out


ENS_LEVELS = (4, 5, 6, 7)
ENS_DTS = (1200.0, 600.0, 300.0, 150.0)


def run_ensemble(args):
    """Members on SCVT levels 4-7 (per-level dt 1200/600/300/150 s), FP32-bulk, in one
    launch per kernel; beside it the same members stepped model by model."""
    import torch

    from dmplex_amd import HIPBackend, ModelTables, TRiSKEnsemble, TRiSKShallowWater, rk4_steps_
    from dmplex_amd.synthetic import earth_parameters, seeded_state

    params = earth_parameters()
    M = args.ens_members
    t = time.perf_counter()
    tabs = []
    for n in ENS_LEVELS:
        dm = hexagon_mesh(n, True, args.mesh_cache)
        tabs.append(ModelTables(dm, centers=dm.generators, **params))
    ens = TRiSKEnsemble(tabs, [M] * len(tabs), backend=HIPBackend(0, "fp32_bulk"))
    y = ens.zeros()
    seeds = {}
    j = 0
    for li, tb in enumerate(tabs):
        for m in range(M):
            seeds[li, m] = j
            ens.set_member(y, li, m, seeded_state(tb.grid, seed_coef=1000 + j, seed_noise=2000 + j))
            j += 1
    log(f"[ensemble] {len(seeds)} members on levels {ENS_LEVELS} ready in {time.perf_counter() - t:.1f} s, "
        f"{ens.cells()} cell-members")
    d0 = ens.diagnostics(y)
    stream = torch.cuda.current_stream()

    def timed(fn):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1)

    t = time.perf_counter()
    if args.ens_schedule:
        ens.set_schedule(json.loads(args.ens_schedule))
        sched, tune_ms = ens.schedule(), None
    else:
        sched, tune_ms = ens.autotune(y, ENS_DTS, steps=4)
    log(f"[ensemble] schedule ({time.perf_counter() - t:.1f} s): {sched}")
    K = args.ens_steps
    ens.rk4_steps_(y, ENS_DTS, 8)
    ens.set_timing(True)
    ms_direct = timed(lambda: ens.rk4_steps_(y, ENS_DTS, K))
    ens.set_timing(False)
    stats = ens.kernel_stats()
    ens.use_graphs(True)
    ens.rk4_steps_(y, ENS_DTS, 8)
    ms = timed(lambda: ens.rk4_steps_(y, ENS_DTS, K))
    d1 = ens.diagnostics(y)
    cells = ens.cells()
    value = per_s(cells * K, ms)
    step_bytes = 4 * sum(k["bytes"] for k in stats)
    log(f"[ensemble] {ms / K:.3f} ms/step, {value:.3e} cell-member-steps/s")
    out = {
        "workload": f"{len(seeds)} members, {M} each on SCVT levels {list(ENS_LEVELS)}, dt {list(ENS_DTS)} s, "
                    f"FP32-bulk, TRiSK PV flux, cell KE",
        "unit": "cell-member-steps/s", "value": value, "ms_per_step": ms / K, "steps": K,
        "ms_per_step_direct_with_events": ms_direct / K, "graphs": True,
        "cell_members": cells, "schedule": sched, "autotune_kernel_ms": tune_ms,
        # one launch per kernel for every member of every level; an RK stage is three such
        # launches (state, cell, edge), each of which gathers what the one before wrote
        # anywhere on the meshes; "fused_stage" below times the one-launch stage (cooperative,
        # grid-wide barriers), which is slower
        "launches_per_rk_stage": 3,
        "persistent_kernels": [k for k, sc in zip(("state", "cell", "edge"), sched) if sc.get("blocks_per_cu")],
        "kernel_bytes_per_step": step_bytes, "hbm_frac": frac_of_peak(step_bytes, ms / K),
        "roofline": roofline_of(stats, key=lambda k: k["avg_ms"]),
        "kernels": stats,
        "max_member_mass_drift": float(np.max(np.abs((d1[:, 0] - d0[:, 0]) / d0[:, 0]))),
    }
    # one cooperative launch per RK stage (state, cell and edge phases with grid-wide barriers
    # between them), direct launches; every member rounds as in the three-launch stage
    try:
        ens.use_fused(True)
    except (ValueError, RuntimeError) as exc:
        out["fused_stage"] = {"available": False, "why": str(exc)}
    else:
        ens.rk4_steps_(y, ENS_DTS, 4)
        ens.set_timing(True)
        msf_direct = timed(lambda: ens.rk4_steps_(y, ENS_DTS, K))
        ens.set_timing(False)
        fstats = ens.kernel_stats()
        msf = timed(lambda: ens.rk4_steps_(y, ENS_DTS, K))
        ens.use_fused(False)
        fbytes = 4 * sum(k["bytes"] for k in fstats)
        out["fused_stage"] = {"available": True, "launches_per_rk_stage": 1, "graphs": False,
                              "value": per_s(cells * K, msf), "ms_per_step": msf / K,
                              "ms_per_step_direct_with_events": msf_direct / K,
                              "hbm_frac": frac_of_peak(fbytes, msf / K), "kernels": fstats}
        log(f"[ensemble] fused stage (one cooperative launch per RK stage): {msf / K:.3f} ms/step")
    if args.no_member_loop:
        return out
    # the same members, one model per level, stepped member after member (graph replay)
    loop_ms = 0.0
    for li, tb in enumerate(tabs):
        m = TRiSKShallowWater(tb.grid, backend=HIPBackend(0, "fp32_bulk"), path="pair", **params)
        m.use_graphs(True)
        states = [m.upload_state(seeded_state(tb.grid, seed_coef=1000 + seeds[li, k], seed_noise=2000 + seeds[li, k]))
                  for k in range(M)]
        for ys in states:
            rk4_steps_(ys, ENS_DTS[li], 8, m)  # capture and warm-up
        loop_ms += timed(lambda: [rk4_steps_(ys, ENS_DTS[li], K, m) for ys in states])
        del m, states
    torch.cuda.empty_cache()
    loop_value = per_s(cells * K, loop_ms)
    log(f"[ensemble] member-by-member loop: {loop_ms / K:.3f} ms/step, {loop_value:.3e} cell-member-steps/s")
    out["member_loop"] = {"value": loop_value, "ms_per_step": loop_ms / K,
                          "how": "one FP32-bulk cell-pair model per level, each member stepped K steps in turn "
                                 "(graph replay)"}
    out["speedup_vs_member_loop"] = value / loop_value if value and loop_value else None
    return out


def run_adjoint(args):
    """Gradient of J = sum_i A_i (h_i(T) - h*_i)^2 after 40 RK4 steps of 150 s on SCVT level 7,
    by the checkpointed adjoint (checkpoints every 8 steps), FP64 and FP32-bulk; beside it the
    plain forward run and four central-difference checks (eps = 1e-6)."""
    import torch

    from dmplex_amd import HIPBackend, TRiSKShallowWater, height_misfit, rk4_adjoint, rk4_steps, rk4_steps_
    from dmplex_amd.synthetic import earth_parameters, seeded_state

    dm = hexagon_mesh(7, True, args.mesh_cache)
    params = earth_parameters()
    dt, N, ck = 150.0, 40, 8
    out = {"workload": f"SCVT level 7 ({dm.nc} cells), J = sum A (h_T - h*)^2 after {N} RK4 steps of {dt} s, "
                       f"checkpoints every {ck} steps", "unit": "ms"}
    grads = {}
    for prec in ("fp64", "fp32_bulk"):
        m = TRiSKShallowWater(dm, centers=dm.generators, path="pair", backend=HIPBackend(0, prec), **params)
        m.use_graphs(True)  # forward steps and the adjoint sweep replay captured graphs
        y0 = m.upload_state(seeded_state(m.grid))
        hstar = m.upload_state(seeded_state(m.grid, seed_coef=3, seed_noise=4))[m.grid.num_edges():].clone()
        y = y0.clone()
        rk4_steps_(y, dt, N, m)  # warm-up
        torch.cuda.synchronize()
        t = time.perf_counter()
        reps = 5
        for _ in range(reps):
            y = y0.clone()
            rk4_steps_(y, dt, N, m)
        torch.cuda.synchronize()
        fwd = (time.perf_counter() - t) / reps * 1e3
        J0, ybar = height_misfit(m, y, hstar)
        rk4_adjoint(m, y0, dt, N, ybar, ck)  # warm-up (allocates the adjoint work)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            g, _ = rk4_adjoint(m, y0, dt, N, ybar, ck)
        torch.cuda.synchronize()
        grad_ms = (time.perf_counter() - t) / reps * 1e3
        grads[prec] = g
        fd = []
        if prec == "fp64":
            gen = torch.Generator(device="cuda").manual_seed(5)
            ne = m.grid.num_edges()
            su, sh = float(torch.std(y0[:ne])), float(torch.std(y0[ne:]))
            for _ in range(4):
                v = torch.randn(m.state_length, dtype=torch.float64, device="cuda", generator=gen)
                v[:ne] *= su
                v[ne:] *= sh
                eps = 1e-6
                Jp = height_misfit(m, rk4_steps(m, y0 + eps * v, dt, N), hstar)[0]
                Jm = height_misfit(m, rk4_steps(m, y0 - eps * v, dt, N), hstar)[0]
                ad = float(g @ v)
                fd.append(abs((Jp - Jm) / (2 * eps) - ad) / abs(ad))
        out[prec] = {"forward_ms": fwd, "gradient_ms": grad_ms, "gradient_over_forward": grad_ms / fwd, "J": J0,
                     "fd_max_rel_diff": max(fd) if fd else None}
        log(f"[adjoint] {prec}: forward {fwd:.2f} ms, gradient {grad_ms:.2f} ms ({grad_ms / fwd:.2f}x)")
        del m
    a, b = grads["fp32_bulk"].cpu().numpy(), grads["fp64"].cpu().numpy()
    out["fp32_bulk_gradient_rel_l2_vs_fp64"] = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    torch.cuda.empty_cache()
    return out


def run_cpu(grid, y0, args, params):
    from dmplex_amd import ModelTables
    from oracle.oracle import Oracle

    t = ModelTables(grid, **params)
    o = Oracle(t)
    threads = o.threads()
    y = y0.copy()
    o.tendency(y)  # warm-up tendency, as the reference's timing driver does
    wall = time.perf_counter()
    o.rk4_steps(y, args.dt, args.cpu_steps)
    wall = time.perf_counter() - wall
    nc = grid.num_cells()
    return dict(value=nc * args.cpu_steps / wall, unit="cell-steps/s", cores=threads, kind="port",
                sample=f"{args.cpu_steps} RK4 steps of the level-{args.level} workload (FP64 oracle: the "
                       f"reference's kernel decomposition, int64 column-major tables, OpenMP), {wall:.1f} s")


def field_errors(m, a, b):
    """Max and L2 errors of state a against the reference state b (host, reference order)."""
    g = m.grid
    ne = g.num_edges()
    ua, ub, ha, hb = a[:ne], b[:ne], a[ne:], b[ne:]
    A = g.cell_areas
    w = g.primal_lengths * g.dual_lengths
    return dict(
        h_max=float(np.max(np.abs(ha - hb)) / np.max(np.abs(hb))),
        h_l2=float(np.sqrt(np.sum(A * (ha - hb) ** 2) / np.sum(A * hb ** 2))),
        u_max=float(np.max(np.abs(ua - ub)) / np.max(np.abs(ub))),
        u_l2=float(np.sqrt(np.sum(w * (ua - ub) ** 2) / np.sum(w * ub ** 2))))


def run_icon_accuracy(args):
    """FP32-bulk ICON (with ICON's divergence damping) against the FP64 trajectory from the
    seeded state after 1 and ``args.icon_accuracy`` simulated days; the FP64 trajectory is
    first checked bitwise equal to the oracle over a few steps."""
    import torch

    from dmplex_amd import HIPBackend, ICONShallowWater, create_spherical_mesh, diagnostics, rk4_steps_
    from dmplex_amd.synthetic import earth_parameters, seeded_state
    from oracle.oracle import IconOracle

    dm = create_spherical_mesh(args.icon_level, centroidal=not args.no_scvt, device=0)
    dt = args.icon_dt
    nu = 0.0025 * (4 * np.pi / dm.nc) ** 2 / dt
    models = {"fp64": ICONShallowWater(dm, divergence_damping=nu, **earth_parameters())}
    grid = models["fp64"].grid
    models["fp32_bulk"] = ICONShallowWater(grid, divergence_damping=nu, backend=HIPBackend(0, "fp32_bulk"),
                                           **earth_parameters())
    y0 = seeded_state(grid)
    n_or = args.oracle_steps
    yo = IconOracle(models["fp64"]).rk4_steps(y0.copy(), dt, n_or)
    yg = models["fp64"].upload_state(y0)
    rk4_steps_(yg, dt, n_or, models["fp64"])
    report = {"fp64_bitwise_equal_to_oracle": {"steps": n_or,
                                               "equal": bool(np.array_equal(models["fp64"].download_state(yg), yo))}}
    log(f"[icon accuracy] FP64 GPU vs oracle after {n_or} steps: {report['fp64_bitwise_equal_to_oracle']}")
    states, mass0, energy0 = {}, {}, {}
    for key, m in models.items():
        m.use_graphs(True)
        states[key] = m.upload_state(y0)
        mass0[key], energy0[key], _ = diagnostics(m, states[key])
    spd = int(round(86400.0 / dt))
    t0 = time.perf_counter()
    for day in range(1, args.icon_accuracy + 1):
        for key, m in models.items():
            rk4_steps_(states[key], dt, spd, m)
        torch.cuda.synchronize()
        log(f"[icon accuracy] day {day} done ({time.perf_counter() - t0:.0f} s)")
        if day not in (1, args.icon_accuracy):
            continue
        ref = models["fp64"].download_state(states["fp64"])
        entry = {}
        for key, m in models.items():
            mass, energy, _ = diagnostics(m, states[key])
            e = {"mass_drift": (mass - mass0[key]) / mass0[key], "energy_drift": (energy - energy0[key]) / energy0[key]}
            if key != "fp64":
                e.update(field_errors(m, m.download_state(states[key]), ref))
            entry[key] = e
        report[f"day_{day}"] = entry
    return {"accuracy": report, "level": args.icon_level, "cells": dm.nc, "dt": dt, "steps_per_day": spd,
            "divergence_damping_nu": nu, "reference": "FP64 ICON (bitwise equal to the FP64 oracle)"}


def run_accuracy(args, grid, y0, params):
    import torch

    from dmplex_amd import HIPBackend, TRiSKShallowWater, diagnostics, rk4_steps_, vertex_pv_

    runs = {"fp64": ("fp64", "auto"), "fp64_pair": ("fp64", "pair"), "fp32_bulk": ("fp32_bulk", "auto")}
    if params.get("pv_flux", "trisk") != "trisk" or params.get("kinetic_energy", "cell") != "cell":
        del runs["fp64_pair"]  # the cell-pair path covers the TRiSK PV flux with cell kinetic energy only
    models, states, mass0, energy0 = {}, {}, {}, {}
    for key, (prec, path) in runs.items():
        m = TRiSKShallowWater(grid, backend=HIPBackend(0, prec), path=path, **params)
        m.use_graphs(True)
        models[key] = m
        states[key] = m.upload_state(y0)
        mass0[key], energy0[key], _ = diagnostics(m, states[key])
    spd = int(round(86400.0 / args.dt))
    report = {}
    # the FP64 reference trajectory is the FP64 table path; tie it to the oracle (the CPU
    # restatement of the reference) on the first steps of the same state
    from dmplex_amd import ModelTables
    from oracle.oracle import Oracle

    n_or = args.oracle_steps
    yo = Oracle(ModelTables(grid, **params)).rk4_steps(y0.copy(), args.dt, n_or)
    yg = models["fp64"].upload_state(y0)
    rk4_steps_(yg, args.dt, n_or, models["fp64"])
    report["fp64_bitwise_equal_to_oracle"] = {"steps": n_or,
                                              "equal": bool(np.array_equal(models["fp64"].download_state(yg), yo))}
    log(f"[accuracy] FP64 GPU vs oracle after {n_or} steps: {report['fp64_bitwise_equal_to_oracle']}")
    t0 = time.perf_counter()
    for day in range(1, args.accuracy + 1):
        for key, m in models.items():
            rk4_steps_(states[key], args.dt, spd, m)
        torch.cuda.synchronize()
        log(f"[accuracy] day {day} done ({time.perf_counter() - t0:.0f} s)")
        if day not in (1, args.accuracy):
            continue
        ref_m = models["fp64"]
        ref = ref_m.download_state(states["fp64"])
        q_ref = vertex_pv_(ref_m, states["fp64"])[0].cpu().numpy()
        Av = grid.vertex_areas
        entry = {}
        for key, m in models.items():
            mass, energy, _ = diagnostics(m, states[key])
            e = {"mass_drift": (mass - mass0[key]) / mass0[key],
                 "energy_drift": (energy - energy0[key]) / energy0[key]}
            if key != "fp64":
                e.update(field_errors(m, m.download_state(states[key]), ref))
                # PV of both states from the same FP64 diagnostic kernel, vertices in device order
                q = vertex_pv_(ref_m, ref_m.upload_state(m.download_state(states[key])))[0].cpu().numpy()
                e["q_max"] = float(np.max(np.abs(q - q_ref)) / np.max(np.abs(q_ref)))
                e["q_l2"] = float(np.sqrt(np.sum(Av[ref_m.permutation()[2]] * (q - q_ref) ** 2) /
                                          np.sum(Av[ref_m.permutation()[2]] * q_ref ** 2)))
            entry[key] = e
        report[f"day_{day}"] = entry
    return {"accuracy": report, "level": args.level, "dt": args.dt, "steps_per_day": spd,
            "reference": "FP64 table path (bitwise equal to the FP64 oracle)"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (multi-GPU halo exchange is "
                                   "design-only)"}))
        sys.exit(2)
    import torch

    from dmplex_amd import TRiSKGrid
    from dmplex_amd.synthetic import earth_parameters, seeded_state

    if args.calibrate_copy:
        n = args.calibrate_copy * 2 ** 20 // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy} MiB done")
        del a, b
    if args.ensemble_only:
        print(json.dumps({"ensemble": run_ensemble(args)}), flush=True)
        return
    if args.adjoint_only:
        print(json.dumps({"adjoint": run_adjoint(args)}), flush=True)
        return
    if args.icon_accuracy:
        print(json.dumps({"icon_accuracy": run_icon_accuracy(args)}), flush=True)
        return
    if args.icon_only:
        print(json.dumps({"icon": run_icon(args)}), flush=True)
        return
    if args.mpas_only:
        print(json.dumps({"mpas": run_mpas(args)}), flush=True)
        return
    t = time.perf_counter()
    if args.grid_file:
        from dmplex_amd import read_grid

        dm = read_grid(args.grid_file, dual=args.grid_dual)
        log(f"[bench] read {args.grid_file} ({dm.nc} cells) in {time.perf_counter() - t:.2f} s")
        grid = TRiSKGrid(dm)
        workload = f"grid file {os.path.basename(args.grid_file)}{' (dual)' if args.grid_dual else ''}"
    else:
        log(f"[bench] generating SCVT hexagon mesh level {args.level}")
        dm = hexagon_mesh(args.level, not args.no_scvt, args.mesh_cache, log_fn=lambda m: log(f"[bench] {m}"))
        grid = TRiSKGrid(dm, centers=dm.generators)
        workload = f"SCVT hexagon level {args.level}"
    params = dict(earth_parameters(), pv_flux=args.pv_flux, kinetic_energy=args.kinetic_energy)
    trisk_cell = args.pv_flux == "trisk" and args.kinetic_energy == "cell"
    y0 = seeded_state(grid)
    log(f"[bench] mesh + tables + state in {time.perf_counter() - t:.1f} s: "
        f"{grid.num_cells()} cells, {grid.num_edges()} edges, {grid.num_vertices()} vertices")
    if args.accuracy:
        print(json.dumps(run_accuracy(args, grid, y0, params)), flush=True)
        return
    if args.precision == "both":
        runs = [("fp64", "fp64", None), ("fp64_pair", "fp64", "pair"), ("fp32_bulk", "fp32_bulk", None)]
        if not trisk_cell:  # the cell-pair path covers the TRiSK PV flux with cell kinetic energy only
            runs = [r for r in runs if r[0] != "fp64_pair"]
    else:
        runs = [(args.precision, args.precision, None)]
    profiled = trisk_cell and args.level == 9 and not args.grid_file and not args.no_scvt
    res = {key: run_gpu(grid, y0, prec, args, params, path, trisk_cell, profiled) for key, prec, path in runs}
    head = res.get("fp32_bulk", res.get("fp64"))
    head_p = "fp32_bulk" if "fp32_bulk" in res else "fp64"
    cpu = None
    if args.cpu_steps > 0:
        log(f"[bench] CPU baseline: {args.cpu_steps} RK4 steps on the host")
        cpu = run_cpu(grid, y0, args, params)
        log(f"[bench] CPU baseline {cpu['value']:.3e} cell-steps/s on {cpu['cores']} threads")
    line = {
        "metric": METRIC, "value": head["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if head_p == "fp32_bulk" else "f64",
        "data": "synthetic: seeded streamfunction state (l<=8 harmonics, max|u|=30 m/s, linear balance, "
                "N(0,5 m) noise), Earth parameters on the unit sphere",
        "config": {"workload": f"{workload}, {args.pv_flux} PV flux, {args.kinetic_energy} KE, RK4 dt={args.dt} s",
                   "cells": grid.num_cells(), "edges": grid.num_edges(), "vertices": grid.num_vertices(),
                   "precision": head_p,
                   "kernel_bytes_per_cell_step": head["kernel_bytes_per_cell_step"],
                   "reference_layout_bytes_per_cell_step": head["reference_layout_bytes_per_cell_step"]},
        # % of HBM peak: the kernels' own algorithmic bytes (each element a launch must touch,
        # counted once) per step over the step time; beside it the HBM bytes the counters
        # measured for the same launches, and the effective rate of the reference's table
        # layout (bytes the reference's kernels would move for the same cell-steps)
        "hbm_frac": head["hbm_frac"],
        "hbm_frac_measured": head["hbm_frac_measured"],
        "hbm_frac_reference_layout_effective": head["hbm_frac_reference_layout_effective"],
        "roofline": head.get("roofline"),
        "cpu_baseline": cpu,
    }
    for p, r in res.items():
        line[p] = {k: v for k, v in r.items() if k != "kernels"}
        line[p]["kernels"] = r["kernels"]
    if "fp64" in res and "fp32_bulk" in res:
        line["speedup_fp32_bulk_vs_fp64"] = res["fp32_bulk"]["value"] / res["fp64"]["value"]
    if "fp64_pair" in res and "fp32_bulk" in res:
        line["speedup_fp32_bulk_vs_fp64_pair"] = res["fp32_bulk"]["value"] / res["fp64_pair"]["value"]
    if not args.no_ensemble and trisk_cell:
        log("[bench] batched ensemble (levels 4-7)")
        line["ensemble"] = run_ensemble(args)
    if not args.no_adjoint and trisk_cell:
        log("[bench] adjoint (level 7, 40 steps)")
        line["adjoint"] = run_adjoint(args)
    if not args.no_icon and trisk_cell and not args.grid_file:
        log(f"[bench] ICON triangle model (level {args.icon_level})")
        line["icon"] = run_icon(args)
    if not args.no_mpas and trisk_cell and not args.grid_file:
        log(f"[bench] MPAS-Gassmann stage operators (hexagon {args.mpas_level} x 30 levels)")
        line["mpas"] = run_mpas(args)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
