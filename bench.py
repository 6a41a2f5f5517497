"""Benchmark of the full coupled super-droplet timestep on one MI355X.

Workload (the metric's configuration): 128 x 128 x 128 = 2,097,152 gridboxes of 25 m cubes,
Poisson(64) superdroplets per gridbox (~134.2M drops, SoA FP64/u64), each step = predictor-
corrector motion under the 2-D divergence-free updraft (peak w 2 m/s) + stable sort by
sdgbxindex + implicit-Euler condensation (FP64 Newton, rtol 1e-6, atol 1e-8, maxniters 50) +
Long-kernel collision-coalescence/breakup (nfrags 5, coaleff 0.85), dt = 1 s, replayed as a
hipGraph.  Inputs are synthetic and already resident in HBM when the timed region starts.

Prints ONE JSON line (rank 0): superdroplet-steps/s, roofline of the dominant kernel measured
with HIP events on the library's stream, and the CPU baseline (the OpenMP build of the
reference-structured oracle) timed on this host on a bounded sample of the same workload.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the CPUs this process may use, read before an OpenMP runtime binds the thread to one core
HOST_CPUS = len(os.sched_getaffinity(0))
# the CPU baseline's OpenMP threads: one per core, packed (set before any OpenMP runtime starts)
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

METRIC = "superdroplet-steps/sec on 1 GPU (plain vs MLP warm-start), Newton iters/drop"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6

# algorithmic HBM bytes (per drop, per gridbox) of each stage's main kernel, by variant
BYTES = {
    # read r, msol, xi; write r | read State (press, temp, qv, qc, vol), refs; write temp, qv, qc
    "k_condensation": (32, 5 * 8 + 8 + 3 * 8),
    # read sdgbxindex, r, msol, xi; write r, dm | saturation ratio and diffusion factor, temp, refs
    "k_cond_drops": (44, 2 * 8 + 8 + 8),
    "k_cond_queue": (44, 2 * 8 + 8 + 8),
    # read xi, r, msol of every paired drop; write its slot of the shuffled order (the drops stay
    # in place; enacted pairs, a few per cent, also write their attributes) | refs, nact, volume
    "k_collisions": (28, 8 + 4 + 8),
    "k_coll_tiles": (28, 8 + 4 + 8),
    # read coords, radius, sdgbxindex; write coords, sdgbxindex | bounds, winds, counts
    "k_motion": (64, 6 * 8 + 6 * 8 + 28 * 4),
    "k_motion_flat": (64, 6 * 8 + 6 * 8 + 28 * 4),
    # fused motion + sort, moving pass: read coords, radius, xi, msol, id; write the 60-byte record
    # at its sorted place | bounds, winds, refs, run starts
    "k_motion_scatter": (112, 6 * 8 + 6 * 8 + 8 + 4 * 8),
    # counting pass of the motion in the sort's gather: read coords, radius, sdgbxindex; write the
    # step | bounds, winds, counts
    "k_motion_keys": (37, 6 * 8 + 6 * 8 + 28 * 4),
    # the sort's gather with the condensation solves: read the source position and the 60-byte
    # record; write the record, dm and qm | saturation ratio, diffusion factor, temperature
    "k_sort_gather_cond": (140, 2 * 8 + 8),
}


def measured_traffic(kernel, ndims):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/traffic.json: KiB per dispatch, each scaled by what the counter reads for a copy of
    known size), when it was measured on this workload; else None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        t = json.load(f)
    if list(t.get("ndims", [])) != list(ndims) or kernel not in t.get("kernels", {}):
        return None
    k = t["kernels"][kernel]
    return 1024.0 * (k["fetch_kib"] * t["fetch_scale"] + k["write_kib"] * t["write_scale"])


def stage_kernel(stage, tiles):
    """Name of the main kernel of a stage (0 condensation, 1 collisions, 2 motion) under a tiling."""
    if stage == 0:
        return {0: "k_condensation", 1: "k_cond_drops", 2: "k_cond_queue"}[tiles.cond_flat]
    if stage == 1:
        return "k_coll_tiles" if tiles.coll_flat else "k_collisions"
    return {0: "k_motion", 1: "k_motion_flat", 2: "k_motion_scatter", 3: "k_motion_keys"}[tiles.motion_flat]


def host_cpu():
    """CPU model and the cores this process may use (for the CPU baseline's record)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(model=model, cpus_available=HOST_CPUS, omp_proc_bind=os.environ.get("OMP_PROC_BIND"),
                omp_places=os.environ.get("OMP_PLACES"), omp_num_threads=os.environ.get("OMP_NUM_THREADS"))


def parse():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nz", type=int, default=128, help="gridboxes along z (128 = full workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--no-mlp", action="store_true", help="skip the learned warm-start leg")
    ap.add_argument("--cpu-ny", type=int, default=16, help="CPU sample: 128 x 128 x ny gridboxes")
    ap.add_argument("--cpu-steps", type=int, default=4)
    ap.add_argument("--tiles", type=str, default=None,
                    help="launch tiling as 8 comma-separated cleo_tiles fields (skips the autotuner)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the coalescence/breakup, condensation and learned-condensation workloads")
    ap.add_argument("--config-steps", type=int, default=10)
    ap.add_argument("--copy-probe", action="store_true",
                    help="negate one 1 GiB device buffer first (reference point for byte counters)")
    return ap.parse_args()


def _ptr(t, ctype):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(ctype))


def make_params(abi):
    cond = abi.CondParams(100, 1, 50, 1e-6, 1e-8, 0.001, 0, 10)
    coll = abi.CollParams(100, abi.PROB_LONG, 1.0, abi.ENACT_COALBURE, 5.0, 0.85)
    mot = abi.MotionParams(100, abi.TV_SIMMEL)
    s = abi.SDMParams()
    s.enable_cond = s.enable_colls = s.enable_motion = 1
    s.cond, s.coll, s.motion = cond, coll, mot
    return s


def run_gpu(args):
    import torch

    from cleo_amd import abi, lib
    from cleo_amd import populations as P

    b = lib.load()
    L = b.lib
    dev = torch.device("cuda:0")
    if args.copy_probe:
        # reference point for byte counters: one elementwise kernel that reads 1 GiB and writes 1 GiB
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst
    ndims = (args.nz, 128, 128)
    t_gen = time.perf_counter()
    arrs, st, halves = P.coupled_device(torch, dev, ndims=ndims)
    torch.cuda.synchronize()
    n = int(arrs["xi"].numel())
    t_gen = time.perf_counter() - t_gen
    soa = abi.SupersSoA()
    soa.sdgbxindex = _ptr(arrs["sdgbxindex"], C.c_uint32)
    soa.coord3, soa.coord1, soa.coord2 = (_ptr(arrs[k], C.c_double) for k in ("coord3", "coord1", "coord2"))
    soa.xi = _ptr(arrs["xi"], C.c_uint64)
    soa.radius, soa.msol = _ptr(arrs["radius"], C.c_double), _ptr(arrs["msol"], C.c_double)
    soa.sdid = _ptr(arrs["sdid"], C.c_uint64)
    sdm = make_params(abi)
    span = 100

    def new_domain():
        d = abi.Domain(b, n, ndims, *halves)
        d.check(L.cleo_upload_supers_device(d.h, C.byref(soa), n), "upload_supers_device")
        d.ntot = n
        d.set_state(st)
        d.set_solute(*P.solute_dimless(P.AMMONIUM_SULPHATE))
        d.set_sdm(sdm)
        d.rng_counter(20261202)
        return d

    def elapsed(d, a, z):
        ms = C.c_float()
        d.check(L.cleo_event_elapsed(d.h, a, z, C.byref(ms)), "event_elapsed")
        return float(ms.value)

    def ndomain(d):
        v = C.c_uint64()
        d.check(L.cleo_ndomain(d.h, C.byref(v)), "ndomain")
        return int(v.value)

    def timed_steps(d, t0, k):
        d.reset_stats()
        torch.cuda.synchronize()
        d.check(L.cleo_synchronize(d.h), "sync")
        nd0 = ndomain(d)
        w0 = time.perf_counter()
        L.cleo_event_record(d.h, 0)
        d.check(L.cleo_run_steps_graph(d.h, t0, span, k), "timed steps")
        L.cleo_event_record(d.h, 1)
        d.check(L.cleo_synchronize(d.h), "sync")
        torch.cuda.synchronize()
        wall = time.perf_counter() - w0
        ms = elapsed(d, 0, 1)
        nd1 = ndomain(d)
        stats = d.stats()
        return dict(ms_total=ms, wall_s=wall, stats=stats, drop_steps=0.5 * (nd0 + nd1) * k,
                    iters_per_drop=stats["newton_iters"] / max(stats["cond_drops"], 1))

    d = new_domain()
    tiles = abi.Tiles()
    t = 0
    if args.warmup:
        d.check(L.cleo_run_steps_graph(d.h, t, span, args.warmup), "warmup")
        t += args.warmup * span
    # tilings do not change results, so the choice can be made on the warmed-up population
    if args.tiles:
        tiles = abi.Tiles(*[int(v) for v in args.tiles.split(",")])
        d.check(L.cleo_set_tiles(d.h, C.byref(tiles)), "set_tiles")
    elif not args.no_autotune:
        d.check(L.cleo_autotune(d.h, C.byref(tiles)), "autotune")
    dl = None
    if not args.no_mlp:
        # identical twin for the learned leg: same upload, same warm-up, same random streams
        dl = new_domain()
        L.cleo_set_tiles(dl.h, C.byref(tiles))
        if args.warmup:
            dl.check(L.cleo_run_steps_graph(dl.h, 0, span, args.warmup), "warmup learned")
    del arrs, soa
    torch.cuda.empty_cache()

    plain = timed_steps(d, t, args.steps)

    learned = None
    if dl is not None:
        learned = learned_leg(args, b, dl, sdm, t, timed_steps, torch)
        dl.close()
    t += args.steps * span

    # per-stage eager timing on the following step (stage = the launches of one C-ABI call)
    rec = L.cleo_event_record
    rec(d.h, 2)
    d.check(d.motion(sdm.motion), "motion")
    rec(d.h, 3)
    reps = 3
    for _ in range(reps):
        d.check(d.condensation(sdm.cond), "condensation")
    rec(d.h, 4)
    d.check(d.collisions(sdm.coll), "collisions")
    rec(d.h, 5)
    stage_ms = dict(motion_sort_qcond=elapsed(d, 2, 3), condensation=elapsed(d, 3, 4) / reps,
                    collisions=elapsed(d, 4, 5))
    # main kernel of each stage alone: HIP events around its launches on the library's stream
    kernel_ms = {}
    for stage in range(3):
        ms = C.c_float()
        d.check(L.cleo_time_kernel(d.h, stage, 3, C.byref(ms)), "time_kernel")
        kernel_ms[stage_kernel(stage, tiles)] = float(ms.value)
    if tiles.cond_flat == 1 and tiles.motion_flat in (1, 3):
        # the step solves condensation inside the sort's gather (k_sort_gather_cond; with motion
        # variant 3 the drops also move in it) instead of in the separate per-drop kernel
        del kernel_ms[stage_kernel(0, tiles)]
        ms = C.c_float()
        d.check(L.cleo_time_kernel(d.h, 3, 3, C.byref(ms)), "time_kernel fused")
        kernel_ms["k_sort_gather_cond"] = float(ms.value)
    ngbx = int(np.prod(ndims))
    nd = ndomain(d)
    # the mass-moment observer's reduction on the state the steps left (not part of the step)
    ms = C.c_float()
    d.check(L.cleo_time_kernel(d.h, 4, 3, C.byref(ms)), "time_kernel mass moments")
    observer_ms = float(ms.value)
    # one coupling step of prescribed dynamics from device arrays (the State is set from them; the
    # drops are not touched and nothing runs after this)
    coupling_ms = time_receive_dynamics(d, L, ndims, torch, rec, elapsed)
    d.check(L.cleo_synchronize(d.h), "final sync")
    err = d.error()
    d.close()
    return dict(n=n, ngbx=ngbx, ndims=ndims, plain=plain, stage_ms=stage_ms, kernel_ms=kernel_ms, nd=nd, tiles=tiles,
                gen_s=t_gen, learned=learned, err=err, observer_ms=observer_ms, coupling_ms=coupling_ms)


def time_receive_dynamics(d, L, ndims, torch, rec, elapsed, reps=5):
    """Mean time of cleo_receive_dynamics over `reps` calls, HIP events on the library's stream:
    thermodynamics per gridbox and winds on the faces of the C grid, synthetic, in HBM."""
    from cleo_amd import abi
    nz, nx, ny = ndims
    ngbx = nz * nx * ny
    dev = dict(press=torch.full((ngbx,), 1.0), temp=torch.full((ngbx,), 1.05), qvap=torch.full((ngbx,), 0.01),
               qcond=torch.zeros(ngbx), wvel_zfaces=torch.zeros((nz + 1) * nx * ny),
               uvel_xfaces=torch.zeros(nz * (nx + 1) * ny), vvel_yfaces=torch.zeros(nz * nx * (ny + 1)))
    dev = {k: v.to(dtype=torch.float64, device="cuda:0") for k, v in dev.items()}
    torch.cuda.synchronize()
    step = abi.DynamicsStep()
    for k, v in dev.items():
        setattr(step, k, v.data_ptr())
    L.cleo_receive_dynamics.argtypes = [C.c_void_p, C.POINTER(abi.DynamicsStep)]
    d.check(L.cleo_receive_dynamics(d.h, C.byref(step)), "receive_dynamics")  # warm-up
    rec(d.h, 6)
    for _ in range(reps):
        d.check(L.cleo_receive_dynamics(d.h, C.byref(step)), "receive_dynamics")
    rec(d.h, 7)
    return elapsed(d, 6, 7) / reps


def learned_leg(args, b, d, sdm, t, timed_steps, torch):
    """Train the warm-start MLP on samples the plain solver produces from the current population,
    then time the same steps with the learned guess."""
    from cleo_amd import abi
    L = b.lib
    cap = 10_000_000
    xd = torch.zeros((cap, 8), dtype=torch.float32, device="cuda")
    yd = torch.zeros(cap, dtype=torch.float32, device="cuda")
    nw = C.c_uint64()
    t0 = time.perf_counter()
    d.check(L.cleo_mlp_make_samples(d.h, C.byref(sdm.cond), xd.data_ptr(), yd.data_ptr(), cap, C.byref(nw)),
            "make_samples")
    t_samples = time.perf_counter() - t0
    ns = int(nw.value)
    tp = abi.MLPTrainParams(20, 65536, 1e-3, 0.9, 0.999, 1e-8, 20261202)
    loss = C.c_float()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.check(L.cleo_mlp_train(d.h, xd.data_ptr(), yd.data_ptr(), ns, C.byref(tp), C.byref(loss)), "mlp_train")
    t_train = time.perf_counter() - t0
    del xd, yd
    torch.cuda.empty_cache()
    sdm_l = make_params(abi)
    sdm_l.cond.warm_start = 1
    sdm_l.cond.warm_restart_iters = 10
    d.set_sdm(sdm_l)
    res = timed_steps(d, t, args.steps)
    res.update(samples=ns, train_s=t_train, samples_s=t_samples, final_loss=float(loss.value),
               warm_restarts=res["stats"]["warm_restarts"])
    return res


# ------------------------------------------------------------------------------------------------
# the other workloads: collisions alone, condensation alone, condensation with the learned start


def _device_domain(torch, b, arrs, st, halves, ndims, solute, seed=20261202):
    from cleo_amd import abi
    from cleo_amd import populations as P
    L = b.lib
    # the library copies on its own stream: the generator's kernels on torch's stream finish first
    torch.cuda.synchronize()
    n = int(arrs["xi"].numel())
    soa = abi.SupersSoA()
    soa.sdgbxindex = _ptr(arrs["sdgbxindex"], C.c_uint32)
    soa.coord3, soa.coord1, soa.coord2 = (_ptr(arrs[k], C.c_double) for k in ("coord3", "coord1", "coord2"))
    soa.xi = _ptr(arrs["xi"], C.c_uint64)
    soa.radius, soa.msol = _ptr(arrs["radius"], C.c_double), _ptr(arrs["msol"], C.c_double)
    soa.sdid = _ptr(arrs["sdid"], C.c_uint64)
    d = abi.Domain(b, n, ndims, *halves)
    d.check(L.cleo_upload_supers_device(d.h, C.byref(soa), n), "upload_supers_device")
    d.ntot = n
    d.set_state(st)
    d.set_solute(*P.solute_dimless(solute))
    d.rng_counter(seed)
    return d, n


def _graph_steps(d, L, torch, t0, k, span=100):
    """k steps replayed as a hipGraph, HIP events on the library's stream: (ms per step, stats)."""
    d.reset_stats()
    d.check(L.cleo_synchronize(d.h), "sync")
    L.cleo_event_record(d.h, 8)
    d.check(L.cleo_run_steps_graph(d.h, t0, span, k), "graph steps")
    L.cleo_event_record(d.h, 9)
    d.check(L.cleo_synchronize(d.h), "sync")
    ms = C.c_float()
    d.check(L.cleo_event_elapsed(d.h, 8, 9, C.byref(ms)), "event_elapsed")
    return float(ms.value) / k, d.stats()


def _host_arrays(arrs):
    """device-generator output (torch, on the CPU) as the numpy SoA the oracle uploads"""
    out = {}
    for k, v in arrs.items():
        a = v.numpy()
        out[k] = a.view(np.uint32) if k == "sdgbxindex" else (a.view(np.uint64) if k in ("xi", "sdid") else a)
    return out


def _cpu_leg(torch, gen, ndims, solute, procs, steps, warm=1, **kw):
    """OpenMP oracle on a bounded sample of a workload: superdroplet-steps/s and the sample text."""
    from cleo_amd import abi
    from cleo_amd import populations as P
    from oracle import pyoracle
    b = pyoracle.load("omp")
    arrs, st, halves = gen(torch, "cpu", ndims=ndims, **kw)
    sup = _host_arrays(arrs)
    d = abi.Domain(b, len(sup["xi"]), ndims, *halves)
    d.upload(sup)
    d.set_state(st)
    d.set_solute(*P.solute_dimless(solute))
    d.rng_counter(20261202)
    assert d.set_sdm_chain(procs) == 0
    for s_ in range(warm):
        d.run_step(100 * s_, 100 * (s_ + 1))
    d.reset_stats()
    t0 = time.perf_counter()
    for s_ in range(warm, warm + steps):
        rc = d.run_step(100 * s_, 100 * (s_ + 1))
        if rc != 0:
            raise RuntimeError(f"oracle run_step rc={rc}")
    dt = time.perf_counter() - t0
    n = len(sup["xi"])
    st_ = d.stats()
    d.close()
    return dict(value=n * steps / dt, unit="superdroplet-steps/s", cores=int(b.lib.orc_num_threads()), kind="port",
                sample=f"{ndims[0]}x{ndims[1]}x{ndims[2]} gridboxes ({n} drops), {steps} steps after {warm} "
                       f"warm-up, {dt:.1f} s, OpenMP oracle"), st_


def run_configs(args):
    """The other workloads, each on its own: a device error in one is reported under its name."""
    import torch

    from cleo_amd import lib

    b = lib.load()
    out = {}
    for leg in (_leg_golovin_box, _leg_coal_breakup, _leg_condensation):
        try:
            out.update(leg(args, b, torch))
        except RuntimeError as e:
            out[leg.__name__[5:]] = {"error": str(e)}
        torch.cuda.empty_cache()
    return out


def _leg_golovin_box(args, b, torch):
    from cleo_amd import abi
    from cleo_amd import populations as P
    from oracle import pyoracle
    L = b.lib
    out = {}
    # the Golovin box (the reference's CPU-runnable case): 1 gridbox of 1e6 m^3, 8192 drops, Golovin
    # kernel, 3600 steps of 1 s replayed as hipGraphs, against the oracle on the same box
    gbox = P.golovin_box()
    kg = abi.CollParams(100, abi.PROB_GOLOVIN, 0.0, abi.ENACT_COAL, 5.0, 0.85)
    res = {}
    for name, backend in (("gpu", b), ("cpu", pyoracle.load("omp"))):
        dd = abi.Domain(backend, gbox.n, gbox.ndims, *gbox.halves)
        dd.upload(gbox.supers)
        dd.set_state(gbox.state)
        dd.rng_counter(20261202)
        assert dd.set_sdm_chain([abi.coll_process(kg)]) == 0
        t0 = time.perf_counter()
        if name == "gpu":
            d_ms = _graph_steps(dd, L, torch, 0, 3600)[0]
            res[name] = d_ms
        else:
            for s_ in range(3600):
                dd.run_step(100 * s_, 100 * (s_ + 1))
            res[name] = (time.perf_counter() - t0) * 1e3 / 3600
        xi_end = float(np.sum(dd.download()["xi"], dtype=np.float64))
        res[name + "_xi_ratio"] = xi_end / float(np.sum(gbox.supers["xi"], dtype=np.float64))
        dd.close()
    out["golovin_box"] = dict(
        workload="1 gridbox of 1e6 m^3, 8192 drops, exponential-in-volume radii (30.531 um), Golovin kernel, "
                 "3600 steps of 1 s, counter RNG, hipGraph", ndrops=gbox.n, ms_per_step=res["gpu"],
        value=gbox.n * 1e3 / res["gpu"], unit="superdroplet-steps/s",
        real_droplets_left=res["gpu_xi_ratio"],
        cpu_baseline=dict(value=gbox.n * 1e3 / res["cpu"], unit="superdroplet-steps/s", cores=1, kind="port",
                          sample="the same 3600 steps (one gridbox runs on one thread), OpenMP oracle",
                          real_droplets_left=res["cpu_xi_ratio"]),
        note="one gridbox of 8192 drops is far too small to fill the GPU: the step is launch-bound")
    return out


def _leg_coal_breakup(args, b, torch):
    from cleo_amd import abi
    from cleo_amd import populations as P
    L = b.lib
    dev = torch.device("cuda:0")
    k = args.config_steps
    out = {}
    # collision-coalescence + breakup alone: 64^3 gridboxes of 50 m, Poisson(128) drops, Long kernel,
    # DoCoalBuRe<ConstNFrags(5), ConstCoalBuReFlag(0.85)>, dt = 1 s
    kp = abi.CollParams(100, abi.PROB_LONG, 1.0, abi.ENACT_COALBURE, 5.0, 0.85)
    ndims = (64, 64, 64)
    arrs, st, halves = P.coal_breakup_device(torch, dev, ndims=ndims)
    d, n = _device_domain(torch, b, arrs, st, halves, ndims, P.NACL)
    del arrs
    torch.cuda.empty_cache()
    assert d.set_sdm_chain([abi.coll_process(kp)]) == 0
    d.check(L.cleo_run_steps_graph(d.h, 0, 100, 2), "warmup")
    tiles = abi.Tiles()
    d.check(L.cleo_autotune(d.h, C.byref(tiles)), "autotune")
    ms, stats = _graph_steps(d, L, torch, 200, k)
    d.close()
    cpu, _ = _cpu_leg(torch, P.coal_breakup_device, (64, 64, 16), P.NACL, [abi.coll_process(kp)], 10)
    val = n * 1e3 / ms
    out["coal_breakup_64cubed"] = dict(
        workload="64x64x64 gridboxes of 50 m, Poisson(128) drops, Long kernel, CoalBuRe(nfrags 5, coaleff 0.85), "
                 "counter RNG, hipGraph", ndrops=n, ms_per_step=ms, value=val, unit="superdroplet-steps/s",
        coll_events_per_step=stats["coll_events"] / k,
        hbm_model_bytes_per_drop=48, hbm_model_GBs=48 * n / (ms / 1e3) / 1e9, cpu_baseline=cpu,
        gpu_over_cpu=val / cpu["value"], tiles={f: int(getattr(tiles, f)) for f, _ in tiles._fields_})
    return out


def _leg_condensation(args, b, torch):
    from cleo_amd import abi
    from cleo_amd import populations as P
    L = b.lib
    dev = torch.device("cuda:0")
    k = args.config_steps
    ndims = (64, 64, 64)
    out = {}
    # condensation alone: 64^3 gridboxes, Poisson(122), ammonium sulphate at Koehler equilibrium, per
    # gridbox T, p, S; FP64 Newton rtol 1e-6, atol 1e-8, maxniters 50, thermodynamic feedback
    cp = abi.CondParams(100, 1, 50, 1e-6, 1e-8, 0.001, 0, 10)
    arrs, st, halves = P.condensation_device(torch, dev, ndims=ndims)
    d, n = _device_domain(torch, b, arrs, st, halves, ndims, P.AMMONIUM_SULPHATE)
    dl, _ = _device_domain(torch, b, arrs, st, halves, ndims, P.AMMONIUM_SULPHATE)  # twin for the learned start
    del arrs
    torch.cuda.empty_cache()
    for x in (d, dl):
        assert x.set_sdm_chain([abi.cond_process(cp)]) == 0
        x.check(L.cleo_run_steps_graph(x.h, 0, 100, 2), "warmup")
    tiles = abi.Tiles()
    d.check(L.cleo_autotune(d.h, C.byref(tiles)), "autotune")
    L.cleo_set_tiles(dl.h, C.byref(tiles))
    ms, stats = _graph_steps(d, L, torch, 200, k)
    iters = stats["newton_iters"] / max(stats["cond_drops"], 1)
    cpu, cst = _cpu_leg(torch, P.condensation_device, (64, 64, 8), P.AMMONIUM_SULPHATE, [abi.cond_process(cp)], 3)
    # FP64 operation count of one drop's solve: the setup (saturation ratio, diffusion and ventilation
    # factors, Koehler factors, critical radius and timestep) and each Newton update (residual and
    # derivative), with a pow / exp / sqrt counted as 20 operations
    flop = 400 + 220 * iters
    val = n * 1e3 / ms
    out["condensation_64cubed"] = dict(
        workload="64x64x64 gridboxes of 50 m, Poisson(122) drops, ammonium sulphate at Koehler equilibrium, "
                 "T 270-300 K, p 7e4-1e5 Pa, S-1 in [-0.05, 0.01], FP64 Newton (rtol 1e-6, atol 1e-8), hipGraph",
        ndrops=n, ms_per_step=ms, value=val, unit="superdroplet-steps/s", newton_iters_per_drop=iters,
        fp64_flop_per_drop_est=flop, fp64_TFLOPs=flop * n / (ms / 1e3) / 1e12, fp64_peak_TFLOPs=FP64_PEAK_TFLOPS,
        hbm_model_bytes_per_drop=32, hbm_model_GBs=32 * n / (ms / 1e3) / 1e9, cpu_baseline=cpu,
        cpu_newton_iters_per_drop=cst["newton_iters"] / max(cst["cond_drops"], 1), gpu_over_cpu=val / cpu["value"],
        tiles={f: int(getattr(tiles, f)) for f, _ in tiles._fields_})

    # the same population with the learned warm start: MLP trained on 10M samples of the plain solver
    # (Adam lr 1e-3, batch 65536, 20 epochs), Newton unchanged, restart from the plain guess after 10
    cap = 10_000_000
    xd = torch.zeros((cap, 8), dtype=torch.float32, device="cuda")
    yd = torch.zeros(cap, dtype=torch.float32, device="cuda")
    nw = C.c_uint64()
    dl.check(L.cleo_mlp_make_samples(dl.h, C.byref(cp), xd.data_ptr(), yd.data_ptr(), cap, C.byref(nw)), "samples")
    tp = abi.MLPTrainParams(20, 65536, 1e-3, 0.9, 0.999, 1e-8, 20261202)
    loss = C.c_float()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dl.check(L.cleo_mlp_train(dl.h, xd.data_ptr(), yd.data_ptr(), int(nw.value), C.byref(tp), C.byref(loss)), "train")
    t_train = time.perf_counter() - t0
    del xd, yd
    torch.cuda.empty_cache()
    cpl = abi.CondParams(100, 1, 50, 1e-6, 1e-8, 0.001, 1, 10)
    assert dl.set_sdm_chain([abi.cond_process(cpl)]) == 0
    msl, statsl = _graph_steps(dl, L, torch, 200, k)
    itl = statsl["newton_iters"] / max(statsl["cond_drops"], 1)
    rp, rl = d.download()["radius"], dl.download()["radius"]
    rel = np.abs(rl - rp) / rp
    d.close()
    dl.close()
    out["condensation_learned_64cubed"] = dict(
        workload="the condensation population and solver with the MLP warm start (6-64-64-1 SiLU, bf16 weights, "
                 "fp32 accumulation)", ndrops=n, ms_per_step=msl, value=n * 1e3 / msl, unit="superdroplet-steps/s",
        newton_iters_per_drop=itl, plain_newton_iters_per_drop=iters, iters_cut_vs_plain=1.0 - itl / iters,
        warm_restarts_per_step=statsl["warm_restarts"] / k, plain_ms_per_step=ms,
        radius_rel_diff_vs_plain={"mean": float(rel.mean()), "max": float(rel.max()),
                                  "after_steps": 2 + k},
        train={"samples": int(nw.value), "seconds": t_train, "final_mse_normalised": float(loss.value)})
    return out


def run_cpu(args):
    from cleo_amd import abi
    from cleo_amd import populations as P
    from oracle import pyoracle

    host = host_cpu()
    b = pyoracle.load("omp")
    ndims = (128, 128, args.cpu_ny)
    pop = P.coupled(ndims=ndims)
    d = abi.Domain(b, pop.n, pop.ndims, *pop.halves)
    d.upload(pop.supers)
    d.set_state(pop.state)
    d.set_solute(*P.solute_dimless(pop.solute))
    sdm = make_params(abi)
    d.set_sdm(sdm)
    d.rng_counter(20261202)
    for s in range(2):  # same warm-up as the GPU leg
        d.run_step(s * 100, (s + 1) * 100)
    t0 = time.perf_counter()
    for s in range(2, 2 + args.cpu_steps):
        rc = d.run_step(s * 100, (s + 1) * 100)
        if rc != 0:
            raise RuntimeError(f"oracle run_step rc={rc}")
    dt = time.perf_counter() - t0
    val = pop.n * args.cpu_steps / dt
    cores = int(b.lib.orc_num_threads())
    d.close()
    # the same code on one thread, one step of a 128 x 128 x 2 slab of the workload
    one = P.coupled(ndims=(128, 128, 2))
    d1 = abi.Domain(b, one.n, one.ndims, *one.halves)
    d1.upload(one.supers)
    d1.set_state(one.state)
    d1.set_solute(*P.solute_dimless(one.solute))
    d1.set_sdm(sdm)
    d1.rng_counter(20261202)
    b.lib.orc_set_num_threads(1)
    d1.run_step(0, 100)
    t1 = time.perf_counter()
    d1.run_step(100, 200)
    dt1 = time.perf_counter() - t1
    b.lib.orc_set_num_threads(cores)
    d1.close()
    return dict(value=val, unit="superdroplet-steps/s", cores=cores, kind="port",
                sample=f"coupled workload on 128x128x{args.cpu_ny} gridboxes ({pop.n} drops, Poisson(64)), "
                       f"{args.cpu_steps} steps after 2 warm-up steps, {dt:.1f} s, OpenMP oracle "
                       f"(gridbox-parallel, 72-byte AoS, atomic counting sort, g++ -O3 -march=x86-64-v3)",
                one_thread=dict(value=one.n / dt1, unit="superdroplet-steps/s",
                                sample=f"one step of 128x128x2 gridboxes ({one.n} drops) after one warm-up step"),
                host=host)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): the library has no multi-GPU "
              "transport (the spatial sharding it would use is described with the design).", file=sys.stderr)
        return 2
    if int(os.environ.get("RANK", "0")) != 0:
        return 0
    g = run_gpu(args)
    p = g["plain"]
    ms_step = p["ms_total"] / args.steps
    value = p["drop_steps"] / (p["ms_total"] / 1e3)
    iters_plain = p["iters_per_drop"]
    # dominant kernel: the stage kernel with the longest live HIP-event time; algorithmic bytes per
    # launch over its mean launch duration
    nd = g["nd"]
    km = g["kernel_ms"]
    dom = max(km, key=km.get)
    bpd, bpg = BYTES[dom]
    if dom == "k_sort_gather_cond" and g["tiles"].motion_flat == 3:
        bpg += 6 * 8  # the drops move in the gather: the face winds of each gridbox
    alg_bytes = bpd * nd + bpg * g["ngbx"]
    achieved = alg_bytes / (km[dom] / 1e3) / 1e9
    traffic = measured_traffic(dom, g["ndims"])
    cpu = None if args.no_cpu else run_cpu(args)
    configs = None if args.no_configs else run_configs(args)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "superdroplet-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"coupled SDM step: {g['ndims'][0]}x{g['ndims'][1]}x{g['ndims'][2]} gridboxes of 25 m, "
                        f"Poisson(64) drops/gridbox, motion+sort+condensation+Long coal/breakup, dt=1 s, hipGraph",
            "ndrops": g["n"],
            "ngbx": g["ngbx"],
            "tiles": {f: int(getattr(g["tiles"], f)) for f, _ in g["tiles"]._fields_},
        },
        "newton_iters_per_drop": iters_plain,
        "wall_s": p["wall_s"],
        "gen_s": g["gen_s"],
        "stage_ms": g["stage_ms"],
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": "profiles/traffic.json: rocprofv3 FETCH_SIZE and WRITE_SIZE of this kernel "
                                       "on this workload, scaled by a copy kernel of known size (not measured in "
                                       "this run)" if traffic is not None else None,
                     "algorithmic_bytes": alg_bytes, "kernel_ms": km[dom]},
        "kernel_ms": km,
        # MassMomentsObserver's reduction: read xi, r, msol | write mom0, mom1, mom2 per gridbox
        "observer": {"kernel": "k_mass_moments", "ms": g["observer_ms"],
                     "GB/s": (24 * nd + 16 * g["ngbx"]) / (g["observer_ms"] / 1e3) / 1e9},
        # FromFileComms::receive_dynamics from device arrays: read 4 thermodynamic values and 6
        # face winds, write 10 State values per gridbox
        "coupling": {"kernel": "k_receive_dynamics", "ms": g["coupling_ms"],
                     "GB/s": 20 * 8 * g["ngbx"] / (g["coupling_ms"] / 1e3) / 1e9},
        "cpu_baseline": cpu,
    }
    if g["learned"]:
        lr = g["learned"]
        out["learned"] = {
            "value": lr["drop_steps"] / (lr["ms_total"] / 1e3),
            "ms_per_step": lr["ms_total"] / args.steps,
            "newton_iters_per_drop": lr["iters_per_drop"],
            "iters_cut_vs_plain": 1.0 - lr["iters_per_drop"] / iters_plain,
            "warm_restarts": lr["warm_restarts"],
            "train": {"samples": lr["samples"], "epochs": 20, "batch": 65536, "lr": 1e-3, "seconds": lr["train_s"],
                      "final_mse_normalised": lr["final_loss"]},
        }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    if configs:
        out["configs"] = configs
    if g["err"][0] != 0:
        out["device_error"] = g["err"]
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
