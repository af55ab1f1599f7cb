"""Benchmark of the pair-force hot path on one MI355X.

Workload (plj_1m): PerturbedLennardJones, N = 1,048,576 particles on a
jittered FCC 64^3 lattice at density 0.85 (L = 107.249), epsilon = sigma = 1,
attraction_scale_factor = 0.5, r_cut = 3.0, r_buff = 0.4, mode none, one
particle type.  Particles are in Morton cell order.

A step is one force evaluation of the whole system in the mixed form
(positions -> 32-bit fixed point, then the pair kernel: FP32 pair arithmetic,
FP64 accumulation of force, energy and virial), with the neighbor list built
beforehand (it is rebuilt every ~10-20 MD steps, not every evaluation).  The
line also carries the plain FP64 form, the timesteps/s of full Langevin MD
steps (kT = 1.2, gamma = 1, dt = 0.005, rebuild checks and rebuilds included),
the mixed form's force error against the FP64 form, the roofline of the pair
kernel and the CPU baseline (the FP64 oracle on one thread, and threaded over
particles with a full list on every thread the job may use).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-tps]
"""

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "pair interactions/s & TPS at 1M particles, mixed vs FP64 (force rel err ≤1e-5)"
HBM_PEAK_GBS = 8000.0
# bytes per particle besides the list: pos 16 (fixed) + n_neigh 4 + head 4 + force 32 + virial 48
MIXED_BYTES_PER_PARTICLE = 104
MIXED_NO_VIRIAL_BYTES_PER_PARTICLE = 56
FP64_BYTES_PER_PARTICLE = 124
# non-packed VALU issue: 256 CUs x 4 SIMDs x 16 lanes x 2.4 GHz
VALU_PEAK_LANE_OPS = 256 * 4 * 16 * 2.4e9
# the headline kernel as rocprofv3 names it: PerturbedLJ, mixed, virial, one
# type, no xplor, cubic box, tiled list, FP32 partial sums
HEADLINE_KERNEL = ("void azp::k_pair<azp::EvalPerturbedLJ, true, true, true, false, true, 1, false>"
                   "(azp::PairKernelArgs)")
# the sources each counter summary's kernels are compiled from: a summary
# measured on other sources is not reported
KERNEL_SOURCES = {
    "pmc_plj_1m.json": ("azplugins_amd/csrc/pair_kernel.h", "azplugins_amd/csrc/evaluators.h",
                        "azplugins_amd/csrc/device_common.h", "azplugins_amd/csrc/tiles.h",
                        "azplugins_amd/csrc/k_pair_tiled.hip", "include/azp.h"),
    "pmc_dpd_2m.json": ("azplugins_amd/csrc/k_dpd.hip", "azplugins_amd/csrc/device_common.h",
                        "azplugins_amd/csrc/portable_math.h", "azplugins_amd/csrc/tiles.h", "include/azp.h"),
    "pmc_tpm_batch.json": ("azplugins_amd/csrc/k_batch.hip", "azplugins_amd/csrc/tpm.h",
                           "azplugins_amd/csrc/device_common.h", "include/azp.h"),
    "pmc_yukawa_262k.json": ("azplugins_amd/csrc/k_adjoint.hip", "azplugins_amd/csrc/pair_kernel.h",
                             "azplugins_amd/csrc/evaluators.h", "azplugins_amd/csrc/device_common.h",
                             "include/azp.h"),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    # untimed steps before the timed ones: the GPU's clocks ramp up over the
    # first tens of milliseconds of load (at 1M the step takes 200.7 us after
    # 5 warm-up steps, 192.5 after 50 and 187.5 after 200, the sustained rate
    # an MD run sees)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tps", action="store_true")
    ap.add_argument("--tile", type=int, default=1, choices=(0, 1, 256, 512),
                    help="tiled list for the mixed form: 0 off, 1 auto (512 else 256), or a fixed size")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the DPD, adjoint, batched-replica and Hertz workloads")
    ap.add_argument("--extras", default="dpd_2m,yukawa_262k_adjoint,tpm_batch_512,hertz_4096,walls_1m,gsd_frame_1m",
                    help="comma-separated workloads to run besides the headline")
    ap.add_argument("--no-autotune", action="store_true",
                    help="batched replicas: keep the default launch shape (short profiling passes)")
    ap.add_argument("--no-headline", action="store_true",
                    help="run only the workloads in --extras (profiling passes of those kernels)")
    ap.add_argument("--tps-steps", type=int, default=200)
    ap.add_argument("--cpu-cells", type=int, default=64, help="FCC cells per side of the CPU sample")
    ap.add_argument("--sweep-blocks", action="store_true", help="time the pair kernels per block size")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy (HBM counter calibration) and exit")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                          "(multi-GPU sharding is designed but not built)"}))
        sys.exit(2)

    from azplugins_amd import _azp, _azplugins, configs

    torch = _azp.torch
    dev = torch.device("cuda:0")
    if args.calibrate_copy:
        # a plain copy kernel of known size: reads and writes 2^30 bytes each
        src = torch.ones(2**27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("bench: copy calibration, 1073741824 bytes read + written per launch")
        return
    if args.no_headline:
        print(json.dumps({"workloads": run_extras(args)}))
        return
    c = configs.plj_1m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    log(f"bench: plj_1m N={N} L={L[0]:.4f}")
    pos = torch.tensor(pos4, device=dev)
    # no d_pos_fixed: the tiled kernel converts positions itself, and over the
    # plain list the launcher packs them into its scratch
    parts = _azp.Particles(N, _azp.Box(*L), pos.data_ptr(), None, None, None, None)
    box = _azp.Box(*L)
    stream = _azp.stream_ptr()

    pp = _azplugins.PotentialPairPerturbedLennardJonesGPU(1)
    pp.setParams(0, 0, [1.0, 1.0, 0.5])
    pp.setRCut(0, 0, 3.0)
    nl = _azplugins.NeighborListGPU(0.4)
    nl.set_r_cut(1, pp.getRCutMatrix())
    t0 = time.perf_counter()
    nl.update(parts, force_rebuild=True)
    torch.cuda.synchronize()
    t_nlist = time.perf_counter() - t0
    # the mixed form reads the tiled list (16-bit entries, tile neighbors in LDS)
    nl.set_tiled(args.tile)
    t0 = time.perf_counter()
    nl.update(parts, force_rebuild=True)
    torch.cuda.synchronize()
    t_nlist_tiled = time.perf_counter() - t0
    tiles = nl.get_tiles()
    tile_order = nl.get_tile_order()
    log(f"bench: tile size {tiles[3]} (0 = plain list), LDS slots {tiles[4]}, "
        f"{'per-tile walk' if nl.walk_kind() else 'plain walk'}")
    d_nn, d_list, pitch, mx = nl.get()
    nn_host, _ = nl.download(N)
    nlist_entries = int(nn_host.sum())

    # device params / cutoffs for the launcher ABI
    pod = (ctypes.c_char * 32)()
    _azp.call("azp_param_pack", 2, _azp.dbl_array([1.0, 1.0, 0.5]), pod)
    d_params = torch.tensor(np.frombuffer(bytes(pod), dtype=np.uint8).copy(), device=dev)
    rcsq = torch.tensor([9.0], dtype=torch.float64, device=dev)
    vpitch = (N + 63) // 64 * 64
    force = torch.zeros((N, 4), dtype=torch.float64, device=dev)
    virial_t = torch.zeros((6, vpitch), dtype=torch.float64, device=dev)

    def make_args(precision, virial=1, tiled=False):
        a = _azp.PairArgs()
        a.particles = parts
        a.out.d_force = force.data_ptr()
        a.out.d_virial = virial_t.data_ptr()
        a.out.virial_pitch = vpitch
        a.out.d_torque = None
        a.d_n_neigh = d_nn
        a.d_nlist = d_list
        a.nlist_pitch = pitch
        a.max_neighbors = mx
        a.d_rcutsq = rcsq.data_ptr()
        a.d_ronsq = None
        a.ntypes = 1
        a.shift_mode = 0
        a.compute_virial = virial
        a.precision = precision
        a.block_size = 0
        a.stream = stream
        if tiled and tiles[3]:
            a.d_tile_count, a.d_tile_jset, a.d_tile_list, a.tile_size, a.tile_slots = tiles
            a.d_tile_order, a.d_tile_nn = tile_order
        return a

    # the forms measured: the headline (mixed, tiled list, with the per-particle
    # virial), the same without the virial (what a Langevin run without
    # pressure logging computes), FP32 pair arithmetic with every entry added
    # to FP64 accumulators, and the FP64 form over the plain and the tiled list
    forms = {
        "mixed": make_args(1, tiled=True),
        "mixed_no_virial": make_args(1, virial=0, tiled=True),
        "mixed_acc64": make_args(2, tiled=True),
        "fp64": make_args(0),
        "fp64_tiled": make_args(0, tiled=True),
    }

    def launch(name):
        _azp.call("azp_gpu_compute_pair_forces", ctypes.byref(forms[name]), 2, _azp.dptr(d_params))

    def time_steps(fn, steps, warmup):
        for _ in range(warmup):
            fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    def time_kernel(name, reps):
        """Average duration of the pair kernel alone: HIP events on its stream
        (the launches go to torch's current stream) around `reps` back-to-back
        launches, the sustained rate of a run of steps."""
        for _ in range(3):
            launch(name)
        start = torch.cuda.Event(enable_timing=True)
        end = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(reps):
            launch(name)
        end.record()
        torch.cuda.synchronize()
        return start.elapsed_time(end) / 1e3 / reps

    def results(name):
        launch(name)
        torch.cuda.synchronize()
        return force.clone(), virial_t.clone()

    # errors of each form against the FP64 form over the plain list
    f64, v64 = results("fp64")
    errors = {}
    for name in ("mixed", "mixed_acc64", "fp64_tiled"):
        f, v = results(name)
        errors[name] = {
            "max_rel_force": (f[:, :3] - f64[:, :3]).abs().max().item() / f64[:, :3].abs().max().item(),
            "l2_rel_force": ((f[:, :3] - f64[:, :3]).norm() / f64[:, :3].norm()).item(),
            "max_rel_energy": (f[:, 3] - f64[:, 3]).abs().max().item() / f64[:, 3].abs().max().item(),
            "max_rel_virial": (v - v64).abs().max().item() / v64.abs().max().item(),
        }
        log(f"bench: {name} vs fp64: force rel err {errors[name]['max_rel_force']:.2e} "
            f"(l2 {errors[name]['l2_rel_force']:.2e}) energy {errors[name]['max_rel_energy']:.2e} "
            f"virial {errors[name]['max_rel_virial']:.2e}")

    # count of unordered pairs inside r_cut (the unit of work)
    pairs_in_cut = count_pairs_in_cut(torch, pos, nn_host, nl, N, L, 9.0)
    log(f"bench: nlist entries {nlist_entries} ({nlist_entries / N:.1f}/particle), "
        f"pairs in cut {pairs_in_cut} ({2 * pairs_in_cut / N:.1f}/particle), nlist build {t_nlist * 1e3:.1f} ms")

    if args.sweep_blocks:
        for name in ("mixed", "fp64"):
            if name == "mixed" and tiles[3]:
                continue  # the tiled kernel's workgroup is its tile
            for b in (64, 128, 256, 512):
                forms[name].block_size = b
                time_kernel(name, 3)
                log(f"bench: sweep {name} block={b} kernel {time_kernel(name, 10) * 1e6:.1f} us")
            forms[name].block_size = 0

    reps = max(10, args.steps)
    t_step = {}
    for name in ("mixed", "fp64", "fp64_tiled"):
        t_step[name] = time_steps(lambda: launch(name), args.steps, args.warmup)
    k_us = {name: time_kernel(name, reps) for name in forms}
    log("bench: step " + ", ".join(f"{k} {v * 1e6:.1f} us" for k, v in t_step.items()) + "; kernel "
        + ", ".join(f"{k} {v * 1e6:.1f} us" for k, v in k_us.items()))
    t_mixed = t_step["mixed"]

    # algorithmic bytes per launch: per-particle streams + 4 bytes per list entry
    bytes_mixed = N * MIXED_BYTES_PER_PARTICLE + 4 * nlist_entries
    bytes_mixed_nv = N * MIXED_NO_VIRIAL_BYTES_PER_PARTICLE + 4 * nlist_entries
    bytes_fp64 = N * FP64_BYTES_PER_PARTICLE + 4 * nlist_entries
    achieved = bytes_mixed / k_us["mixed"] / 1e9
    kernel_name = HEADLINE_KERNEL if tiles[3] else "k_pair<EvalPerturbedLJ, mixed, virial, one type, column list>"
    counters = load_counters("pmc_plj_1m.json", HEADLINE_KERNEL) if tiles[3] else None
    roofline = {
        "bound": "hbm",
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": counters["traffic_bytes"] if counters else None,
        "kernel": kernel_name,
        "kernel_us": round(k_us["mixed"] * 1e6, 2),
        "algorithmic_bytes": bytes_mixed,
        "counters": counters["provenance"] if counters else "none measured on these kernel sources",
    }
    # the kernel is bound by VALU issue: its wavefront VALU instructions per
    # launch (rocprofv3 counters) over the live kernel time, against the
    # non-packed VALU issue peak
    if counters and counters.get("insts_valu"):
        lane_ops = counters["insts_valu"] * 64
        roofline["valu"] = {
            "instructions_per_launch": counters["insts_valu"],
            "lane_instructions_per_list_entry": round(lane_ops / nlist_entries, 2),
            "achieved_lane_ops_per_s": lane_ops / k_us["mixed"],
            "peak_lane_ops_per_s": VALU_PEAK_LANE_OPS,
            "frac": round(lane_ops / k_us["mixed"] / VALU_PEAK_LANE_OPS, 4),
        }

    def form_line(name, nbytes):
        return {"kernel_us": round(k_us[name] * 1e6, 2), "value": pairs_in_cut / k_us[name],
                "unit": "pair interactions/s (kernel time)", "algorithmic_bytes": nbytes,
                "achieved_GBs": round(nbytes / k_us[name] / 1e9, 1),
                "frac": round(nbytes / k_us[name] / 1e9 / HBM_PEAK_GBS, 4)}

    variants = {
        "mixed_no_virial": dict(form_line("mixed_no_virial", bytes_mixed_nv),
                                note="forces and energies only, as the Langevin run computes them"),
        "mixed_acc64": dict(form_line("mixed_acc64", bytes_mixed), errors=errors["mixed_acc64"],
                            dtype="f32 pair arithmetic, every entry added to f64 accumulators"),
        "fp64_tiled": dict(form_line("fp64_tiled", bytes_fp64), ms_per_step=t_step["fp64_tiled"] * 1e3,
                           errors=errors["fp64_tiled"]),
    }
    out = {
        "metric": METRIC,
        "value": pairs_in_cut / t_mixed,
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 pair arithmetic; f32 sums of 32 list entries added to f64 accumulators",
        "data": "synthetic",
        "config": {
            "workload": "plj_1m",
            "potential": "PerturbedLennardJones",
            "N": N,
            "density": 0.85,
            "r_cut": 3.0,
            "r_buff": 0.4,
            "pairs_in_cut": pairs_in_cut,
            "nlist_entries": nlist_entries,
            "precision": "mixed",
            "virial": True,
            "list": "tiled" if tiles[3] else "column",
        },
        "roofline": roofline,
        "fp64": {
            "value": pairs_in_cut / t_step["fp64"],
            "ms_per_step": t_step["fp64"] * 1e3,
            "kernel_us": round(k_us["fp64"] * 1e6, 2),
            "achieved_GBs": round(bytes_fp64 / k_us["fp64"] / 1e9, 1),
            "list": "column (plain)",
        },
        "mixed_over_fp64": t_step["fp64"] / t_mixed,
        "mixed_over_fp64_tiled": t_step["fp64_tiled"] / t_mixed,
        "max_rel_force_err_vs_fp64": errors["mixed"]["max_rel_force"],
        "l2_rel_force_err_vs_fp64": errors["mixed"]["l2_rel_force"],
        "errors_vs_fp64": errors["mixed"],
        "variants": variants,
        "nlist_build_ms": t_nlist * 1e3,
        "nlist_build_with_tiles_ms": t_nlist_tiled * 1e3,
        "tile_size": tiles[3],
        "tile_slots": tiles[4],
    }
    if not args.no_tps:
        out["tps"] = run_tps(c, args.tps_steps)
        out["tps_fp64"] = run_tps(c, args.tps_steps, "fp64")
    if not args.no_extras:
        out["workloads"] = run_extras(args)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_cells)
    print(json.dumps(out))


def run_extras(args):
    """The other workloads (DPD, adjoint, batched replicas, Hertz NVE, walls,
    GSD frame) selected by --extras."""
    extras = {}
    wanted = set(args.extras.split(","))
    for name, fn in (("dpd_2m", bench_dpd), ("yukawa_262k_adjoint", bench_adjoint),
                     ("tpm_batch_512", bench_batch), ("hertz_4096", bench_hertz),
                     ("walls_1m", bench_walls), ("gsd_frame_1m", bench_gsd)):
        if name not in wanted:
            continue
        try:
            extras[name] = fn(args)
        except Exception as exc:  # a failing extra must not hide the headline line
            log(f"bench: {name} failed: {exc!r}")
            extras[name] = {"error": repr(exc)}
    return extras


def count_pairs_in_cut(torch, pos, nn_host, nl, N, L, rcutsq):
    """Unordered pairs with r < r_cut, counted on the GPU from the full list."""
    _, lst = nl.download(N)
    return count_list_pairs(torch, pos, nn_host, lst, N, L, rcutsq)


def count_list_pairs(torch, pos, nn_host, lst, N, L, rcutsq, typeid=None, ntypes=1):
    """Unordered pairs of a full list with r^2 < rcutsq[ti, tj] (L: (3,) or
    per-particle (N, 3)).  The list goes to the device in one copy (it views
    pinned memory after a download) and is counted there in row chunks."""
    dev = pos.device
    total = 0
    chunk = 16
    nn_t = torch.from_numpy(np.ascontiguousarray(nn_host).view(np.int32)).to(dev).long()
    Lt = torch.tensor(np.asarray(L, dtype=np.float64), device=dev)
    rc = torch.tensor(np.atleast_1d(np.asarray(rcutsq, dtype=np.float64)), device=dev)
    lst_d = torch.from_numpy(np.ascontiguousarray(lst).view(np.int32)).to(dev)
    for k0 in range(0, lst_d.shape[0], chunk):
        rows = lst_d[k0:k0 + chunk, :N].long()
        valid = (torch.arange(k0, k0 + rows.shape[0], device=dev)[:, None] < nn_t[None, :])
        jj = rows.clamp(max=N - 1)
        pj = pos[jj, :3]
        d = pos[None, :, :3] - pj
        d = d - Lt * torch.round(d / Lt)
        rsq = (d * d).sum(-1)
        if typeid is None:
            lim = rc[0]
        else:
            lim = rc[(typeid[None, :] * ntypes + typeid[jj]).long()]
        total += int(((rsq < lim) & valid).sum().item())
    del lst_d
    return total // 2


def _events(torch, fn, reps, warmup=2):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    s = torch.cuda.Event(enable_timing=True)
    e = torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        fn()
    e.record()
    torch.cuda.synchronize()
    return s.elapsed_time(e) / 1e3 / reps


def roofline_line(nbytes, seconds, kernel, counters_file, entries, note=None):
    """roofline object of one workload kernel: algorithmic bytes over the
    measured time against the HBM peak, with counters when measured on the
    current sources."""
    r = {"bound": "hbm", "algorithmic_bytes": nbytes, "us": round(seconds * 1e6, 2),
         "achieved": round(nbytes / seconds / 1e9, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(nbytes / seconds / 1e9 / HBM_PEAK_GBS, 4)}
    if note:
        r["note"] = note
    c = load_counters(counters_file, kernel) if counters_file else None
    if c:
        r["traffic"] = c["traffic_bytes"]
        r["counters"] = c["provenance"]
        if c.get("insts_valu"):
            lane_ops = c["insts_valu"] * 64
            r["valu"] = {"instructions_per_launch": c["insts_valu"],
                         "lane_instructions_per_list_entry": round(lane_ops / entries, 2),
                         "frac": round(lane_ops / seconds / VALU_PEAK_LANE_OPS, 4)}
    return r


def bench_dpd(args):
    """DPD general weight, N = 2,000,000 at density 3, A/B 50/50 (FP64, Philox random forces)."""
    from azplugins_amd import _azp, _azplugins, configs

    torch = _azp.torch
    dev = torch.device("cuda:0")
    c = configs.dpd_2m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    rng = np.random.default_rng(7)
    vel = np.zeros((N, 4))
    vel[:, :3] = rng.normal(scale=np.sqrt(c["kT"]), size=(N, 3))
    vel[:, 3] = 1.0
    pos = torch.tensor(pos4, device=dev)
    velt = torch.tensor(vel, device=dev)
    tag = torch.arange(N, dtype=torch.int32, device=dev)
    parts = _azp.Particles(N, _azp.Box(*L), pos.data_ptr(), None, velt.data_ptr(), tag.data_ptr(), None)
    pp = _azplugins.PotentialPairDPDThermoGeneralWeightGPU(2)
    for (a, b), v in c["params"].items():
        pp.setParams(a, b, [v["A"], v["gamma"], v["s"]])
        pp.setRCut(a, b, c["r_cut"])
    pp.setThermostat(c["seed"], c["dt"], c["kT"])
    nl = _azplugins.NeighborListGPU(c["r_buff"])
    nl.set_r_cut(2, pp.getRCutMatrix())
    nl.update(parts, force_rebuild=True)
    vpitch = (N + 63) // 64 * 64
    force = torch.zeros((N, 4), dtype=torch.float64, device=dev)
    virial = torch.zeros((6, vpitch), dtype=torch.float64, device=dev)
    out = _azp.Outputs(force.data_ptr(), virial.data_ptr(), vpitch, None)
    step = [100]

    def fn():
        pp.compute(nl.handle, parts, out, step[0])
        step[0] += 1

    def result():
        pp.compute(nl.handle, parts, out, 99)
        torch.cuda.synchronize()
        return force.clone(), virial[:, :N].clone()

    # the plain list (global gathers), then the tiled list (neighbors in LDS)
    t_plain = _events(torch, fn, max(5, args.steps // 5))
    f_plain, v_plain = result()
    nl.set_tiled(True)
    nl.update(parts, force_rebuild=True)
    tiles = nl.get_tiles()
    t = _events(torch, fn, max(5, args.steps // 5))
    f_tiled, v_tiled = result()
    same = bool(torch.equal(f_plain, f_tiled) and torch.equal(v_plain, v_tiled))
    del f_plain, v_plain, f_tiled, v_tiled
    nn, lst = nl.download(N)
    pairs = count_list_pairs(torch, pos, nn, lst, N, L, [1.0])
    entries = int(nn.sum())
    # pos 32 + n_neigh 4 + head 8 + force 32 + virial 48 + vel 32 + tag 4 per particle, 4 per list entry
    nbytes = N * 160 + 4 * entries
    kernel = "void azp::k_dpd_tiled<true>(azp::DPDKernelArgs)" if tiles[3] else "void azp::k_dpd<true, true>(azp::DPDKernelArgs)"
    res = {"value": pairs / t, "unit": "pair interactions/s", "ms_per_eval": t * 1e3, "N": N,
           "pairs_in_cut": pairs, "nlist_entries": entries, "precision": "fp64",
           "list": "tiled" if tiles[3] else "column", "tile_size": tiles[3], "tile_slots": tiles[4],
           "plain_list_ms_per_eval": t_plain * 1e3, "tiled_equals_plain_bitwise": same,
           "roofline": roofline_line(nbytes, t, kernel, "pmc_dpd_2m.json", entries),
           "note": "one DPD force evaluation (conservative + dissipative + Philox random) per step"}
    log(f"bench: dpd_2m plain list {t_plain * 1e3:.3f} ms/eval; tiled (size {tiles[3]}, slots {tiles[4]}) "
        f"{t * 1e3:.3f} ms/eval, bit-identical {same}")
    log(f"bench: dpd_2m {pairs / t:.3e} pairs/s")
    return res


def bench_adjoint(args):
    """ExpandedYukawa force matching at N = 262,144: forward + adjoint backward."""
    from azplugins_amd import _azp, configs
    from azplugins_amd.adjoint import PairForceField

    torch = _azp.torch
    dev = torch.device("cuda:0")
    c = configs.yukawa_262k()
    pos4, L = c["pos4"], c["L"]
    keys = [(0, 0), (0, 1), (1, 1)]
    theta = np.array([[c["params"][k]["epsilon"], c["params"][k]["kappa"], c["params"][k]["delta"]] for k in keys])
    typeid = torch.tensor(pos4[:, 3], device=dev)
    pos = torch.tensor(pos4[:, :3], device=dev)
    res = {}
    for precision in ("mixed", "fp64"):
        ff = PairForceField("ExpandedYukawa", 2, L, c["r_cut"], precision=precision)
        with torch.no_grad():
            tgt, _ = ff(pos, typeid, torch.tensor(theta * 1.05, device=dev))
        th = torch.tensor(theta, device=dev, requires_grad=True)
        pr = pos.clone().requires_grad_(True)

        def fwd_bwd():
            th.grad = None
            pr.grad = None
            F, E = ff(pr, typeid, th)
            loss = ((F - tgt) ** 2).mean() + 1e-3 * E.sum()
            loss.backward()

        t = _events(torch, fwd_bwd, 10)
        with torch.no_grad():
            t_f = _events(torch, lambda: ff(pos, typeid, th), 10)
        N = pos4.shape[0]
        entries = int(ff._nl.download(N)[0].sum())
        # forward: the pair kernel's model (104 B per particle mixed, 124 FP64,
        # 4 per list entry); backward: pos 32 + dL/dF,dL/dE 32 + dL/dr 32 +
        # n_neigh 4 per particle and 4 per list entry
        fwd_bytes = N * (104 if precision == "mixed" else 124) + 4 * entries
        bwd_bytes = N * 100 + 4 * entries
        res[precision] = {"value": 1.0 / t, "unit": "loss+gradient evaluations/s", "ms_forward_backward": t * 1e3,
                          "ms_forward": t_f * 1e3, "nlist_entries": entries,
                          "roofline_forward": roofline_line(fwd_bytes, t_f, None, None, entries),
                          "roofline_backward": roofline_line(bwd_bytes, t - t_f, None, None, entries,
                                                             note="backward = forward+backward minus forward: "
                                                                  "the adjoint kernel, its final reduction "
                                                                  "and the torch glue")}
        log(f"bench: yukawa_262k adjoint {precision}: forward {t_f * 1e3:.3f} ms, forward+backward {t * 1e3:.3f} ms")
    res["N"] = pos4.shape[0]
    return res


def bench_batch(args):
    """512 TwoPatchMorse replicas (sizes log-uniform in [500, 20000]): graph-captured Langevin steps."""
    from azplugins_amd import _azp, configs
    from azplugins_amd.batch import from_config

    torch = _azp.torch
    cfg = configs.tpm_batch()
    b = from_config(cfg, precision="mixed")
    b.set_timestep(0)
    b.compute(force_rebuild=True)
    tune = None if args.no_autotune else b.autotune()
    t_force = _events(torch, lambda: b.compute(), 10)
    steps = 20 if args.no_autotune else 100
    b.run(20)
    torch.cuda.synchronize()
    st0 = b.stats()
    t0 = time.perf_counter()
    b.run(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    st1 = b.stats()
    nn, lst = b.nlist()
    Lper = np.repeat(np.stack([r["L"] for r in cfg["replicas"]]), b.sizes.astype(np.int64), axis=0)
    pairs = count_list_pairs(torch, b.pos, nn, lst, b.N, Lper, [cfg["r_cut"] ** 2])
    entries = int(nn.sum())
    # the batched force launch: 104 + orientation 16 + torque 32 per particle, 4 per list entry
    force_roofline = roofline_line(b.N * 152 + 4 * entries, t_force,
                                   "void azp::k_b_forces<true, true, true>(azp::BatchForceArgs)",
                                   "pmc_tpm_batch.json", entries,
                                   note="time of azp_batch_compute: the fixed-point pack and the persistent "
                                        "force launch")
    res = {"value": steps / dt, "unit": "timesteps/s (all replicas)", "replica_steps_per_s": b.R * steps / dt,
           "particle_steps_per_s": b.N * steps / dt, "pairs_in_cut": pairs,
           "pair_interactions_per_s_in_steps": pairs * steps / dt, "force_ms": t_force * 1e3,
           "replicas": b.R, "N_total": b.N, "nlist_entries": entries, "roofline": force_roofline,
           "tuning": tune, "nlist_builds": st1["n_builds"] - st0["n_builds"],
           "replica_rebuilds": st1["replica_builds"] - st0["replica_builds"],
           "graph_launches": st1["graph_launches"] - st0["graph_launches"], "precision": "mixed",
           "method": "Langevin (translation + rotation, I = 1), dt 0.005, per-replica kT"}
    log(f"bench: tpm_batch_512 {steps / dt:.1f} steps/s ({b.N} particles), force {t_force * 1e3:.3f} ms, "
        f"tuning {tune}")
    return res


def bench_hertz(args):
    """Hertz N = 4096: 1000 NVE steps through the drop-in API (GPU)."""
    import azplugins_amd.harness as hoomd
    from azplugins_amd import _azp, configs

    torch = _azp.torch
    c = configs.hertz_4096()
    snap = hoomd.Snapshot()
    N = c["pos4"].shape[0]
    snap.configuration.box = [*c["L"], 0, 0, 0]
    snap.particles.N = N
    snap.particles.position[:] = c["pos4"][:, :3]
    sim = hoomd.Simulation(seed=1)
    sim.create_state_from_snapshot(snap)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=c["kT"])
    integ = hoomd.md.Integrator(dt=c["dt"])
    nl = hoomd.md.nlist.Cell(buffer=c["r_buff"])
    hz = hoomd.azplugins.pair.Hertz(nlist=nl, default_r_cut=c["r_cut"])
    hz.params[("A", "A")] = dict(epsilon=1.0)
    integ.forces = [hz]
    integ.methods = [hoomd.md.methods.ConstantVolume(hoomd.filter.All())]
    sim.operations.integrator = integ
    sim.run(10)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(1000)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    log(f"bench: hertz_4096 GPU NVE {1000 / dt:.1f} steps/s")
    return {"value": 1000 / dt, "unit": "timesteps/s", "steps": 1000, "N": N, "precision": "fp64",
            "method": "NVE velocity Verlet dt 0.005"}


def bench_walls(args):
    """LJ 9-3 walls on the 1M-particle PerturbedLJ configuration: two planes
    and a sphere, forces and energies without the virial (FP64)."""
    from azplugins_amd import _azp, _azplugins, configs

    torch = _azp.torch
    dev = torch.device("cuda:0")
    c = configs.plj_1m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    pos = torch.tensor(pos4, device=dev)
    parts = _azp.Particles(N, _azp.Box(*L), pos.data_ptr(), None, None, None, None)
    w = _azplugins.WallsPotentialLJ93GPU(1)
    w.setParams(0, [1.0, 1.0, 3.0, 0.0])
    h = 0.45 * float(L[2])
    w.setGeometry(spheres=[[0.0, 0.0, 0.0, 0.45 * float(L[0]), 1.0, 0.0]], cylinders=[],
                  planes=[[0.0, 0.0, -h, 0.0, 0.0, 1.0, 1.0], [0.0, 0.0, h, 0.0, 0.0, -1.0, 1.0]])
    force = torch.zeros((N, 4), dtype=torch.float64, device=dev)
    out = _azp.Outputs(force.data_ptr(), None, 0, None)
    t = _events(torch, lambda: w.compute(parts, out, compute_virial=False), max(10, args.steps))
    bytes_alg = N * (32 + 32)
    res = {"value": N / t, "unit": "particle evaluations/s", "us_per_eval": t * 1e6, "N": N, "walls": 3,
           "achieved_GBs": bytes_alg / t / 1e9, "algorithmic_bytes": bytes_alg,
           "note": "algorithmic bytes: position 32 B read + force 32 B written per particle"}
    log(f"bench: walls_1m {t * 1e6:.1f} us/eval, {bytes_alg / t / 1e9:.0f} GB/s")
    return res


def bench_gsd(args):
    """One 1M-particle GSD frame: the state written (device to host and the
    file) and a new state created from it."""
    import tempfile

    import azplugins_amd.harness as hoomd
    from azplugins_amd import _azp, configs

    torch = _azp.torch
    c = configs.plj_1m()
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    snap = hoomd.Snapshot()
    snap.configuration.box = [*L, 0, 0, 0]
    snap.particles.N = N
    snap.particles.position[:] = pos4[:, :3]
    snap.particles.velocity[:] = np.random.default_rng(5).normal(size=(N, 3))
    sim = hoomd.Simulation(seed=1)
    sim.create_state_from_snapshot(snap)
    with tempfile.TemporaryDirectory() as d:
        fn = os.path.join(d, "frame.gsd")
        tw = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            hoomd.write.GSD.write(state=sim.state, filename=fn)
            tw.append(time.perf_counter() - t0)
        size = os.path.getsize(fn)
        tr = []
        for _ in range(3):
            t0 = time.perf_counter()
            s2 = hoomd.Simulation(seed=1)
            s2.create_state_from_gsd(fn)
            torch.cuda.synchronize()
            tr.append(time.perf_counter() - t0)
            del s2
    w, r = float(np.median(tw)), float(np.median(tr))
    log(f"bench: gsd_frame_1m write {w * 1e3:.1f} ms, create state {r * 1e3:.1f} ms ({size / 1e6:.1f} MB)")
    return {"value": size / w / 1e6, "unit": "MB/s written", "write_ms": w * 1e3, "create_state_ms": r * 1e3,
            "file_MB": size / 1e6, "N": N,
            "note": "host path: device-to-host copy of the state, float32 schema fields, file write (median of 3)"}


def run_tps(c, steps, precision="mixed"):
    """Timesteps/s of full Langevin MD steps through the drop-in API."""
    import azplugins_amd.harness as hoomd
    from azplugins_amd import _azp

    torch = _azp.torch
    snap = hoomd.Snapshot()
    N = c["pos4"].shape[0]
    snap.configuration.box = [*c["L"], 0, 0, 0]
    snap.particles.N = N
    snap.particles.position[:] = c["pos4"][:, :3]
    sim = hoomd.Simulation(seed=2)
    sim.create_state_from_snapshot(snap)
    sim.state.thermalize_particle_momenta(hoomd.filter.All(), kT=1.2)
    integ = hoomd.md.Integrator(dt=0.005)
    nl = hoomd.md.nlist.Cell(buffer=0.4)
    plj = hoomd.azplugins.pair.PerturbedLennardJones(nlist=nl, default_r_cut=3.0)
    plj.params[("A", "A")] = dict(epsilon=1.0, sigma=1.0, attraction_scale_factor=0.5)
    plj.precision = precision
    integ.forces = [plj]
    integ.methods = [hoomd.md.methods.Langevin(hoomd.filter.All(), kT=1.2, default_gamma=1.0)]
    sim.operations.integrator = integ
    sim.run(20)
    torch.cuda.synchronize()
    b0 = nl.num_builds
    t0 = time.perf_counter()
    sim.run(steps)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    res = {"value": steps / dt, "unit": "timesteps/s", "steps": steps,
           "nlist_builds": nl.num_builds - b0, "precision": precision, "method": "Langevin kT=1.2 gamma=1 dt=0.005"}
    log(f"bench: TPS {precision} {steps / dt:.1f} ({nl.num_builds - b0} rebuilds in {steps} steps)")
    return res


def cpu_baseline(cells):
    """FP64 oracle (half list, Newton's third law, one thread) on a bounded sample."""
    import oracle
    from azplugins_amd import configs

    if not os.path.exists(oracle._LIB_PATH):
        return None
    c = configs.plj_1m(n=cells)
    pos4, L = c["pos4"], c["L"]
    N = pos4.shape[0]
    pairs = oracle.nlist(pos4, L, 1, oracle.r_listsq_matrix([3.0], 0.4), "cell")
    packed = oracle.pack_matrix(oracle.PERTURBED_LJ, c["params"], 1)
    n_cut = oracle.count_pairs_in_cut(pos4, L, pairs, 1, np.array([9.0]))
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.pair_forces(oracle.PERTURBED_LJ, pos4, L, pairs, 1, packed, np.array([9.0]))
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    log(f"bench: CPU oracle {n_cut / t:.3e} pair interactions/s on N={N} ({t:.2f} s per evaluation)")
    res = {
        "value": n_cut / t,
        "unit": "pair interactions/s",
        "cores": 1,
        "kind": "port",
        "sample": f"PerturbedLJ FCC {cells}^3 (N={N}) at density 0.85, r_cut 3.0, 3 force "
                  f"evaluations (median {t:.2f} s), FP64 half list with third law; host "
                  f"{os.cpu_count()} logical CPUs ({host_cpu_model()})",
    }
    # the same forces over a full list, OpenMP-threaded over particles, on the
    # threads this job may use (OMP_NUM_THREADS, else every logical CPU)
    nthreads = int(os.environ.get("OMP_NUM_THREADS") or os.cpu_count() or 1)
    csr = oracle.full_list(N, pairs)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        oracle.pair_forces_threaded(oracle.PERTURBED_LJ, pos4, L, csr, 1, packed, np.array([9.0]),
                                    nthreads=nthreads)
        times.append(time.perf_counter() - t0)
    tt = float(np.median(times))
    res["threaded"] = {"value": n_cut / tt, "unit": "pair interactions/s", "cores": nthreads,
                       "sample": f"the same configuration, FP64 full list threaded over particles "
                                 f"(OpenMP, {nthreads} threads), 3 evaluations (median {tt:.3f} s)"}
    log(f"bench: CPU oracle {nthreads} threads {n_cut / tt:.3e} pair interactions/s ({tt:.3f} s per evaluation)")
    # Hertz N = 4096, 1000 NVE steps (the CPU-only workload), same oracle, one thread
    h = configs.hertz_4096()
    rng = np.random.default_rng(1)
    vel = np.zeros((h["pos4"].shape[0], 4))
    vel[:, :3] = rng.normal(scale=np.sqrt(h["kT"]), size=(vel.shape[0], 3))
    vel[:, :3] -= vel[:, :3].mean(axis=0)
    vel[:, 3] = 1.0
    t0 = time.perf_counter()
    oracle.run_nve(oracle.HERTZ, h["pos4"], vel, h["L"], 1, oracle.pack_matrix(oracle.HERTZ, h["params"], 1),
                   [h["r_cut"]], h["r_buff"], "none", h["dt"], 1000)
    th = time.perf_counter() - t0
    res["hertz_4096_nve"] = {"value": 1000 / th, "unit": "timesteps/s", "steps": 1000, "cores": 1}
    log(f"bench: CPU oracle Hertz 4096 NVE {1000 / th:.1f} steps/s")
    return res


def host_cpu_model():
    """The host CPU's model name (Linux /proc/cpuinfo), for the baseline's record."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown CPU"


def kernel_source_hash(fname):
    """SHA-256 (first 16 hex digits) of the sources of the kernels that the
    counter summary `fname` covers: the provenance recorded with it."""
    import hashlib

    h = hashlib.sha256()
    for rel in KERNEL_SOURCES[fname]:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(rel.encode() + b"\0" + fh.read())
    return h.hexdigest()[:16]


def load_counters(fname, kernel):
    """Counters of `kernel` from profiles/<fname> (rocprofv3 --pmc summaries),
    only when they were measured on the current kernel sources."""
    p = os.path.join(ROOT, "profiles", fname)
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        d = json.load(fh)
    src = kernel_source_hash(fname)
    if d.get("source_sha256") != src:
        log(f"bench: profiles/{fname} was measured on sources {d.get('source_sha256')}, these are {src}: "
            f"its counters are not reported")
        return None
    k = d.get("kernels", {}).get(kernel)
    if k is None:
        return None
    return {"traffic_bytes": k.get("traffic_bytes"), "insts_valu": k.get("SQ_INSTS_VALU"),
            "provenance": f"profiles/{fname}, kernel sources {src}"}


if __name__ == "__main__":
    main()
