"""Benchmark: FP64 multi-particle tracking on one MI355X.

Workload (one GPU): the 4000-element ring with synchrotron motion -- FODO
cells with thick sextupoles (4 kick-drift slices), thin octupoles, sector bends
with hard-edge fringe and 8 x 1 MV 500 MHz rf cavities, 6 GeV electrons,
1,000,000 Gaussian particles (eps_x 10 nm, eps_y 0.1 nm, sigma_z 1 mm,
sigma_pz 1e-3, matched to the periodic Twiss), elliptical 20 x 10 mm
aperture, FP64.  One step = TURNS_PER_STEP turns of the whole bunch through
the ring in one bmx_track_bunch call; the bunch stays resident in HBM and is
tracked on from step to step, as the long-term tracking turn loop does.

Printed: one JSON line with the metric (particle-turns/s), the roofline of the
tracking kernel (algorithmic FP64 FLOPs from the reference's operation count
per particle-element-turn / kernel time from HIP events on the library's
stream) and the CPU baseline (the oracle's reference-structured OpenMP
tracking on a bounded sample, timed on this host).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bmad-ecosystem_amd"))

import numpy as np  # noqa: E402

TURNS_PER_STEP = 10
N_PARTICLES = 1_000_000
FP64_PEAK_TFLOPS = 78.6


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "bmx_track_bunch_kernel.ring4000_e6GeV_1e6particles.md")


def measured_traffic(particles, turns_per_step):
    """HBM bytes per launch of the bunch kernel on the default workload as
    the committed rocprofv3 counter runs measured them (FETCH_SIZE and
    WRITE_SIZE in separate runs, each scaled by what the counter reads for a
    copy of known size): the `traffic_bytes_per_launch:` line of the
    kernel's profile summary.  None for another workload or without it."""
    if particles != N_PARTICLES or turns_per_step != TURNS_PER_STEP or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.strip().startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1].split()[0])
    return None


def measured_hw_flops_per_pet(particles, turns_per_step):
    """FP64 flops per particle-element-turn the hardware counters counted on
    the default workload (64 lanes x (ADD + MUL + 2 FMA + TRANS) of the
    committed counter run): the `hardware_fp64_flops_per_pet:` line of the
    kernel's profile summary.  None for another workload or without it."""
    if particles != N_PARTICLES or turns_per_step != TURNS_PER_STEP or not os.path.exists(TRAFFIC_PROFILE):
        return None
    with open(TRAFFIC_PROFILE) as f:
        for line in f:
            if line.strip().startswith("hardware_fp64_flops_per_pet:"):
                return float(line.split(":", 1)[1].split()[0])
    return None


def physical_cores(cpus):
    """Physical cores (distinct package, core id) among the logical CPUs
    this process may run on, from the kernel's CPU topology (what lscpu
    reports as cores); the logical count when the topology is unreadable."""
    ids = set()
    for c in cpus:
        base = f"/sys/devices/system/cpu/cpu{c}/topology/"
        try:
            with open(base + "physical_package_id") as f:
                pkg = f.read().strip()
            with open(base + "core_id") as f:
                core = f.read().strip()
        except OSError:
            return len(cpus)
        ids.add((pkg, core))
    return len(ids) or len(cpus)


def cpu_baseline(eles, tw, budget_s=15.0):
    """Oracle (reference loop structure: element-outer, OpenMP over particles,
    bmad/multiparticle/beam_utils.f90:69-74) on a bounded sample of the same
    workload: sample particles x full ring x 1 turn, repeated until ~budget_s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    from bmx import abi, beam

    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count()))
    threads = max(1, min(len(cpus), 64))
    cores = physical_cores(cpus)
    n = 4096
    coords = beam.matched_gaussian(n, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=20261203)
    arr = oracle.abi_array(eles)
    turns = 0
    t0 = time.perf_counter()
    while True:
        oracle.track_bunch(arr, abi.ELECTRON, 1, coords, n_threads=threads)
        turns += 1
        if time.perf_counter() - t0 > budget_s or turns >= 50:
            break
    dt = time.perf_counter() - t0
    return {"value": n * turns / dt, "unit": "particle-turns/s", "cores": cores, "threads": threads, "kind": "port",
            "sample": f"{n} particles x {len(eles)} elements x {turns} turns (1 turn per call), "
                      f"element-outer OpenMP loop, {dt:.1f} s",
            "pet_per_s": n * turns * len(eles) / dt}


# The DA workload: a C2-style error-seeded ring (125 cells = 1,000 elements of
# the C1 cell pattern, 8 rf cavities, quad k1 x (1 + 1e-3 g), sextupole
# k2 x (1 + 5e-3 g), 50 um offsets) under 32 sextupole-family knobs
# U(+-20 %), polar grid 64 angles x 128 amplitudes (0..30 sigma_x, y at the
# same amplitude) x 5 pz in [-2 %, +2 %], 1,024 turns, FP64.
DA_CELLS = 125
DA_TURNS = 1024


def da_spec():
    from bmx import beam, lattice, surrogate
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=DA_CELLS, n_rf=0)[:8])
    amp = 30.0 * (10e-9 * tw[0]) ** 0.5
    return surrogate.ScanSpec(n_turns=DA_TURNS, n_angle=64, n_amp=128, n_pz=5, amp_max_x=amp, amp_ratio_y=1.0,
                              pz_min=-0.02, pz_max=0.02, n_cells=DA_CELLS, n_rf=8, error_seed=20261204)


def da_scan(ctx, n_train=4096, n_test=1024, train_ics=8192, epochs=6):
    """DA lattices/hour, surrogate-screened vs plain FP64 grid tracking, on
    n_test unseen knob settings of the DA workload above.  The network (4 x
    256 SiLU, BF16 MFMA, FP32 accumulation) is trained first on FP64 tracking
    of train_ics random grid ICs of each of n_train other settings; that
    one-time cost and the number of scanned lattices that repays it are
    reported beside the rates.  Headline (band mode): every IC within +-10 %
    of the predicted boundary amplitude is re-tracked in FP64, in one round,
    and decides the ray, the network standing for the untracked amplitudes.
    Second figure (bracket mode, labelled): two FP64 rounds around the
    predicted boundary, rays still open from resolved neighbours."""
    from bmx import surrogate
    spec = da_spec()
    r = surrogate.compare_scans(ctx, spec, n_train=n_train, n_test=n_test, epochs=epochs, workers=2,
                                train_ics=train_ics, modes=[surrogate.BAND_MODE, surrogate.BRACKET_MODE],
                                scan_chunk=256)
    band, bracket = r["modes"]
    keep = ("plain_lattices_per_hour", "plain_s", "ics_tracked_plain", "pet_plain", "n_train", "n_test",
            "train_ics_per_lattice", "train_samples", "train_scan_s", "train_data_s", "train_s", "loss_first",
            "loss_last", "mlp_forward_tflops", "mlp_forward_fused_tflops", "mlp_train_step_tflops")
    out = {k: r[k] for k in keep}
    out["band"] = band
    out["bracket"] = {k: bracket[k] for k in bracket if k != "rounds"}
    out["surrogate_lattices_per_hour"] = band["surrogate_lattices_per_hour"]
    out["speedup"] = band["speedup"]
    out["da_area_err_mean"] = band["da_area_err_mean"]
    out["da_area_err_max"] = band["da_area_err_max"]
    out["frac_within_2pct"] = band["frac_within_2pct"]
    out["da_area_err_all_slices_frac_within_2pct"] = band["da_area_err_all_slices_frac_within_2pct"]
    out["da_area_err_summed_slices_mean"] = band["da_area_err_summed_slices_mean"]
    out["error_definition"] = ("DA area = polygon through the surviving prefix of every ray on the on-momentum "
                               "slice; relative error against the plain scan of the same lattice, over the "
                               "lattices whose plain area is nonzero; all-slices and summed-slices figures "
                               "beside it")
    out["method"] = ("band: every IC within +-10 % of the predicted boundary amplitude re-tracked in FP64 in one "
                     "round, the network standing for the untracked amplitudes; no neighbour estimate")
    out["workload"] = (f"error-seeded {8 * DA_CELLS}-element ring with 8 rf cavities, 32 sextupole-family knobs "
                       f"U(+-20%), grid 64 x 128 x 5 (0..{spec.amp_max_x * 1e3:.2f} mm = 30 sigma_x, pz -2..2 %), "
                       f"{DA_TURNS} turns, FP64; training set {n_train} lattices x {train_ics} ICs "
                       f"(the configured set is 20,000 x 8,192)")
    return out


def opt_search(ctx, n_gen, n_init=2048):
    """The surrogate-guided sextupole optimisation (bmx.optimize) on the DA
    workload tracked for 2,048 turns: 4,096 candidates per generation
    screened by the network, the top 5 % tracked in FP64 on the full grid,
    the network retrained every 5 generations; then the plain FP64-only
    search with the same wall time.  n_gen = 20 is the full loop."""
    from bmx import optimize
    spec = da_spec()
    spec.n_turns = 2 * DA_TURNS
    r = optimize.compare_search(ctx, spec, n_gen=n_gen, n_init=n_init)
    r["workload"] = (f"error-seeded {8 * DA_CELLS}-element ring with 8 rf cavities, 32 sextupole-family knobs, "
                     f"{n_gen} generations x 4096 candidates, top 5 % tracked in FP64 on the 64 x 128 x 5 grid for "
                     f"{spec.n_turns} turns, retraining every 5 generations; plain search: FP64 only, same wall time")
    return r


def error_study_grid(ctx, n_var, per_variant=True):
    """The batched-DA error study (n_var = 256 is the whole study; fewer is a
    sample): the first n_var of the 256 seeded error variants (504..4000-element C1-style rings with 8 rf
    cavities; quad k1 x (1 + 1e-3 g), sextupole k2 x (1 + 5e-3 g), 50 um
    offsets), each scanned on the full polar grid of 64 angles x 128
    amplitudes (0..30 sigma_x at eps_x = 10 nm, y at the same amplitude) x 5
    pz in [-2 %, +2 %] for 1,024 turns, FP64.  One bmx_track_grid call over
    the whole grid, after the autotuner has picked the turn-block schedule,
    so the long-lived tails of the five pz slices share their blocks.
    lattices/hour = n_var / the wall time of the whole study (upload and
    compile of the variants, tuning, scan); the scan alone is reported
    beside it.  per_variant=False reports the spread of the variants'
    lengths and DA areas instead of one entry per variant."""
    from bmx import beam, lattice, tracker
    vs = lattice.da_variants(n_var=n_var)
    t_study = time.perf_counter()
    offs = [0]
    for v in vs:
        offs.append(offs[-1] + len(v))
    lat = ctx.lattice([e for v in vs for e in v], abi_species(), variant_offsets=offs)
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=500, n_rf=0)[:8])
    amp = 30.0 * (10e-9 * tw[0]) ** 0.5
    n_turns = 1024
    sched = tracker.tune_grid(lat, n_turns, n_angle=64, n_amp=128, n_pz=5, amp_max_x=amp, amp_ratio_y=1.0,
                              pz_min=-0.02, pz_max=0.02)
    print(f"error study: schedule {sched}", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    surv, area, _, st = tracker.track_grid(lat, n_turns, n_angle=64, n_amp=128, n_pz=5, amp_max_x=amp,
                                           amp_ratio_y=1.0, pz_min=-0.02, pz_max=0.02, return_stats=True)
    t1 = time.perf_counter()
    wall, wall_study = t1 - t0, t1 - t_study
    pet = st["pet"]
    surv = np.asarray(surv)
    print(f"error study: done in {wall:.1f} s, survivors per pz slice "
          f"{[round(float(np.mean(surv[:, i] == n_turns)), 4) for i in range(5)]}", file=sys.stderr, flush=True)
    lat.close()
    n_ele = [len(v) for v in vs]
    area = np.asarray(area)
    rec = {"workload": f"{n_var} of the 256 seeded error variants, grid 64 x 128 x 5 (0..{amp * 1e3:.2f} mm, "
                       f"pz -2..2 %), 1024 turns, FP64",
           "lattices_per_hour": n_var / wall_study * 3600.0,
           "wall_s": wall_study, "scan_only_lattices_per_hour": n_var / wall * 3600.0, "scan_wall_s": wall,
           "pet_per_s": pet / wall, "schedule": sched}
    if per_variant:
        rec["n_elements"] = n_ele
        rec["da_area_per_variant_m2"] = area.tolist()
    else:
        rec["n_elements_min_max"] = [min(n_ele), max(n_ele)]
        rec["da_area_m2_min_mean_max"] = [float(area.min()), float(area.mean()), float(area.max())]
    return rec


def abi_species():
    from bmx import abi
    return abi.ELECTRON


def radiation_rates(ctx, eles, coords, steps=3, warmup=1, turns=10):
    """The workload's tracking rate with radiation on (track1_radiation at
    both edges of every element: bmx_lattice_upload_rad, with the maps of the
    ring's 500 bends about the design orbit from bmx.radiation), damping
    alone and damping with fluctuations, against the same bunch without."""
    import torch
    from bmx import abi, radiation, tracker

    t0 = time.perf_counter()
    maps = radiation.design_orbit_maps(ctx, eles, abi.ELECTRON)
    t_maps = time.perf_counter() - t0
    out = {"maps_s": t_maps, "radiating_elements": int(np.sum(np.any(maps[:, :, 6:] != 0, axis=(1, 2)))),
           "energy_loss_per_turn_rel": float(-maps[:, :, 47].sum())}
    for name, rad in (("plain", None), ("damping", dict(damping_on=True)),
                      ("damping_fluctuations", dict(damping_on=True, fluctuations_on=True, seed=20261204))):
        lat = ctx.lattice(eles, abi.ELECTRON) if rad is None else ctx.lattice(eles, abi.ELECTRON, rad_maps=maps,
                                                                             radiation=rad)
        b = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
        for _ in range(warmup):
            tracker.track_bunch(lat, b, n_turns=turns, sync=False)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kms = 0.0
        for _ in range(steps):
            kms += tracker.track_bunch(lat, b, n_turns=turns, sync=False)["kernel_ms"]
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        n_live = int(b.state.eq(abi.ALIVE).sum().item())
        out[name] = {"particle_turns_per_s": b.n * turns * steps / wall, "kernel_ms_per_launch": kms / steps,
                     "particles_alive_end": n_live}
        lat.close()
        del b
    for name in ("damping", "damping_fluctuations"):
        out[name]["rate_vs_plain"] = out[name]["particle_turns_per_s"] / out["plain"]["particle_turns_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--particles", type=int, default=N_PARTICLES)
    ap.add_argument("--turns-per-step", type=int, default=TURNS_PER_STEP)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-da", action="store_true", help="skip the surrogate-vs-plain DA scan comparison")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 512 MiB device copy (reference for HBM byte counters) and exit")
    ap.add_argument("--da-only", action="store_true",
                    help="only run the surrogate-vs-plain DA scan comparison and print its record")
    ap.add_argument("--da-train", type=int, default=4096, help="training lattices of the DA comparison")
    ap.add_argument("--da-test", type=int, default=1024, help="test lattices of the DA comparison")
    ap.add_argument("--da-epochs", type=int, default=6, help="training epochs of the DA comparison")
    ap.add_argument("--opt-init", type=int, default=2048,
                    help="training lattices (x 8,192 ICs) of the optimisation's initial network")
    ap.add_argument("--opt-gens", type=int, default=3,
                    help="generations of the surrogate-guided optimisation (20: the full loop; 0: skip it)")
    ap.add_argument("--opt-only", action="store_true",
                    help="only run the surrogate-guided optimisation and print its record")
    ap.add_argument("--no-rad", action="store_true", help="skip the tracking rates with radiation")
    ap.add_argument("--no-error-study", action="store_true", help="skip the batched-DA error study")
    ap.add_argument("--rad-only", action="store_true", help="only run the tracking rates with radiation")
    ap.add_argument("--error-study-sample", type=int, default=0, metavar="N_VAR",
                    help="only run the batched-DA error study on its first N_VAR variants and print its record")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); particle tracking "
              f"shards by particles with no exchange, see the design notes.", file=sys.stderr)
        sys.exit(2)

    import torch
    from bmx import abi, beam, lattice, tracker

    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes_read": src.numel() * 8, "copy_calibration_bytes_written":
                          dst.numel() * 8, "launches": 3}), flush=True)
        return

    if args.da_only:
        print(json.dumps({"da_scan": da_scan(tracker.Context(0), n_train=args.da_train, n_test=args.da_test,
                                             epochs=args.da_epochs)}), flush=True)
        return
    if args.opt_only:
        print(json.dumps({"opt_search": opt_search(tracker.Context(0), args.opt_gens, args.opt_init)}), flush=True)
        return
    if args.error_study_sample > 0:
        print(json.dumps({"error_study": error_study_grid(tracker.Context(0), args.error_study_sample)}), flush=True)
        return

    eles = lattice.ring_c1(n_cells=500, n_rf=8)
    n_ele = len(eles)
    tw = beam.periodic_twiss(lattice.ring_c1(n_cells=500, n_rf=0)[:8])
    coords = beam.matched_gaussian(args.particles, tw, 10e-9, 0.1e-9, 1e-3, 1e-3, seed=20261203)

    if args.rad_only:
        print(json.dumps({"radiation": radiation_rates(tracker.Context(0), eles, coords)}), flush=True)
        return

    with open(os.path.join(ROOT, "tests", "golden", "flops_per_pet.json")) as f:
        flops = json.load(f)["C1"]
    flops_per_particle_turn = flops["flops_per_particle_turn"]

    ctx = tracker.Context(0)
    lat = ctx.lattice(eles, abi.ELECTRON)
    bunch = tracker.Bunch(coords, lat.p0c, abi.M_ELECTRON, e_tot=lat.e_tot)
    torch.cuda.synchronize()

    n_live = args.particles
    for _ in range(args.warmup):
        st = tracker.track_bunch(lat, bunch, n_turns=args.turns_per_step, sync=False)
        n_live -= st["n_lost"]

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kernel_ms = 0.0
    particle_turns = 0.0
    for _ in range(args.steps):
        st = tracker.track_bunch(lat, bunch, n_turns=args.turns_per_step, sync=False)
        lost = st["n_lost"]
        particle_turns += (n_live - 0.5 * lost) * args.turns_per_step
        n_live -= lost
        kernel_ms += st["kernel_ms"]
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    value = particle_turns / wall
    kernel_s = kernel_ms / 1e3
    achieved = flops_per_particle_turn * particle_turns / kernel_s / 1e12
    hw_per_pet = measured_hw_flops_per_pet(args.particles, args.turns_per_step)
    result = {
        "metric": "FP64 particle-turns/s",
        "value": value,
        "unit": "particle-turns/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "4000-element FODO ring with thick sextupoles, thin octupoles, 8 rf cavities; "
                               "6 GeV e-, 1e6 matched Gaussian particles, elliptical 20x10 mm aperture",
                   "particles": args.particles, "elements": n_ele, "turns_per_step": args.turns_per_step,
                   "particle_element_turns_per_s": value * n_ele, "particles_alive_end": int(n_live)},
        "roofline": {"bound": "fp64-vector", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": measured_traffic(args.particles, args.turns_per_step),
                     "flops_per_pet": flops["flops_per_pet"], "kernel_ms_per_launch": kernel_ms / args.steps,
                     "hw_fp64_flops_per_pet": hw_per_pet,
                     "hw_achieved": (None if hw_per_pet is None else
                                     hw_per_pet * particle_turns * n_ele / kernel_s / 1e12),
                     "hw_frac": (None if hw_per_pet is None else
                                 hw_per_pet * particle_turns * n_ele / kernel_s / 1e12 / FP64_PEAK_TFLOPS)},
    }
    lat.close()
    del bunch
    if not args.no_rad:
        result["radiation"] = radiation_rates(ctx, eles, coords)
    if not args.no_error_study:
        result["error_study"] = error_study_grid(ctx, 256, per_variant=False)
    if not args.no_da:
        result["da_scan"] = da_scan(ctx, n_train=args.da_train, n_test=args.da_test, epochs=args.da_epochs)
        if args.opt_gens > 0:
            result["opt_search"] = opt_search(ctx, args.opt_gens, 512)
    ctx.close()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(eles, tw)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
