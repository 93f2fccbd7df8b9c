#!/usr/bin/env python3
"""Throughput of the lattice evolution hot path on one MI355X.

Workload: two-field preheating, V = m^2 phi^2/2 + g^2 phi^2 chi^2/2 with
g^2/m^2 = 2.5e5, cosmic-time FRW background, pseudospectral Laplacian,
Yoshida-4 (symp4), L = 10/m, dt = dx/40, on a 1024^3 lattice (1.07e9
sites per field).  Initial conditions: homogeneous background
(hamiltonian_cosmic.f90:18-22 values) plus seeded vacuum fluctuations from
the device sampler.  The state is resident in HBM before timing starts.

One "step" = one Yoshida-4 step (3 force evaluations) of every site; the
metric is site-steps/s = N^3 * steps / s, output spectra of both fields
included at the configured cadence.  Both forms run from identical IC:
  plain FP64, and mixed (FP32 fields/FFT, FP64 background and reductions);
`value` is the mixed form, the FP64 form and the mixed-vs-FP64 spectrum
error are reported beside it.  After the headline (and outside its timed
region) the default run measures the other configurations into `extras`:
the 256^3 drop-in configuration (spectral and 7-point FD, both forms), the
96-member ensemble of uneven sizes and the 1024^3 run with the six GW
components (--no-extras skips them).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lattice-dev_amd"))

METRIC = "site-steps/s at 1024^3, mixed vs plain FP64, at stated energy/spectrum error"
PEAK_HBM_GBS = 8000.0

PHI0 = 1.0093430384226378929425913902459
DPHI0 = -0.7137133070120812430962278466136
MPL = 2.0e5
G2 = 2.5e5


def make_lattice(la, n, precision, laplacian=None):
    lap = la.LAP_SPECTRAL if laplacian is None else laplacian
    return la.Lattice(n, nfld=2, length=10.0, precision=precision, laplacian=lap, stencil=la.STENCIL_7PT,
                      time_var=la.COSMIC, model=la.MODEL_TWOFIELD, model_params=(1.0, G2))


def init_state(lat, seed=20240601, g2=G2):
    """Vacuum fluctuations (gamma = -1/4 fields, +1/4 momenta) on the background."""
    m2 = [1.0, g2 * PHI0 ** 2]
    lat.sample_field(0, 0, PHI0, -0.25, m2[0], MPL, seed, 0)
    lat.sample_field(0, 1, DPHI0, 0.25, m2[0], MPL, seed, 1)
    lat.sample_field(1, 0, 0.0, -0.25, m2[1], MPL, seed, 2)
    lat.sample_field(1, 1, 0.0, 0.25, m2[1], MPL, seed, 3)
    lat.set_scalars(1.0, 0.0)
    lat.calc_metric()


def algorithmic_bytes(stage, n, nf, w):
    """Fused-minimum HBM bytes per launch of each stage (w = bytes per real)."""
    nvol = float(n) ** 3
    wc = w * (1.0 + 2.0 / n)  # half-complex bytes per real site
    return {
        "x_r2c": nvol * (3 * w + wc),       # read phi, pi; write phi; write F (one field)
        "y_fwd": nf * nvol * 2 * wc,         # read + write F (all fields)
        "z_lap": nf * nvol * 2 * wc,
        "y_inv": nf * nvol * 2 * wc,
        "x_c2r": nvol * nf * (wc + 3 * w),   # read F, phi; read + write pi
        # separable form: x (one field): read phi, pi; write phi, lap.  y: read
        # phi, read + write lap.  z: read phi, lap, the other field's phi, pi;
        # write pi (all fields)
        "sep_x": nvol * 4 * w,
        "sep_y": nf * nvol * 3 * w,
        "sep_z": nf * nvol * (4 + (1 if nf == 2 else 0)) * w,
    }.get(stage)


def run_form(la, n, precision, steps, warmup, out_every, dt, sep=False, laplacian=None, tune=False, korder=None,
             place=0):
    lat = make_lattice(la, n, precision, laplacian)
    if korder is not None:
        lat.set_kspace_order(korder == "blocked")
    if sep:
        lat.set_laplacian_form(True)
    if tune:
        lat.autotune()  # launch shapes for lattices up to 256 per side
    t0 = time.time()
    init_state(lat)
    lat.synchronize()
    t_init = time.time() - t0
    placement = None
    if place > 1:
        # keep the fastest of `place` placements of the state and k-space
        # arrays in HBM (the state is carried over unchanged)
        kept, ms = lat.place_tune(place)
        placement = {"candidates": place, "kept": kept, "force_eval_ms": [round(x, 3) for x in ms]}
    lat.symp4(dt, warmup)
    lat.synchronize()
    lat.profile(True)
    done = 0
    spectra = None
    t0 = time.perf_counter()
    while done < steps:
        k = min(out_every, steps - done)
        lat.symp4(dt, k)
        done += k
        spectra = [lat.output_spectra(0), lat.output_spectra(1)]
    lat.synchronize()
    elapsed = time.perf_counter() - t0
    prof = lat.profile_read()
    lat.profile(False)
    energies = lat.energies()
    lat.close()
    return dict(elapsed=elapsed, prof=prof, spectra=spectra, energies=energies, t_init=t_init, placement=placement)


def roofline(prof, n, nf, w):
    best = None
    for stage, (ms, cnt) in prof.items():
        b = algorithmic_bytes(stage, n, nf, w)
        if b is None or cnt == 0:
            continue
        if best is None or ms > best[1]:
            best = (stage, ms, cnt, b)
    if best is None:
        return None
    stage, ms, cnt, b = best
    avg_s = ms / cnt / 1e3
    achieved = b / avg_s / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_{stage}_{n}.json")
    if w == 4 and os.path.exists(tpath):
        # measured for the mixed form (FP32 sweeps); no FP64 counterpart
        with open(tpath) as fh:
            traffic = json.load(fh).get("bytes_per_launch")
    return {"bound": "hbm", "kernel": stage, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic,
            "algorithmic_bytes_per_launch": b, "avg_launch_ms": round(ms / cnt, 4), "launches": cnt}


def cpu_baseline(n_sample, steps_sample):
    """The oracle (the C restatement of the reference's FP64 CPU path, OpenMP
    over all host threads) on a bounded sample of the same workload: the
    same two-field physics, dt = dx/40, Yoshida-4, at n_sample^3.  Built here
    with -O3 -march=native -fopenmp for this host (gcc, a few seconds)."""
    import ctypes
    import platform
    import shutil
    import subprocess
    import tempfile

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    tmp = tempfile.mkdtemp(prefix="latoracle_native_")
    so = os.path.join(tmp, "liblatoracle_native.so")
    flags = ["-O3", "-march=native", "-fcx-limited-range", "-fopenmp", "-fPIC", "-std=gnu11", "-shared"]
    subprocess.run(["gcc", *flags, os.path.join(ROOT, "oracle", "latoracle.c"), "-o", so, "-lm"], check=True)
    O._LIB_PATH = so
    O._lib = None
    try:
        dt = (10.0 / n_sample) / 40.0
        sim = O.Simulation(n_sample, 2, 10.0, O.COSMIC, model=O.MODEL_TWOFIELD, mp=(1.0, G2))
        rng = np.random.default_rng(1)
        blk = rng.standard_normal((2, 32, 32, 32))
        reps = n_sample // 32
        sim.fld[0] = PHI0 + 1e-6 * np.tile(blk[0], (reps, reps, reps))
        sim.fld[1] = 1e-6 * np.tile(blk[1], (reps, reps, reps))
        sim.fldp[0] = DPHI0
        sim.yscl = 1.0
        sim.calc_metric()
        t0 = time.perf_counter()
        sim.symp(4, dt, steps_sample)
        el = time.perf_counter() - t0
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or len(os.sched_getaffinity(0))
    model = platform.processor() or "unknown"
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"value": n_sample ** 3 * steps_sample / el, "unit": "site-steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle/latoracle.c FP64 built -O3 -march=native -fopenmp on this host, {threads} OpenMP "
                      f"threads ({model}, nproc {os.cpu_count()}), {n_sample}^3 two-field cosmic spectral "
                      f"Yoshida-4, dt = dx/40, {steps_sample} steps, {el:.1f} s"}


def spectrum_error(sa, sb, floor):
    """max |a-b|/|b| of S_phi, S_pi of both fields over k > 0 bins carrying at
    least `floor` of the peak bin power of the FP64 spectrum."""
    errs = []
    for a, b in zip(sa, sb):
        for ch in (0, 1):
            x, y = a[ch][1:], b[ch][1:]
            m = np.abs(y) >= floor * np.abs(y).max()
            m &= np.abs(y) > 0
            errs.append(float(np.max(np.abs(x[m] - y[m]) / np.abs(y[m]))))
    return max(errs)


def corrector_leg(la, args):
    """The learned corrector as config 4 states it: one network trained on
    --corr-pairs snapshot pairs that the plain integrator generates at
    --corr-train-n (L scaled so that dx and dt equal the evaluation
    lattice's: the per-site dynamics are the same, only the box is smaller),
    along trajectories with g^2/m^2 log-uniform in [1e4, 1e6] and their own
    seeds, --corr-per-traj pairs each (a coupling whose coarse leapfrog step
    is not finite gives no pairs and its trajectory is dropped after one);
    then applied at --corr-n with coarse steps of --corr-cf times dt.
    Reported: a learning curve (held-out one-step error of the corrected
    coarse step over that of the leapfrog alone, against the plain
    Yoshida-4 step, on states the training never saw); and per evaluation
    coupling, over --corr-steps coarse steps, the trajectory error of the
    leapfrog alone, of the always-accepted corrector, of the corrector
    under the 1e-6 constraint bound (plain fallback) and under the judge
    (every --corr-judge-th step re-taken by the plain integrator, switched
    off above --corr-err-tol), with force evaluations (coarse step 1,
    Yoshida-4 step 3) against the plain run.  Errors are relative to the
    fluctuation norm of the plain run, for fields / momenta."""
    cf = args.corr_cf
    n, ntr = args.corr_n, args.corr_train_n
    dt = (10.0 / n) / 40.0
    ltr = 10.0 * ntr / n
    rng = np.random.default_rng(args.corr_seed)
    out = {"metric": "learned corrector: error and force evaluations against plain Yoshida-4",
           "coarse_factor": cf, "train_n": ntr, "train_L": ltr, "eval_n": n, "dt": dt}

    def fluct_err(st, ref):
        e = []
        for k in (0, 1):
            d = np.linalg.norm(st[k] - ref[k])
            r = np.linalg.norm(ref[k] - ref[k].mean(axis=(1, 2, 3), keepdims=True))
            e.append(float(d / r))
        return e

    # held-out states: own seeds, couplings where the coarse step is stable
    val = [(float(g), 90000 + q, w) for q, (g, w) in enumerate([(1e4, 40), (2e4, 80), (4e4, 40), (1e4, 160)])]

    def one_step_errors(corr):
        res = []
        for g2, seed, warm in val:
            with la.Lattice(ntr, nfld=2, length=ltr, model_params=(1.0, g2)) as lat:
                init_state(lat, seed=seed, g2=g2)
                lat.symp4(dt, warm)
                start = lat.get_state()
                lat.symp4(dt, cf)
                fine = lat.get_state()
                lat.set_state(*start)
                lat.symp(2, cf * dt, 1)
                leap = lat.get_state()
                lat.set_state(*start)
                corr.step(lat, dt, cf, nsteps=1, threshold=1e300)
                cor = lat.get_state()
            el, ec = fluct_err(leap, fine), fluct_err(cor, fine)
            res.append({"g2": g2, "warm_steps": warm, "err_leapfrog": el, "err_corrected": ec,
                        "ratio": [ec[0] / el[0], ec[1] / el[1]]})
        return res

    # the coarse leapfrog step is stable while (k_max^2 + g^2 phi^2) (cf dt)^2
    # < 4, k_max^2 = 3 (pi / dx)^2 for the spectral Laplacian and |phi| up to
    # PHI0: couplings above g2max have no usable coarse step
    g2max = (4.0 - 3.0 * math.pi ** 2 * (cf / 40.0) ** 2) / ((cf * dt) ** 2 * PHI0 ** 2)
    out["g2_stable_max"] = g2max
    t0 = time.perf_counter()
    curve, trajs, used = [], [], 0
    checkpoints = sorted({int(args.corr_pairs * f) for f in (0.05, 0.1, 0.25, 0.5, 0.75, 1.0)})
    nxt = 0

    def new_stream(q, g2=None):
        if g2 is None:
            g2 = float(np.exp(rng.uniform(np.log(1e4), np.log(min(g2max, 1e6)))))
        lat = la.Lattice(ntr, nfld=2, length=ltr, model_params=(1.0, g2))
        init_state(lat, seed=50000 + q, g2=g2)
        warm = int(rng.integers(0, 400))
        if warm:
            lat.symp4(dt, warm)
        return {"lat": lat, "g2": g2, "warm": warm, "pairs": 0, "losses": []}

    with la.Corrector(seed=args.corr_seed) as corr:
        # interleaved trajectories, one pair from each in turn; the first
        # (which fixes the target scale) at the middle of the stable range
        nq = 0
        streams = []
        for k in range(args.corr_streams):
            streams.append(new_stream(nq, g2=math.sqrt(1e4 * min(g2max, 1e6)) if k == 0 else None))
            nq += 1
        while used < args.corr_pairs:
            for k, st in enumerate(streams):
                if used >= args.corr_pairs:
                    break
                # learning rate decaying exponentially to a tenth over the run
                lr = args.corr_lr * 0.1 ** (used / args.corr_pairs)
                loss = corr.train(st["lat"], dt, cf, npairs=1, lr=lr)[0]
                if np.isfinite(loss):
                    used += 1
                    st["pairs"] += 1
                    st["losses"].append(float(loss))
                if not np.isfinite(loss) or st["pairs"] >= args.corr_per_traj:
                    trajs.append({"g2": st["g2"], "warm": st["warm"], "pairs": st["pairs"],
                                  "loss_first": float(np.mean(st["losses"][:5])) if st["losses"] else None,
                                  "loss_last": float(np.mean(st["losses"][-5:])) if st["losses"] else None})
                    st["lat"].close()
                    streams[k] = new_stream(nq)
                    nq += 1
                if used >= nxt:
                    nxt += 500
                    print(f"[bench] corrector: {used} pairs, {time.perf_counter() - t0:.0f} s", file=sys.stderr,
                          flush=True)
                if checkpoints and used >= checkpoints[0]:
                    checkpoints.pop(0)
                    ev = one_step_errors(corr)
                    curve.append({"pairs": used, "seconds": round(time.perf_counter() - t0, 1),
                                  "median_ratio_fields": float(np.median([r["ratio"][0] for r in ev])),
                                  "median_ratio_momenta": float(np.median([r["ratio"][1] for r in ev])),
                                  "states": ev})
                    print(f"[bench] corrector: {used} pairs, held-out ratio "
                          f"{curve[-1]['median_ratio_fields']:.3f} / {curve[-1]['median_ratio_momenta']:.3f}",
                          file=sys.stderr, flush=True)
        for st in streams:
            st["lat"].close()
        stable = [t["g2"] for t in trajs if t["pairs"] > 0]
        out["train"] = {"pairs": used, "trajectories": len(trajs), "usable_trajectories": len(stable),
                        "largest_usable_g2": max(stable) if stable else None,
                        "seconds": round(time.perf_counter() - t0, 1), "ratio": corr.get_norm().tolist()}
        out["learning_curve"] = curve
        runs = []
        for q, g2 in enumerate(float(x) for x in args.corr_g2.split(",")):
            run = {"g2": g2}
            with la.Lattice(n, nfld=2, length=10.0, model_params=(1.0, g2)) as lat:
                init_state(lat, seed=2000 + q, g2=g2)
                e0 = lat.energies()["resid"]
                start = lat.get_state()
                lat.symp4(dt, cf * args.corr_steps)
                ref = lat.get_state()
                d_plain = lat.energies()["resid"] - e0
                lat.set_state(*start)
                lat.symp(2, cf * dt, args.corr_steps)
                leap = lat.get_state()
                lat.set_state(*start)
                corr.step(lat, dt, cf, nsteps=args.corr_steps, threshold=1e300)
                always = lat.get_state()
                lat.set_state(*start)
                st_b = corr.step(lat, dt, cf, nsteps=args.corr_steps, threshold=args.corr_bound)
                bounded = lat.get_state()
                lat.set_state(*start)
                st_j = corr.step_judged(lat, dt, cf, nsteps=args.corr_steps, threshold=args.corr_bound,
                                        judge_every=args.corr_judge, err_tol=args.corr_err_tol)
                judged = lat.get_state()
            plain = 3.0 * cf * args.corr_steps
            run.update({"coarse_steps": args.corr_steps, "plain_resid_change": d_plain,
                        "err_leapfrog": fluct_err(leap, ref), "err_corrected": fluct_err(always, ref),
                        "bound": args.corr_bound, "bound_fallbacks": st_b["fallbacks"],
                        "bound_saving": plain / st_b["force_evals"], "err_bound": fluct_err(bounded, ref),
                        "judge_every": args.corr_judge, "err_tol": args.corr_err_tol,
                        "judged": st_j["judged"], "max_judged_error": st_j["max_judged_error"],
                        "max_judged_leapfrog_error": st_j["max_judged_leapfrog_error"],
                        "switched_off_at": st_j["switched_off_at"], "judged_saving": plain / st_j["force_evals"],
                        "err_judged": fluct_err(judged, ref)})
            runs.append(run)
            print(f"[bench] corrector eval g2={g2:g}: leapfrog {run['err_leapfrog']}, corrected "
                  f"{run['err_corrected']}, judged {run['err_judged']} saving {run['judged_saving']:.2f}",
                  file=sys.stderr, flush=True)
        out["eval"] = runs
    print(json.dumps(out))


def ensemble_leg(la, args):
    """Batched ensemble: 96 realisations of the two-field model, 32 x 64^3 +
    48 x 128^3 + 16 x 192^3, fixed L = 10/m (dt = dx/40 per size class),
    g^2/m^2 log-uniform in [1e4, 1e6] (seed 11), realisation r seeded 3000 + r.
    One launch per FFT sweep for all size classes; the Yoshida-4 call is
    captured into a hipGraph on its first use and replayed.  Reported:
    aggregate site-steps/s with the graph, eagerly, and with one launch per
    sweep per class (graph), and the spectra of one realisation per class
    against a single-context run of the same realisation."""
    classes = [(64, 32, 10.0), (128, 48, 10.0), (192, 16, 10.0)]
    rng = np.random.default_rng(11)
    g2s = np.exp(rng.uniform(np.log(1e4), np.log(1e6), sum(c for _, c, _ in classes)))
    params = [(1.0, float(g)) for g in g2s]
    prec = la.MIXED if args.ens_form == "mixed" else la.F64
    dts = [(10.0 / n) / 40.0 for n, _, _ in classes]
    sizes = [n for n, c, _ in classes for _ in range(c)]
    sites = float(sum(n ** 3 for n in sizes))
    samples = [0, 32, 80]  # one member of each class
    k = args.ens_steps
    out = {"metric": "aggregate site-steps/s of a 96-realisation ensemble of uneven sizes",
           "classes": [{"n": n, "count": c, "L": L} for n, c, L in classes], "form": args.ens_form,
           "steps": k, "sites_per_field": sites}
    got = {}
    # single: one launch per FFT sweep for all classes; per_class: one per class
    for name, single, graph in (("graph", True, True), ("eager", True, False), ("graph_per_class", False, True),
                                ("eager_per_class", False, False)):
        with la.Ensemble(classes, precision=prec, members=params) as ens:
            out.setdefault("autotuner_picked_single_launch", ens.single_launch(None))
            assert ens.single_launch(single) == single
            ens.use_graph(graph)
            views = [ens.member(i) for i in range(ens.size)]
            for i, v in enumerate(views):
                init_state(v, seed=3000 + i, g2=float(g2s[i]))
            ens.symp(4, dts, k)  # first call: captured (graph) or warm-up (eager)
            ens.synchronize()
            t0 = time.perf_counter()
            ens.symp(4, dts, k)
            ens.synchronize()
            t = time.perf_counter() - t0
            out[name] = {"seconds": t, "site_steps_per_s": sites * k / t}
            if name == "graph":
                got = {i: views[i].output_spectra(0) for i in samples}
            for v in views:
                v.close()
    errs = []
    for i in samples:
        n = sizes[i]
        with la.Lattice(n, nfld=2, length=10.0, precision=prec, model_params=params[i]) as lat:
            init_state(lat, seed=3000 + i, g2=float(g2s[i]))
            lat.symp4(dts[[c[0] for c in classes].index(n)], 2 * k)
            ref = lat.output_spectra(0)
        a, b = got[i][0][1:], ref[0][1:]
        m = np.abs(b) >= 1e-6 * np.abs(b).max()
        errs.append(float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m]))))
    # bins >= 1e-6 of the peak; the ensemble's autotuned launch shapes may
    # differ from the single context's, which changes the rounding only
    out["spectrum_rel_err_vs_single_runs"] = errs
    # the mode the ensemble's autotuner keeps
    best = "graph" if out["autotuner_picked_single_launch"] else "graph_per_class"
    out["value"] = out[best]["site_steps_per_s"]
    out["value_mode"] = best
    return out


def config1_leg(la, args):
    """The 256^3 drop-in configuration: two-field preheating (g^2/m^2 =
    2.5e5, cosmic FRW, L = 10/m, dt = dx/40, Yoshida-4, spectra of both
    fields every 20 steps) with the spectral and the 7-point FD Laplacian,
    FP64 and mixed, launch shapes autotuned.  site-steps/s, and the HBM rate
    of the model bytes per site-step (spectral 3 nf (7w + 8w_c), FD 18 nf w)
    against 8 TB/s."""
    n = args.c1_n
    dt = (10.0 / n) / 40.0
    out = {"metric": f"site-steps/s at {n}^3, spectral and 7-point FD Laplacian", "steps": args.c1_steps,
           "warmup": args.warmup, "spectra_every": 20, "runs": []}
    for lap_name, lap in (("spectral", la.LAP_SPECTRAL), ("fd7", la.LAP_STENCIL)):
        for form, prec in (("f64", la.F64), ("mixed", la.MIXED)):
            r = run_form(la, n, prec, args.c1_steps, args.warmup, 20, dt, laplacian=lap, tune=True)
            w = 8 if form == "f64" else 4
            wc = w * (1.0 + 2.0 / n)
            model = 3 * 2 * (7 * w + 8 * wc) if lap_name == "spectral" else 18 * 2 * w
            v = float(n) ** 3 * args.c1_steps / r["elapsed"]
            out["runs"].append({"laplacian": lap_name, "form": form, "site_steps_per_s": v,
                                "ms_per_step": 1e3 * r["elapsed"] / args.c1_steps,
                                "model_bytes_per_site_step": model, "hbm_frac_vs_8TBs": v * model / 8e12,
                                "stages_ms": {k: round(x[0], 3) for k, x in r["prof"].items() if x[1]}})
            print(f"[bench] {n}^3 {lap_name} {form}: {v:.4e} site-steps/s", file=sys.stderr, flush=True)
    return out


def gw_leg(la, args):
    """The headline physics with the six metric perturbations u_ij (spectral,
    sourced by the field gradients, TT projection in the GW spectrum),
    u = pi_u = 0 at the start, both forms from identical IC: site-steps/s of
    each form and the mixed form's errors against FP64 after the run (field
    spectra and GW spectra over bins >= 1e-6 of the peak, energy)."""
    n = args.gw_n
    dt = (10.0 / n) / 40.0
    res = {}
    for form, prec in (("f64", la.F64), ("mixed", la.MIXED)):
        with la.Lattice(n, nfld=2, length=10.0, precision=prec, laplacian=la.LAP_SPECTRAL, time_var=la.COSMIC,
                        model=la.MODEL_TWOFIELD, model_params=(1.0, G2), gw=True, gw_mpl=1.0) as lat:
            init_state(lat)
            placement = None
            if args.place_tune > 1:
                kept, ms = lat.place_tune(args.place_tune)
                placement = {"candidates": args.place_tune, "kept": kept, "force_eval_ms": [round(x, 3) for x in ms]}
            lat.symp4(dt, args.warmup)
            lat.synchronize()
            t0 = time.perf_counter()
            lat.symp4(dt, args.gw_steps)
            lat.synchronize()
            el = time.perf_counter() - t0
            res[form] = {"site_steps_per_s": float(n) ** 3 * args.gw_steps / el,
                         "ms_per_step": 1e3 * el / args.gw_steps,
                         "spectra": [lat.output_spectra(0), lat.output_spectra(1)],
                         "gw": lat.gw_spectrum(), "e": lat.energies(), "placement": placement}
        print(f"[bench] gw {form} {n}^3: {res[form]['site_steps_per_s']:.4e} site-steps/s", file=sys.stderr,
              flush=True)

    def maxrel(a, b, floor=1e-6):
        a, b = np.asarray(a)[1:], np.asarray(b)[1:]
        m = np.abs(b) >= floor * np.abs(b).max()
        return float(np.max(np.abs(a[m] - b[m]) / np.abs(b[m])))

    d, m = res["f64"], res["mixed"]
    w = 4
    out = {"metric": f"site-steps/s at {n}^3 with the six GW components, mixed vs plain FP64",
           "steps": args.gw_steps, "warmup": args.warmup,
           "mixed": {k: m[k] for k in ("site_steps_per_s", "ms_per_step", "placement")},
           "f64": {k: d[k] for k in ("site_steps_per_s", "ms_per_step", "placement")},
           "mixed_over_fp64": m["site_steps_per_s"] / d["site_steps_per_s"],
           # model bytes per site-step with GW (two fields): 3 [2 (7w + 8wc) + 6 (6w + 8wc) + 2w]
           "hbm_frac_mixed_vs_8TBs": m["site_steps_per_s"] * 3 * (2 * (7 * w + 8 * w) + 6 * (6 * w + 8 * w)
                                                                 + 2 * w) / 8e12,
           "mixed_vs_f64": {"field_spectra": max(maxrel(m["spectra"][l][q], d["spectra"][l][q])
                                                 for l in (0, 1) for q in (0, 1)),
                            "gw_spectrum_pi": maxrel(m["gw"][0], d["gw"][0]),
                            "gw_spectrum_u": maxrel(m["gw"][1], d["gw"][1]),
                            "energy": abs(m["e"]["rho"] - d["e"]["rho"]) / abs(d["e"]["rho"])},
           "after_steps": args.warmup + args.gw_steps}
    out["value"] = out["mixed"]["site_steps_per_s"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1024)
    ap.add_argument("--output-every", type=int, default=20)
    ap.add_argument("--forms", default="f64,mixed")
    ap.add_argument("--lap-form", default="3d", choices=["3d", "sep"],
                    help="mixed form: five-sweep 3-D spectral Laplacian or the separable three-sweep form")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=2)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's plain 2 GiB device copy (4 launches) to calibrate HBM byte "
                         "counters")
    ap.add_argument("--corrector", action="store_true", help="only run the learned-corrector leg")
    ap.add_argument("--corr-train-n", type=int, default=128)
    ap.add_argument("--corr-n", type=int, default=256)
    ap.add_argument("--corr-pairs", type=int, default=20000, help="training pairs in all")
    ap.add_argument("--corr-per-traj", type=int, default=100, help="training pairs per trajectory")
    ap.add_argument("--corr-streams", type=int, default=8, help="trajectories trained on in turn")
    ap.add_argument("--corr-seed", type=int, default=7)
    ap.add_argument("--corr-steps", type=int, default=20)
    ap.add_argument("--corr-lr", type=float, default=3e-4)
    ap.add_argument("--corr-cf", type=int, default=8, help="coarse step = this many fine steps")
    ap.add_argument("--corr-judge", type=int, default=5,
                    help="judged run: every this many accepted corrected steps is re-taken by the plain integrator")
    ap.add_argument("--corr-err-tol", type=float, default=0.5,
                    help="judged run: switch the corrector off when a judged one-step error exceeds this (or the "
                         "leapfrog step's)")
    ap.add_argument("--ensemble", action="store_true", help="only run the batched-ensemble leg")
    ap.add_argument("--gw", action="store_true", help="only run the leg with the six GW components")
    ap.add_argument("--config1", action="store_true",
                    help="only run the 256^3 leg with the spectral and the 7-point FD Laplacian")
    ap.add_argument("--c1-n", type=int, default=256)
    ap.add_argument("--c1-steps", type=int, default=200)
    ap.add_argument("--gw-n", type=int, default=1024)
    ap.add_argument("--gw-steps", type=int, default=5)
    ap.add_argument("--ens-steps", type=int, default=20)
    ap.add_argument("--ens-form", default="mixed", choices=["mixed", "f64"])
    ap.add_argument("--corr-bound", type=float, default=1e-6,
                    help="fall back when a coarse step moves the relative Friedmann residual by more than this")
    ap.add_argument("--corr-g2", default="1e4,2e4,3e4,4e4,6e4,1e5,1e6")
    ap.add_argument("--kspace-order", default=None, choices=["blocked", "fftw"],
                    help="row order of the force pipeline's k-space arrays (default: the library's)")
    ap.add_argument("--place-tune", type=int, default=4,
                    help="1024^3 headline forms: keep the fastest of this many HBM placements of the state and "
                         "k-space arrays (lat_place_tune; 1 = as allocated)")
    ap.add_argument("--no-extras", action="store_true",
                    help="default run: skip the 256^3 drop-in, ensemble and GW legs after the headline")
    args = ap.parse_args()
    if args.corrector:
        import lattice_amd as la

        corrector_leg(la, args)
        return
    if args.ensemble:
        import lattice_amd as la

        print(json.dumps(ensemble_leg(la, args)))
        return
    if args.gw:
        import lattice_amd as la

        print(json.dumps(gw_leg(la, args)))
        return
    if args.config1:
        import lattice_amd as la

        print(json.dumps(config1_leg(la, args)))
        return
    if args.copy_calibration:
        # the library's plain copy kernel (k_copy16) over 2 GiB, 1 + 3 launches:
        # per launch it reads and writes exactly 2 GiB
        import lattice_amd as la

        with la.Lattice(8, nfld=2, length=1.0) as lat:
            gbs = lat.copy_bandwidth(2 << 30, 3)
        print(json.dumps({"copy_bytes_read_per_launch": 2 << 30, "copy_bytes_written_per_launch": 2 << 30,
                          "copy_GBs": gbs}))
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU only; multi-GPU sharding is a design note, "
              "not a measured configuration (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import lattice_amd as la

    n = args.n
    dt = (10.0 / n) / 40.0
    forms = args.forms.split(",")
    with la.Lattice(8, nfld=2, length=1.0) as probe:
        copy_gbs = probe.copy_bandwidth(2 << 30, 5)
    print(f"[bench] measured copy bandwidth {copy_gbs:.1f} GB/s", file=sys.stderr, flush=True)
    res = {}
    for form in forms:
        prec = la.F64 if form == "f64" else la.MIXED
        print(f"[bench] {form} {n}^3: warmup {args.warmup}, timing {args.steps} steps", file=sys.stderr, flush=True)
        res[form] = run_form(la, n, prec, args.steps, args.warmup, args.output_every, dt,
                             sep=(form == "mixed" and args.lap_form == "sep"), korder=args.kspace_order,
                             place=args.place_tune)
        r = res[form]
        print(f"[bench] {form}: {r['elapsed']:.3f} s, {n ** 3 * args.steps / r['elapsed']:.4e} site-steps/s "
              f"(init {r['t_init']:.1f} s)", file=sys.stderr, flush=True)
    head = "mixed" if "mixed" in res else forms[0]
    r = res[head]
    value = n ** 3 * args.steps / r["elapsed"]
    w = 4 if head == "mixed" else 8
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "site-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r["elapsed"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 fields/FFT + f64 background/reductions" if head == "mixed" else "f64",
        "data": "synthetic: seeded vacuum fluctuations (device Philox sampler) on the homogeneous two-field background",
        "config": {"workload": f"{n}^3 two-field preheating g^2/m^2=2.5e5, cosmic FRW, spectral Laplacian, "
                               f"Yoshida-4, L=10/m, dt=dx/40, spectra every {args.output_every} steps",
                   "form": head,
                   "laplacian_form": ("separable, 3 sweeps" if args.lap_form == "sep" and head == "mixed"
                                      else "3-D, 5 sweeps"),
                   "placement": (f"fastest of {args.place_tune} HBM placements (lat_place_tune)"
                                 if args.place_tune > 1 else "as allocated")},
        "roofline": roofline(r["prof"], n, 2, w),
        "hbm_copy_GBs": round(copy_gbs, 1),
    }
    if line["roofline"]:
        line["roofline"]["frac_of_copy"] = round(line["roofline"]["achieved"] / copy_gbs, 4)
    if "f64" in res:
        f = res["f64"]
        v64 = n ** 3 * args.steps / f["elapsed"]
        line["fp64"] = {"value": v64, "ms_per_step": 1e3 * f["elapsed"] / args.steps,
                        "roofline": roofline(f["prof"], n, 2, 8)}
        if head == "mixed":
            line["mixed_over_fp64"] = value / v64
            line["spectrum_rel_err_vs_fp64"] = {
                "bins_ge_1e-3_peak": spectrum_error(r["spectra"], f["spectra"], 1e-3),
                "bins_ge_1e-6_peak": spectrum_error(r["spectra"], f["spectra"], 1e-6),
                "all_bins": spectrum_error(r["spectra"], f["spectra"], 0.0),
                "after_steps": args.warmup + args.steps}
            line["energy_rel_err_vs_fp64"] = abs(r["energies"]["rho"] - f["energies"]["rho"]) / abs(
                f["energies"]["rho"])
            line["friedmann_resid"] = {"mixed": r["energies"]["resid"], "f64": f["energies"]["resid"]}
    for form in res:
        line.setdefault("stages_ms", {})[form] = {k: round(v[0], 3) for k, v in res[form]["prof"].items() if v[1]}
        if res[form]["placement"]:
            line.setdefault("placement", {})[form] = res[form]["placement"]
    if not args.no_cpu:
        print("[bench] cpu baseline (oracle)", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(args.cpu_n, args.cpu_steps)
    else:
        line["cpu_baseline"] = None
    if not args.no_extras and n == 1024:
        # the other configurations, measured in the same run (not the headline)
        extras = {}
        c1 = config1_leg(la, args)
        extras["config1_256"] = {"steps": c1["steps"], "runs": [
            {k: r[k] for k in ("laplacian", "form", "site_steps_per_s", "ms_per_step", "hbm_frac_vs_8TBs")}
            for r in c1["runs"]]}
        ens = ensemble_leg(la, args)
        extras["ensemble_96"] = {k: ens[k] for k in ("classes", "form", "steps", "value", "value_mode",
                                                      "autotuner_picked_single_launch",
                                                      "spectrum_rel_err_vs_single_runs")}
        extras["ensemble_96"]["modes"] = {m: ens[m]["site_steps_per_s"]
                                          for m in ("graph", "eager", "graph_per_class", "eager_per_class")}
        gw = gw_leg(la, args)
        extras["gw_1024"] = {k: gw[k] for k in ("steps", "mixed", "f64", "mixed_over_fp64",
                                                 "hbm_frac_mixed_vs_8TBs", "mixed_vs_f64", "after_steps")}
        line["extras"] = extras
    print(json.dumps(line))


if __name__ == "__main__":
    main()
