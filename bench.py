"""Benchmark of the RK3/Crank-Nicolson channel time step on one MI355X.

Workload: single periodic channel, flow only, Re_tau = 395 (REN = 6875), grid 256 x 192 x 256
(x, y, z), box 4 pi x 2 x 2 pi, y tanh-stretched with beta = 2.0, dt = 5e-4.  Synthetic start:
the Re_tau 180 field (64^3, 100 steps from laminar + 10 % seeded noise) interpolated onto the
mesh (chs_interp_initial_tg), re-perturbed by 5 % (seeded) and projected by the NS = 0 substep.  A "step" is one full RK3 step (three substeps SOLVERRK3_MOM_tg(1..3)) of the whole
grid with all fields resident in HBM.

Prints one JSON line.  ``value`` is grid-point RK3-steps/s of the form named in
config.form (plain FP64 unless --form mixed).  ``roofline`` is the dominant kernel's
algorithmic HBM bytes per launch over its average launch time, measured with HIP events on
the solver's stream inside the timed region.  ``cpu_baseline`` is the reference Fortran/MPI
program (oracle/_ref/chapsim, built from the reference sources) timed on this host's cores
on a bounded sample of the same workload (a few steps), or the C oracle on one core if the
reference build is absent.

Usage: python bench.py [--gpus 1] [--steps 200] [--warmup 10] [--form plain|mixed]
                       [--no-cpu-baseline] [--graph]
Other workloads (each prints its own JSON line; not the default run):
  --workload channel64   64^3 channel, Re_tau 180, dt 1e-3
  --workload poisson     standalone pressure Poisson 512x384x512, beta 2.2
  --workload ensemble    32 channels of seeded uneven sizes, one hipGraph per step
  --workload heated      heated channel 256x128x256 (energy + variable-property momentum,
                         Crank-Nicolson viscous scheme; --vis explicit for iVisScheme 0)
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
CFG = dict(L=256, M=192, N=256, HX=4 * np.pi, HZ=2 * np.pi, STR2=2.0, REN=6875.0, dt=5.0e-4)
CFG64 = dict(L=64, M=64, N=64, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8, REN=2800.0, dt=1.0e-3)


def alg_bytes(kernel: str, L: int, M: int, N: int, wbytes: int) -> float:
    """Algorithmic HBM bytes of one launch: every array element the kernel must read or
    write, counted once.  F = L*M*N cells; fields u,v,w,p,phi,div are FP64; RHS/history and
    spectral arrays use the work precision (wbytes)."""
    F = float(L * M * N)
    S = float((L // 2 + 1) * M * N)          # complex spectral values
    table = {
        # R u,v,w,p (4F, FP64) + H0 x3; W RHS x3 + H0 x3
        "conv_rhs": 4 * F * 8 + 3 * F * wbytes + 6 * F * wbytes,
        # R RHS x3, W RHS x3 (v skips row 1); one launch for the three components
        "cn_x": 6 * F * wbytes,
        # z stencils, one launch for the three components: R + W RHS x3
        "cn_z": 6 * F * wbytes,
        # x and z together, three components: R + W RHS x3
        "cn_xz": 6 * F * wbytes,
        # y stencil + Q update of one component: R RHS, R q, W predicted q (one launch per
        # component, or the three in one launch)
        "cn_y": F * wbytes + 2 * F * 8,
        # z and y solves in one pass, three components: R RHS x3, R q x3, W predicted q x3
        "cn_zy": 3 * F * wbytes + 6 * F * 8,
        # R predicted u,v,w; W spectrum (the correction recomputes the divergence)
        "divg_fftx": 3 * F * 8 + S * 2 * wbytes,
        "fftz_fwd": 2 * S * 2 * wbytes,
        # R spectrum + pivots (work precision; FP64 for six low modes), W spectrum
        "ytdma": 2 * S * 2 * wbytes + S * wbytes,
        "fftz_inv": 2 * S * 2 * wbytes,
        # R spectrum, W phi
        "fftx_inv": S * 2 * wbytes + F * 8,
        # R phi, predicted u, v, w, p; W u, v, w, p
        "update": 9 * F * 8,
        # refinement correction: R phi, u, v, w, p; W u, v, w, p
        "refine_update": 9 * F * 8,
    }
    return table[kernel]


def canonical_step_bytes(L: int, M: int, N: int, wbytes: int) -> float:
    """Algorithmic bytes of one RK3 step by the fixed count of compulsory passes, independent of
    how the passes are fused: per substep 48 field passes, 22 of FP64 state (u, v, w, p reads
    and writes, the divergence's velocity reads, the correction's reads and writes) and 26 of
    work-precision arrays (RHS, history, spectra): explicit RHS with the x-solve 13, z-solves 6,
    y-solves with the velocity update 9, divergence + x-transform 4, z-transform 2, spectral
    y-Thomas 2, inverse z-transform 2, inverse x-transform + correction 10.  FP64: 1152 B per
    grid-point RK3 step.  The fused z+y pass does not move the z-solves' 6 passes; what the
    launched kernels must move is reported beside it (kernel_sum_bytes_per_step)."""
    F = float(L * M * N)
    return 3 * (22 * 8 + 26 * wbytes) * F


KERNELS = ["conv_rhs", "cn_x", "cn_z", "cn_xz", "cn_y", "cn_zy", "divg_fftx", "fftz_fwd", "ytdma",
           "fftz_inv", "fftx_inv", "update", "refine_update"]


def load_pkg():
    import __graft_entry__
    return __graft_entry__.load()


def start_field(pkg, c):
    """The workload's start: the Re_tau 180 field (64^3, beta 1.8, REN 2800: 100 RK3 steps from
    the laminar profile with 10 % seeded noise, projected) interpolated onto the workload's mesh
    (INITIAL_INTERP_tg restated, chs_interp_initial_tg) and re-perturbed by 5 % (seeded); the
    NS = 0 substep then projects it.  For the 64^3 workload itself: laminar + 5 %."""
    from chapsim_amd import init
    c1 = CFG64
    if (c["L"], c["M"], c["N"]) == (c1["L"], c1["M"], c1["N"]):
        return pkg.poiseuille_perturbed(pkg.channel_setup(c1["L"], c1["M"], c1["N"], c1["HX"], c1["HZ"], c1["STR2"]),
                                        seed=395, amp=0.05)
    co1 = pkg.channel_setup(c1["L"], c1["M"], c1["N"], c1["HX"], c1["HZ"], c1["STR2"])
    ch1 = pkg.ChannelTG(co1, pkg.PLAIN_FP64)
    q1, p1 = pkg.poiseuille_perturbed(co1, seed=180, amp=0.1)
    ch1.upload(q1, p1)
    ch1.SOLVERRK3_MOM_tg(0, c1["dt"], 1.0 / c1["REN"])
    for _ in range(100):
        ch1.rk3_step(c1["dt"], 1.0 / c1["REN"])
    q1, p1 = ch1.download()
    ch1.close()
    q, pr = pkg.interp_initial_tg(q1, p1, c1["HX"], c1["HZ"], init.channel_ynd(c1["M"], c1["STR2"]),
                                  c["L"], c["M"], c["N"], c["HX"], c["HZ"], c["STR2"])
    return init.add_perturbation(q, seed=395, amp=0.05), pr


def run_gpu(pkg, form: str, steps: int, warmup: int, c=None, graph=False, autotune=False, shapes=None,
            roundtrip=0, start=None):
    c = c or CFG
    coeffs = pkg.channel_setup(c["L"], c["M"], c["N"], c["HX"], c["HZ"], c["STR2"])
    mode = pkg.MIXED if form == "mixed" else pkg.PLAIN_FP64
    ch = pkg.ChannelTG(coeffs, mode)
    q, pr = start if start is not None else start_field(pkg, c)
    ch.upload(q, pr)
    cv = 1.0 / c["REN"]
    dt = c["dt"]
    ch.SOLVERRK3_MOM_tg(0, dt, cv)
    q0, p0 = ch.download()          # projected start, shared with the CPU baseline
    if form == "mixed":
        ch.set_refinement(0, 2)
    tuning = None
    if shapes:
        ch.set_tuning(**shapes)
        tuning = dict(shapes)
    if autotune:
        from chapsim_amd import tune
        tuning = tune.autotune(ch, dt, cv, steps=3, verbose=True)["best"]
    # profiling pre-pass: which kernel dominates a step
    ch.profile_reset(None)
    for _ in range(2):
        ch.rk3_step(dt, cv)
    ch.sync()
    per, launches = {}, {}
    for k in KERNELS:
        avg, n = ch.profile_get(k)
        per[k] = avg * n / 2.0          # ms per step
        launches[k] = n / 2.0
    dom = max(per, key=per.get)
    pre_avg, pre_n = ch.profile_get(dom)
    for _ in range(warmup):
        ch.rk3_step(dt, cv)
    if graph:                       # replayed hipGraph: no per-kernel events in the timed region
        ch.set_graph(True)
        ch.rk3_step(dt, cv)
        ch.sync()
        ch.profile_reset("none")
    else:                           # timed region; events only around the dominant kernel
        ch.profile_reset(dom)
    ch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        ch.rk3_step(dt, cv)
    ch.sync()
    t1 = time.perf_counter()
    dom_avg_ms, dom_n = ch.profile_get(dom)
    if graph:                       # kernel time from the profiling pre-pass instead
        dom_avg_ms, dom_n = pre_avg, pre_n
    cfl = ch.cfl(dt)               # CFLMM * DT of the last step's new field, formed inside it
    div = ch.max_div()
    mon = ch.monitor(c["REN"])
    final = ch.download()
    rt = host_roundtrip(pkg, ch, dt, cv, roundtrip) if roundtrip else None
    ch.close()
    ms = (t1 - t0) * 1e3 / steps
    return dict(ms_per_step=ms, per_kernel_ms=per, launches=launches, dominant=dom,
                dom_avg_ms=dom_avg_ms, dom_count=dom_n, cfl=cfl, max_div=div,
                wbytes=4 if form == "mixed" else 8, start=(q0, p0), tuning=tuning, roundtrip_ms=rt,
                final=final, monitor=mon, steps_run=2 + warmup + (1 if graph else 0) + steps)


def host_roundtrip(pkg, ch, dt, cv, steps):
    """ms per RK3 step of the per-substep drop-in (SOLVERRK3_MOM_tg with host arrays): each
    substep copies Q, PR, CONVH0 host -> HBM, runs, and copies them back (pageable host
    buffers allocated once, as the Fortran glue holds them)."""
    from chapsim_amd.tg import _p
    q, p, h = ch.download(with_history=True)
    lib = pkg.lib()
    ch.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        for ns in (1, 2, 3):
            if lib.chs_upload_fields(ch._h, _p(q), _p(p), _p(h)) != 0:
                raise RuntimeError("chs_upload_fields failed")
            ch.SOLVERRK3_MOM_tg(ns, dt, cv)
            if lib.chs_download_fields(ch._h, _p(q), _p(p), _p(h)) != 0:
                raise RuntimeError("chs_download_fields failed")
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps


def _monitor_cput(work, prefix):
    logs = [f for f in os.listdir(os.path.join(work, "0_log_monitors")) if f.startswith(prefix)]
    cput = []
    if logs:
        with open(os.path.join(work, "0_log_monitors", logs[0])) as fh:
            for line in fh:
                if line.lstrip().startswith("#") or not line.strip():
                    continue
                parts = line.split()
                if re.fullmatch(r"\d+", parts[0]):
                    cput.append(float(parts[2]))
    return cput


def host_cores(M: int, N: int):
    """Ranks for the reference on this host: its physical cores among the CPUs this process may
    use (/proc/cpuinfo physical id + core id), capped by the CPU share the environment states
    (OMP_NUM_THREADS, set to the share of one GPU on the GPU box) and by min(M, N), which the
    reference's y / z decomposition needs (BASELINE: P = physical cores, P <= min(M, N))."""
    aff = set(os.sched_getaffinity(0))
    phys, cur, seen = set(), {}, False
    try:
        with open("/proc/cpuinfo") as fh:
            for line in list(fh) + ["\n"]:
                if not line.strip():
                    if "processor" in cur and int(cur["processor"]) in aff:
                        phys.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                        seen = True
                    cur = {}
                elif ":" in line:
                    k, v = line.split(":", 1)
                    cur[k.strip()] = v.strip()
    except OSError:
        pass
    n_phys = len(phys) if seen else len(aff)
    share = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or None
    P = min(n_phys, share or n_phys, M, N)
    return max(2, P), dict(logical_cpus=len(aff), physical_cores=n_phys, cpu_share=share)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "host"


def cpu_baseline_heated(sample_steps: int = 3, cores: int | None = None, vis_scheme: int = 1):
    """Time the reference program on the heated main-domain workload: its own random start
    (INITIAL_FL_THEML_FLD_io), energy from the first step (timeThermoStart = 0), the given
    viscous scheme (iVisScheme), the same synthetic table; median per-step CpuT of the
    main-domain monitor.  The reference's compiled Poisson call for this domain stores its
    solution with a shifted layout (SOLVERRK3_MOM_IO.f90:93), so its field diverges within a
    few steps; the per-step cost of the work it does is still what is timed."""
    from chapsim_amd import thermo
    from oracle import refrun
    if not os.path.exists(refrun.PROGRAM) or not os.path.exists(refrun.MPIRUN):
        return None
    c = CFG_HEATED
    L, M, N = c["L"], c["M"], c["N"]
    P, hc = host_cores(M, N)
    P = cores or P

    def run(steps):
        work = tempfile.mkdtemp(prefix="chs_cpu_")
        try:
            ini = refrun.ini_text(L, M, N, HX=c["HX"], HZ=c["HZ"], STR2=c["STR2"], REN=c["REN"], DT=c["dt"],
                                  NTSTF=steps, iDomain=2, iThermoDynamics=1, T0=c["T0"], Ti=c["T0"],
                                  iThermalWallType=(1, 1), thermalWallBC=(2.0e4, 2.0e4), iVisScheme=vis_scheme,
                                  timeThermoStart=0.0)
            with open(os.path.join(work, "readdata.ini"), "w") as fh:
                fh.write(ini)
            thermo.write_table_file(os.path.join(work, "NIST_WATER_23.5MP.DAT"), thermo.synthetic_table())
            cmd = (f"ulimit -s unlimited 2>/dev/null || ulimit -s $(ulimit -H -s); "
                   f"{refrun.MPIRUN} -np {P} {refrun.PROGRAM}")
            t0 = time.perf_counter()
            res = subprocess.run(["bash", "-c", cmd], cwd=work, capture_output=True, text=True, timeout=900)
            return res, time.perf_counter() - t0, _monitor_cput(work, "monitor.io")
        finally:
            shutil.rmtree(work, ignore_errors=True)

    res, wall, cput = run(sample_steps)
    if res.returncode == 0 and not cput:
        # the reference's field diverges (see above); a run that stops before its monitor line
        # is timed again over its first step only
        res, wall, cput = run(1)
    if res.returncode != 0 or not cput:
        return dict(error=f"reference run failed rc={res.returncode}: {res.stderr[-400:]}")
    s = float(np.median(cput))
    return dict(value=L * M * N / s, unit="grid-point RK3-steps/s", cores=P, kind="reference", host=hc,
                sample=f"reference CHAPSim (oracle/_ref/chapsim, amdflang -O2) mpirun -np {P} on "
                       f"{_cpu_model()} ({hc['physical_cores']} physical cores usable, CPU share "
                       f"{hc['cpu_share']}); heated channel {L}x{M}x{N}, iVisScheme={vis_scheme}, "
                       f"{len(cput)} RK3 steps from its own random start; median per-step CpuT {s:.3f} s; "
                       f"run wall {wall:.1f} s")


def cpu_baseline_reference(start, sample_steps: int = 6, cores: int | None = None, c=None):
    """Time the reference program on this host: restart from the projected start field
    (the reference checks max div < 1e-6 on restart, FLOWSTART.f90:41-46) written in its
    restart format (processing/WRT_VARS_TG.f90:12-104), `sample_steps` steps, median of the
    per-step CpuT monitor column (processing/PP_SCN_MONITOR.f90:236, RK loop only, max over
    ranks)."""
    from oracle import refrun
    if not os.path.exists(refrun.PROGRAM) or not os.path.exists(refrun.MPIRUN):
        return None
    c = c or CFG
    L, M, N = c["L"], c["M"], c["N"]
    P, hc = host_cores(M, N)
    P = cores or P
    work = tempfile.mkdtemp(prefix="chs_cpu_")
    try:
        ini = refrun.ini_text(L, M, N, HX=c["HX"], HZ=c["HZ"], STR2=c["STR2"], REN=c["REN"],
                              DT=c["dt"], NTSTF=sample_steps, iIniField_tg=2, TimeReStart_tg=0.0)
        with open(os.path.join(work, "readdata.ini"), "w") as fh:
            fh.write(ini)
        os.makedirs(os.path.join(work, "1_instant_rawdata"))
        q, pr = start
        hdr_i = np.array([L, M, N, 0], dtype=np.int32)
        hdr_r = np.array([0.0, c["REN"], c["dt"]], dtype=np.float64)
        for name, arr in (("U", q[0]), ("V", q[1]), ("W", q[2]), ("P", pr)):
            fn = os.path.join(work, "1_instant_rawdata", f"DNS_perixz_INSTANT_T{0.0:.9E}_{name}.D")
            with open(fn, "wb") as fh:
                fh.write(hdr_i.tobytes())
                fh.write(hdr_r.tobytes())
                fh.write(np.ascontiguousarray(arr[:, 1:M + 1, :]).tobytes())
        cmd = f"ulimit -s unlimited 2>/dev/null || ulimit -s $(ulimit -H -s); " \
              f"{refrun.MPIRUN} -np {P} {refrun.PROGRAM}"
        t0 = time.perf_counter()
        res = subprocess.run(["bash", "-c", cmd], cwd=work, capture_output=True, text=True, timeout=600)
        wall = time.perf_counter() - t0
        cput = _monitor_cput(work, "monitor.tg")
        if res.returncode != 0 or not cput:
            return dict(error=f"reference run failed rc={res.returncode}: {res.stderr[-400:]}")
        s = float(np.median(cput))
        cpu_model = _cpu_model()
        return dict(value=L * M * N / s, unit="grid-point RK3-steps/s", cores=P, kind="reference", host=hc,
                    sample=f"reference CHAPSim (oracle/_ref/chapsim, amdflang -O2) mpirun -np {P} "
                           f"on {cpu_model or 'host'} ({hc['physical_cores']} physical cores usable, CPU "
                           f"share {hc['cpu_share']}, P = min(cores, share, M, N)); {L}x{M}x{N}, "
                           f"{len(cput)} RK3 steps from a restart field; median per-step CpuT {s:.3f} s; "
                           f"run wall {wall:.1f} s")
    finally:
        shutil.rmtree(work, ignore_errors=True)


def cpu_baseline_port():
    """C oracle (one core) on one substep of the same grid."""
    from oracle import oracle, setup
    c = CFG
    mesh = setup.Mesh(c["L"], c["M"], c["N"], HX=c["HX"], HZ=c["HZ"], STR2=c["STR2"])
    q, pr = oracle.initial_field(mesh, 395, 0.05)
    o = oracle.OracleTG(mesh, c["REN"])
    o.set_fields(q, pr)
    t0 = time.perf_counter()
    o.substep(1, c["dt"])
    s = 3 * (time.perf_counter() - t0)
    return dict(value=c["L"] * c["M"] * c["N"] / s, unit="grid-point RK3-steps/s", cores=1,
                kind="port", sample=f"C oracle, one substep at {c['L']}x{c['M']}x{c['N']}, x3")


def bench_poisson(pkg, args):
    """Standalone pressure Poisson (config 3): 512x384x512, beta 2.2, seeded Gaussian RHS with
    the dy-weighted mean removed.  Times `steps` solves of the plain FP64 form and of the
    mixed form with args.nref FP64 refinement sweeps; reports the mixed form's error vs the
    FP64 solution and both residuals."""
    L, M, N = 512, 384, 512
    co = pkg.channel_setup(L, M, N, 4 * np.pi, 2 * np.pi, 2.2)
    rng = np.random.Generator(np.random.Philox(key=3))
    r = rng.standard_normal((N, M, L))
    w = (1.0 / co["dyfi"])[None, :, None]
    r -= np.sum(r * w) / (np.sum(w) * L * N)
    d_r = pkg.DeviceArray.from_host(r)
    rmax = float(np.max(np.abs(r)))
    del r
    out = {}
    phis = {}
    for form, mode, nref in (("plain_fp64", pkg.PLAIN_FP64, 0), ("mixed", pkg.MIXED, args.nref)):
        ch = pkg.ChannelTG(co, mode)
        d_phi = pkg.DeviceArray(d_r.shape)
        for _ in range(max(1, args.warmup)):
            res = ch.poisson_device(d_r.ptr, d_phi.ptr, nref, True)
        ch.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ch.poisson_device(d_r.ptr, d_phi.ptr, nref, False)
        ch.sync()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        res = ch.poisson_device(d_r.ptr, d_phi.ptr, nref, True)
        phis[form] = d_phi.to_host()
        d_phi.free()
        out[form] = {"ms_per_solve": ms, "grid_point_solves_per_s": L * M * N / (ms * 1e-3),
                     "refinement_sweeps": nref, "max_residual_rel": res / rmax}
        ch.close()
    d_r.free()
    err = float(np.max(np.abs(phis["mixed"] - phis["plain_fp64"])) / np.max(np.abs(phis["plain_fp64"])))
    out["mixed"]["max_rel_error_vs_fp64"] = err
    print(json.dumps({"metric": "grid-point Poisson solves/s", "workload":
                      "standalone Poisson 512x384x512, tanh beta=2.2, periodic x,z, Neumann y",
                      "value": out["mixed"]["grid_point_solves_per_s"], "unit": "grid-point solves/s",
                      "n_gpus": 1, "steps": args.steps, "forms": out,
                      "speedup_mixed_vs_plain": out["plain_fp64"]["ms_per_solve"] / out["mixed"]["ms_per_solve"]}))


def bench_ensemble(pkg, args):
    """32 independent channels of seeded uneven sizes (config 4), one hipGraph per RK3 step
    of the whole ensemble; value = total grid-point RK3-steps/s."""
    from chapsim_amd.init import ensemble_cases
    cases = ensemble_cases()
    form = "plain" if args.form == "plain" else "mixed"     # the batched workload runs mixed
    mode = pkg.MIXED if form == "mixed" else pkg.PLAIN_FP64
    cos = [pkg.channel_setup(c["L"], c["M"], c["N"], c["HX"], c["HZ"], c["STR2"]) for c in cases]
    ens = pkg.Ensemble(cos, mode)
    ens.set_batched(args.ens_batched)
    dt = np.array([c["dt"] for c in cases])
    cv = np.array([1.0 / c["REN"] for c in cases])
    for i, (ch, co) in enumerate(zip(ens.channels, cos)):
        q, p = pkg.poiseuille_perturbed(co, seed=1000 + i, amp=0.1)
        ch.upload(q, p)
        ch.SOLVERRK3_MOM_tg(0, dt[i], cv[i])
    for _ in range(args.warmup):
        ens.step(dt, cv)
    ens.sync()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ens.step(dt, cv)
    ens.sync()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    pts = sum(c["L"] * c["M"] * c["N"] for c in cases)
    divs = [ch.max_div() for ch in ens.channels]
    ens.close()
    print(json.dumps({"metric": "grid-point RK3-steps/s (ensemble total)", "value": pts / (ms * 1e-3),
                      "unit": "grid-point RK3-steps/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": ms, "channels": len(cases),
                      "grid_points": pts, "form": form, "max_div": max(divs),
                      "launches": "batched: one per kernel kind and stage over all channels"
                                  if args.ens_batched else "per-channel, on capture lanes",
                      "config": {"workload": "32 channels, L,N in [64,192], M in [64,128], "
                                             "Re_tau U[180,395], one hipGraph per RK3 step"}}))


CFG_HEATED = dict(L=256, M=128, N=256, HX=4 * np.pi, HZ=2 * np.pi, STR2=1.8, REN=2800.0, dt=1.0e-3,
                  T0=665.0, wallflux=(0.115, -0.115))


def bench_heated(pkg, args):
    """Quasi-incompressible heated channel on the main domain (config 5): per RK3 stage the
    energy substep (table properties, iso-flux walls), the staggered properties and the
    variable-property momentum substep, Crank-Nicolson viscous scheme (iVisScheme = 1; --vis
    explicit for iVisScheme = 0), in SOLVE's order (chs_heated_substep).  --form both (default):
    the plain FP64 step and the mixed form (EXPLT0 and the momentum increment in FP32,
    chs_io_set_mixed), each from the same start, with the mixed fields' error against the
    plain ones; value = the first form's grid-point RK3-steps/s."""
    from chapsim_amd import init, thermo
    c = CFG_HEATED
    L, M, N = c["L"], c["M"], c["N"]
    t = thermo.table_setup(thermo.synthetic_table(), c["T0"])
    co = pkg.channel_setup(L, M, N, c["HX"], c["HZ"], c["STR2"])
    dh, g, pr = init.heated_channel_initial(co, t, seed=180)
    cthecd = 1.0 / c["REN"] / (t["M0"] * t["CP0"] / t["K0"])
    dhmin, dhmax = t["DH"][t["IMIN_DH"] - 1], t["DH"][t["IMAX_DH"] - 1]
    vis = 1 if args.vis == "cn" else 0
    io_shapes = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.set}
    dt = c["dt"]

    def run(mixed):
        eng = thermo.EnergyTG(co, t, (1, 1), cthecd, dhmin, dhmax, c["wallflux"])
        eng.upload(dh, g)
        f0 = eng.download()
        io = pkg.ChannelIO(co, c["REN"], flow_driven=1, vis_scheme=vis)
        eshapes = {k: v for k, v in io_shapes.items() if k.startswith("rhs_")}   # energy-substep keys
        if eshapes:
            eng.set_tuning(**eshapes)
        if len(eshapes) < len(io_shapes):
            io.set_tuning(**{k: v for k, v in io_shapes.items() if k not in eshapes})
        io.set_mixed(mixed)
        io.upload(G=g, PR=pr, D0=f0["D"], M0=f0["M"])

        def step():
            for ns in (1, 2, 3):
                pkg.heated_substep(eng, io, ns, dt)

        for _ in range(args.warmup):
            step()
        io.sync()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        io.sync()
        ms = (time.perf_counter() - t0) * 1e3 / args.steps
        f, fe = io.download(), eng.download()
        eng.close()
        io.close()
        return ms, f, fe

    forms = {"both": (False, True), "plain": (False,), "mixed": (True,)}[args.form]
    res = {mx: run(mx) for mx in forms}
    ms, f, fe = res[forms[0]]
    mixed_leg = None
    if True in res:
        msm, fm, fem = res[True]
        mixed_leg = {"ms_per_step": msm, "value": L * M * N / (msm * 1e-3)}
        if False in res:
            def rerr(a, b):
                return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
            mixed_leg["speedup_vs_plain"] = ms / msm
            mixed_leg["rel_err_vs_plain"] = {
                "u": rerr(fm["G"][0], f["G"][0]), "v": rerr(fm["G"][1], f["G"][1]),
                "w": rerr(fm["G"][2], f["G"][2]), "p": rerr(fm["PR"], f["PR"]), "T": rerr(fem["T"], fe["T"]),
                "after_steps": args.warmup + args.steps}
    # per-step algorithmic bytes, FP64 field passes per stage: Crank-Nicolson 138, the fixed
    # count of compulsory passes (energy 32: RHS + update 10, property update 7, d rho / dt 7,
    # rho / mu staggering 8; momentum 106: u = G / rho 9, explicit terms 23, x / z / y line
    # solves 12 + 12 + 15, divergence + forward transforms 9, y-Thomas 2, inverse transforms 4,
    # p / G / u updates 20); explicit 148, the kernel list of chs_io.hip / chs_thermo.hip
    alg = 3 * (138 if vis else 148) * 8.0 * L * M * N
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_heated(vis_scheme=vis)
        except Exception as e:  # pragma: no cover - host-dependent
            cpu = dict(error=str(e))
    print(json.dumps({
        "metric": "grid-point RK3-steps/s, heated channel (energy + variable-property momentum)",
        "value": L * M * N / (ms * 1e-3), "unit": "grid-point RK3-steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "dtype": "f64", "data": "synthetic: laminar G + 5% seeded noise, T/T0 = 1 + 1% seeded; "
                               "synthetic 2048-point table",
        "config": {"workload": f"heated channel REN={c['REN']:g}, {L}x{M}x{N}, 4pi x 2 x 2pi, tanh "
                               f"beta={c['STR2']:g}, dt={dt:g}, iso-flux walls, "
                               f"{'Crank-Nicolson' if vis else 'explicit'} viscous, RK3"},
        "step_roofline": {"alg_bytes_per_step": alg, "achieved_GBs": alg / (ms * 1e-3) / 1e9,
                          "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cpu_baseline": cpu,
        "form": "plain FP64" if not forms[0] else "mixed (EXPLT0 and momentum increment FP32)",
        "mixed": mixed_leg,
        "checks": {"T_range": [float(fe["T"][:, 1:M + 1].min()), float(fe["T"][:, 1:M + 1].max())],
                   "bulk_G": float(np.mean(f["G"][0][:, 1:M + 1] / np.asarray(co["dyfi"])[None, :, None]))},
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--form", choices=["both", "plain", "mixed"], default="both",
                    help="channel workloads: time the plain FP64 and the mixed form (value = plain), or one")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--vis", choices=["cn", "explicit"], default="cn",
                    help="heated workload's viscous scheme (iVisScheme 1 or 0)")
    ap.add_argument("--graph", action="store_true", help="replay each RK3 step from a hipGraph")
    ap.add_argument("--autotune", action="store_true",
                    help="search the kernels' launch shapes (chapsim_amd.tune) before timing")
    ap.add_argument("--ens-batched", action="store_true",
                    help="ensemble: one launch per kernel kind over all channels instead of each channel's "
                         "own launches on capture lanes")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="launch shape for chs_set_tuning (repeatable), e.g. --set conv_bx=4")
    ap.add_argument("--workload", choices=["channel", "channel64", "poisson", "ensemble", "heated"],
                    default="channel")
    ap.add_argument("--nref", type=int, default=1, help="refinement sweeps (poisson, mixed)")
    ap.add_argument("--host-roundtrip", type=int, default=0, metavar="STEPS",
                    help="also time STEPS steps of the per-substep host-array drop-in (PCIe included)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="only run a plain 512 MiB device copy (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_each_way": x.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (the workload is one channel on one "
              "MI355X); --gpus must be 1", file=sys.stderr)
        sys.exit(2)
    pkg = load_pkg()
    if args.workload == "poisson":
        return bench_poisson(pkg, args)
    if args.workload == "ensemble":
        return bench_ensemble(pkg, args)
    if args.workload == "heated":
        return bench_heated(pkg, args)
    c = CFG64 if args.workload == "channel64" else CFG
    shapes = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in args.set}
    start = start_field(pkg, c)
    forms = ["plain", "mixed"] if args.form == "both" else [args.form]
    res = {f: run_gpu(pkg, f, args.steps, args.warmup, c, args.graph, args.autotune, shapes,
                      args.host_roundtrip if f == forms[0] else 0, start) for f in forms}
    F = c["L"] * c["M"] * c["N"]

    def form_line(f):
        r = res[f]
        dom = r["dominant"]
        b = alg_bytes(dom, c["L"], c["M"], c["N"], r["wbytes"])
        achieved = b / (r["dom_avg_ms"] * 1e-3) / 1e9 if r["dom_avg_ms"] > 0 else 0.0
        traffic = None
        tpath = os.path.join(ROOT, "profiles", "traffic.json")
        if os.path.exists(tpath):
            with open(tpath) as fh:
                traffic = json.load(fh).get(f"{f}:{dom}")
        # per step: the y-solves count three components per substep however they are launched
        kernel_bytes = sum(alg_bytes(k, c["L"], c["M"], c["N"], r["wbytes"]) * (9 if k == "cn_y" and n else n)
                           for k, n in r["launches"].items())
        step_bytes = canonical_step_bytes(c["L"], c["M"], c["N"], r["wbytes"])
        return dict(
            value=F / (r["ms_per_step"] * 1e-3), ms_per_step=r["ms_per_step"],
            roofline={"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                      "alg_bytes_per_launch": b, "avg_launch_ms": r["dom_avg_ms"], "launches": r["dom_count"]},
            step_roofline={"alg_bytes_per_step": step_bytes,
                           "achieved_GBs": step_bytes / (r["ms_per_step"] * 1e-3) / 1e9,
                           "frac": step_bytes / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
                           "count": "48 field passes per substep (22 FP64, 26 work precision)",
                           "kernel_sum_bytes_per_step": kernel_bytes,
                           "kernel_sum_frac": kernel_bytes / (r["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS},
            per_kernel_ms_per_step=r["per_kernel_ms"],
            checks={"cfl_dt": r["cfl"], "max_div": r["max_div"], "monitor": r["monitor"]},
            tuning=r["tuning"] or "defaults")

    main_form = forms[0]
    head = form_line(main_form)
    cpu = None
    if not args.no_cpu_baseline:
        try:
            cpu = cpu_baseline_reference(res[main_form]["start"], c=c)
        except Exception as e:  # pragma: no cover - host-dependent
            cpu = dict(error=str(e))
        if cpu is None or "error" in cpu:
            err = cpu.get("error") if cpu else "reference build absent"
            cpu = cpu_baseline_port()
            cpu["note"] = f"reference run unavailable: {err}"
    out = {
        "metric": "grid-point RK3-steps/s, mixed vs plain FP64 at stated error; % of HBM peak",
        "value": head["value"],
        "unit": "grid-point RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_form == "plain" else "f32+f64",
        "data": "synthetic: Re_tau 180 field (64^3, 100 steps from laminar + 10% seeded noise) "
                "interpolated to the mesh, + 5% seeded (Philox) perturbation, projected by NS=0",
        "config": {"workload": (f"periodic channel REN={c['REN']:g}, {c['L']}x{c['M']}x{c['N']}, "
                                f"4pi x 2 x 2pi, tanh beta={c['STR2']:g}, dt={c['dt']:g}, "
                                "RK3+CN, flow only; a step = 3 substeps + CFL_tg of the new field"),
                   "graph": bool(args.graph),
                   "tuning": head["tuning"],
                   "form": "plain_fp64" if main_form == "plain" else "mixed",
                   "grid": [c["L"], c["M"], c["N"]]},
        "roofline": head["roofline"],
        "step_roofline": head["step_roofline"],
        "per_kernel_ms_per_step": head["per_kernel_ms_per_step"],
        "checks": head["checks"],
        "host_roundtrip_ms_per_step": res[main_form]["roundtrip_ms"],
        "cpu_baseline": cpu,
    }
    if len(forms) == 2:
        m = form_line("mixed")
        qp, pp = res["plain"]["final"]
        qm, pm = res["mixed"]["final"]
        M = c["M"]

        def rel(a, b):
            return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        ump, umm = qp[0][:, 1:M + 1].mean(axis=(0, 2)), qm[0][:, 1:M + 1].mean(axis=(0, 2))
        mp, mm = res["plain"]["monitor"], res["mixed"]["monitor"]
        rt = lambda mo: c["REN"] * 0.5 * (np.sqrt(abs(mo["cf_bottom"]) / 2) + np.sqrt(abs(mo["cf_top"]) / 2))
        m["speedup_vs_plain"] = head["ms_per_step"] / m["ms_per_step"]
        m["error_vs_plain"] = {
            "after_steps": res["mixed"]["steps_run"],
            "u": rel(qm[0], qp[0]), "v": rel(qm[1], qp[1]), "w": rel(qm[2], qp[2]), "p": rel(pm, pp),
            "mean_u_profile": rel(umm, ump), "re_tau": abs(rt(mm) / rt(mp) - 1.0),
            "max_div_mixed": res["mixed"]["max_div"], "max_div_plain": res["plain"]["max_div"]}
        out["mixed"] = m
    print(json.dumps(out))


if __name__ == "__main__":
    main()
