"""The reference program itself on one process with the GPU substep (oracle/_ref/chapsim_gpu:
every reference object unchanged, SOLVERRK3_MOM_tg, CFL_tg, DIVGCK_tg, the post-processing
hook and the y-halo exchange wrapped by chapsim1.0_amd/fortran/chs_tg_single_rank.f90), run as
a fresh child process next to the unmodified reference on two MPI ranks from the same restart
field and readdata.ini: the monitor lines and the instantaneous fields it writes agree."""
import glob
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import refrun

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GPU_PROGRAM = os.path.join(ROOT, "oracle", "_ref", "chapsim_gpu")


def monitor(work):
    logs = glob.glob(os.path.join(work, "0_log_monitors", "monitor.tg.*.log"))
    rows = []
    with open(logs[0]) as fh:
        for line in fh:
            parts = line.split()
            if parts and parts[0].isdigit():
                rows.append([float(x) for x in parts])
    return np.array(rows)


def run(cmd, work):
    res = subprocess.run(["bash", "-c", "ulimit -s unlimited 2>/dev/null || ulimit -s $(ulimit -H -s); " + cmd],
                         cwd=work, capture_output=True, text=True, timeout=300)
    log = "".join(open(f).read() for f in glob.glob(os.path.join(work, "0_log_monitors", "CHAPSim_Solver*")))
    assert res.returncode == 0 and "Error" not in log, (res.returncode, res.stdout[-2000:], log[-2000:])


@pytest.mark.skipif(not os.path.exists(GPU_PROGRAM) or not os.path.exists(refrun.PROGRAM),
                    reason="oracle/_ref not built (make -f oracle/ref.mk)")
def test_single_rank_gpu_program_matches_reference(pkg):
    L, M, N, STR2, REN, dt, nst = 16, 16, 16, 1.8, 2800.0, 1e-3, 4
    print(f"reference program: CPU on 2 ranks vs one process on the GPU, {L}x{M}x{N}, {nst} steps", flush=True)
    co = pkg.channel_setup(L, M, N, STR2=STR2)
    ch = pkg.ChannelTG(co, pkg.PLAIN_FP64)
    q, p = pkg.poiseuille_perturbed(co, seed=21, amp=0.1)
    ch.upload(q, p)
    ch.SOLVERRK3_MOM_tg(0, dt, 1.0 / REN)   # a divergence-free start (FLOWSTART.f90:41-46 checks it)
    q0, p0 = ch.download()
    ch.close()
    # every step a monitor line and (dtSave = DT / 2) an instantaneous field set
    ini = refrun.ini_text(L, M, N, STR2=STR2, REN=REN, DT=dt, NTSTF=nst, iIniField_tg=2, TimeReStart_tg=0.0,
                          dtSave1=dt / 2, iterMonitor=1)
    works = {}
    try:
        for name in ("cpu", "gpu"):
            w = tempfile.mkdtemp(prefix=f"chs_dropin_{name}_")
            works[name] = w
            with open(os.path.join(w, "readdata.ini"), "w") as fh:
                fh.write(ini)
            os.makedirs(os.path.join(w, "1_instant_rawdata"))
            pkg.restart_write_tg(os.path.join(w, "1_instant_rawdata"), q0, p0, 0, 0.0, REN, dt)
        run(f"{refrun.MPIRUN} -np 2 {refrun.PROGRAM}", works["cpu"])
        run(GPU_PROGRAM, works["gpu"])
        mc, mg = monitor(works["cpu"]), monitor(works["gpu"])
        assert mc.shape == mg.shape and mc.shape[0] == nst
        # columns: 1 STEP, 2 PhyT, 3 CpuT (not compared), 4 CFL, 5 DT, 6 MaxDiv, 7.. statistics
        np.testing.assert_array_equal(mc[:, [0, 1, 4]], mg[:, [0, 1, 4]])
        assert np.all(mg[:, 5] < 1e-13) and np.all(mc[:, 5] < 1e-13)
        stats_c, stats_g = mc[:, 6:], mg[:, 6:]
        scale = np.maximum(np.abs(stats_c), 1e-12)
        e = np.max(np.abs(stats_g - stats_c) / scale)
        print(f"  monitor statistics (printed to 6 digits) max relative difference {e:.1e}")
        assert e < 2e-5
        np.testing.assert_allclose(mg[:, 3], mc[:, 3], rtol=1e-3)   # CFL, printed to 3 decimals
        # the last instantaneous field written by each program
        t_end = nst * dt
        fc = pkg.restart_read_tg(os.path.join(works["cpu"], "1_instant_rawdata"), t_end, L, M, N)
        fg = pkg.restart_read_tg(os.path.join(works["gpu"], "1_instant_rawdata"), t_end, L, M, N)
        eq = max(np.max(np.abs(fg[0][c] - fc[0][c])) / np.max(np.abs(fc[0][c])) for c in range(3))
        ep = np.max(np.abs(fg[1] - fc[1])) / np.max(np.abs(fc[1]))
        print(f"  fields at t = {t_end}: velocity {eq:.2e}, pressure {ep:.2e} (relative max-norm)")
        assert fg[2] == fc[2] and abs(fg[3] - fc[3]) < 1e-15
        assert eq < 1e-10 and ep < 1e-10
    finally:
        for w in works.values():
            shutil.rmtree(w, ignore_errors=True)
