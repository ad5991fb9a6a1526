"""The Fortran host run whole on the CPU: volcanor_host.f90 linked against
oracle/vlm_cpu_shim.c, the libvlm C ABI restated over the oracle
(make -C oracle hostcpu -> oracle/_cpu/libvolcanor_host_cpu.so).  This
checks host-side logic (namelists, cursors, slow start, tip plots, restart
records) and the oracle's restatement of whole cases with no GPU: the CL/CT
history and tip plots of tip_rotor against the reference solver's own
outputs, a restart at step 24 of the same case against the uninterrupted
run, and the prescribed far wake against the reference solver."""
import ctypes as C
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from volcanor_amd.host import Stats

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CASES = os.path.join(ROOT, "cases")
CPU_HOST = os.path.join(ROOT, "oracle", "_cpu", "libvolcanor_host_cpu.so")


@pytest.fixture(scope="module")
def cpu_host():
    """The CPU build of the host (rebuilt when a source is newer)."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "hostcpu"],
                       capture_output=True, text=True, timeout=300)
    if r.returncode != 0:
        pytest.fail("make -C oracle hostcpu failed:\n" + r.stdout + r.stderr)
    lib = C.CDLL(CPU_HOST)
    lib.volcanor_run.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int,
                                 C.c_int, C.POINTER(Stats), C.c_int, C.c_int, C.c_double]
    lib.volcanor_run.restype = C.c_int
    lib.volcanor_last_error.restype = C.c_char_p
    return lib


def run_cpu(lib, casedir, max_steps=2000):
    """volcanor_run on the CPU build; (status, CL/CT history, message)."""
    hist = np.zeros(max_steps + 1)
    st = Stats()
    rc = lib.volcanor_run(os.path.abspath(casedir).encode(), 0, 0, 0,
                          hist.ctypes.data_as(C.c_void_p), 1, max_steps + 1, C.byref(st), 0, 0,
                          1e-4)
    return rc, hist[: st.nt + 1].copy(), lib.volcanor_last_error().decode()


def case_copy(name, edits=()):
    d = tempfile.mkdtemp(prefix="vlmcpu_")
    src = os.path.join(CASES, name + ".case")
    for f in os.listdir(src):
        if os.path.isfile(os.path.join(src, f)):
            shutil.copy(os.path.join(src, f), d)
    cfg = os.path.join(d, "config.nml")
    text = open(cfg).read()
    for old, new in edits:
        assert old in text
        text = text.replace(old, new)
    with open(cfg, "w") as fh:
        fh.write(text)
    return d


def read_force(path):
    rows = [l.split() for l in open(path) if l.strip() and not l.lstrip().startswith("iter")]
    return np.array([[float(x) for x in r[1:]] for r in rows])


def tec_numbers(path):
    out = []
    for line in open(path):
        w = line.strip()
        if w and not w.startswith(("Title", "VARIABLES", "DATAPACKING", "VARLOCATION", "Zone")):
            out.extend(float(x) for x in w.split())
    return np.array(out)


def test_cpu_host_tip_rotor_matches_reference_solver(cpu_host):
    """tip_rotor (AM2, slow start over 20 steps, far wake from step 12, tip
    plots every 12 steps) on the CPU build: the CL/CT history and the tip
    plots of steps 12-48 against the reference solver's (the bounds of the
    GPU test of the same case)."""
    d = case_copy("tip_rotor")
    try:
        rc, ct, msg = run_cpu(cpu_host, d)
        assert rc == 0, msg
        ref = read_force(os.path.join(GOLD, "tip_rotor_ForceNonDim.csv"))[:, 0]
        np.testing.assert_allclose(ct, ref, rtol=2e-6, atol=1e-9)
        for step in ("00012", "00024", "00036", "00048"):
            got = tec_numbers(os.path.join(d, "Results", f"r01tip{step}.tec"))
            exp = tec_numbers(os.path.join(GOLD, f"tip_rotor_r01tip{step}.tec"))
            np.testing.assert_allclose(got, exp, rtol=1e-10, atol=1e-12, err_msg=step)
    finally:
        shutil.rmtree(d, ignore_errors=True)


def test_cpu_restart_with_slow_start_and_tip_plots(cpu_host):
    """tip_rotor with restart records every 24 steps, and a copy resumed from
    step 24 in the same process: the same CL/CT from step 25 and byte-equal
    tip files at steps 36 and 48 (the azimuthal ages and the slow-start
    rotation rate travel in the record)."""
    full = case_copy("tip_rotor", [("restartWriteNt = 0", "restartWriteNt = 24")])
    resumed = case_copy("tip_rotor", [("restartFromNt = 0", "restartFromNt = 24")])
    try:
        rc, ct_full, msg = run_cpu(cpu_host, full)
        assert rc == 0, msg
        os.makedirs(os.path.join(resumed, "Restart"))
        shutil.copy(os.path.join(full, "Restart", "restart00024.dat"),
                    os.path.join(resumed, "Restart"))
        rc, ct_res, msg = run_cpu(cpu_host, resumed)
        assert rc == 0, msg
        np.testing.assert_array_equal(ct_res[25:], ct_full[25:])
        for step in ("00036", "00048"):
            for f in (f"r01tip{step}.tec", f"r01tipXYZ{step}.csv"):
                a = open(os.path.join(full, "Results", f)).read()
                b = open(os.path.join(resumed, "Results", f)).read()
                assert a == b, f
    finally:
        shutil.rmtree(full, ignore_errors=True)
        shutil.rmtree(resumed, ignore_errors=True)


def test_cpu_emptied_config_returns_status_and_next_run_works(cpu_host):
    """An empty config.nml (what a caller leaves by truncating the file
    before reading it) is an error status with the runtime's message, not a
    stopped process; the next run in the same process completes."""
    d = case_copy("tip_rotor")
    try:
        open(os.path.join(d, "config.nml"), "w").close()
        rc, _, msg = run_cpu(cpu_host, d)
        assert rc == 1 and "config.nml" in msg and "End of file" in msg, msg
        shutil.copy(os.path.join(CASES, "tip_rotor.case", "config.nml"), d)
        rc, ct, msg = run_cpu(cpu_host, d)
        assert rc == 0 and len(ct) == 49, msg
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name", ["presc_rotor", "presc_gen_rotor"])
def test_cpu_prescribed_far_wake_matches_reference_solver(cpu_host, name):
    """The prescribed far wake (pFwake_update, classdef.f90:998-1066,
    5172-5217) on the CPU build: a climbing rotor whose 240-filament helix
    is rebuilt after every convection from step 31, from every far
    filament (presc_rotor) or, with three axisymmetric blades, from the
    last six (presc_gen_rotor).  The CL/CT history against the reference
    solver's, and every zone of the wing and wake plots at steps 24 and 48,
    the helix (PrescFarWake) included, to 1e-8 relative or 1e-10 absolute:
    the helix switches on with a jump in CT (0.0117 to -0.018 at step 32
    for presc_rotor), after which the wake amplifies round-off differences
    to about 4e-9 relative by step 48."""
    from tests.test_host_gpu import read_tec
    d = case_copy(name)
    try:
        rc, ct, msg = run_cpu(cpu_host, d)
        assert rc == 0, msg
        ref = read_force(os.path.join(GOLD, f"{name}_ForceNonDim.csv"))[:, 0]
        np.testing.assert_allclose(ct, ref, rtol=2e-6, atol=1e-9)
        for step in ("00024", "00048"):
            got = read_tec(os.path.join(d, "Results", f"r01wingNwake{step}.tec"))
            exp = read_tec(os.path.join(GOLD, f"{name}_wingNwake{step}.tec"))
            assert [z[:3] for z in got] == [z[:3] for z in exp]
            for zg, ze in zip(got, exp):
                for k in range(3, 8):
                    np.testing.assert_allclose(zg[k], ze[k], rtol=1e-8, atol=1e-10,
                                               err_msg=f"{step} {zg[0]} column {k - 2}")
        assert any(z[0] == "PrescFarWake" and z[1] == 241 for z in got)
    finally:
        shutil.rmtree(d, ignore_errors=True)


@pytest.mark.parametrize("name,nr,tips", [("image_rotor", 2, False), ("image_rotor_xy", 3, True)])
def test_cpu_image_rotors_match_reference_solver(cpu_host, name, nr, tips):
    """Image rotors (rotor_init classdef.f90:2798-2960, rotor_mirrorVelCP /
    rotor_mirrorGamma / rotor_mirrorWake 4990-5131) on the CPU build: an
    axisymmetric rotor in ground effect and its image in z = 0 (AM2), and a
    rotor with cyclic pitch and its images in x = 0 and y = 0 (AB2).  Every
    column of every rotor's force file, and every zone of every rotor's
    wing and wake plots (and tip plots) at steps 24 and 48, against the
    reference solver's, to 1e-8 relative or 1e-10 absolute (measured: 1.1e-9
    at step 48)."""
    from tests.test_host_gpu import read_tec
    d = case_copy(name)
    try:
        rc, _, msg = run_cpu(cpu_host, d)
        assert rc == 0, msg
        for k in range(1, nr + 1):
            got = read_force(os.path.join(d, "Results", f"r{k:02d}ForceNonDim.csv"))
            ref = read_force(os.path.join(GOLD, f"{name}_r{k:02d}ForceNonDim.csv"))
            np.testing.assert_allclose(got, ref, rtol=2e-6, atol=1e-9, err_msg=f"r{k:02d}")
            for step in ("00024", "00048"):
                for p in ["wingNwake"] + (["tip"] if tips else []):
                    g = read_tec(os.path.join(d, "Results", f"r{k:02d}{p}{step}.tec"))
                    e = read_tec(os.path.join(GOLD, f"{name}_r{k:02d}{p}{step}.tec"))
                    assert [z[:3] for z in g] == [z[:3] for z in e]
                    for zg, ze in zip(g, e):
                        for c in range(3, 8):
                            np.testing.assert_allclose(
                                zg[c], ze[c], rtol=1e-8, atol=1e-10,
                                err_msg=f"r{k:02d} {p}{step} {zg[0]} column {c - 2}")
    finally:
        shutil.rmtree(d, ignore_errors=True)
