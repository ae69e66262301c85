"""Drop-in check: the reference's own algorithm layer (lib/Update,
lib/Action, lib/Main/Gpu, packers, compiled unmodified with -DUSE_GPU) links
against librhmc_amd.so through compat/reference_shim.cc, and its update()
runs on the GPU with every device call served by this library."""
import importlib.util
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))
GOLD = os.path.join(ROOT, "tests", "golden")
EXE = os.path.join(ROOT, "oracle", "_ref", "{}_gpu", "ref_harness_gpu")


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


@pytest.mark.skipif(not os.path.isdir("/root/reference"), reason="reference sources not present")
def test_reference_links_against_library():
    bc = _load("build_compat", os.path.join(ROOT, "rhmc-on-gpu_amd", "compat", "build_compat.py"))
    exe = bc.build("t4", force=True)
    assert exe and os.path.exists(exe)
    # no undefined reference symbols remain (linked with --no-undefined); the
    # lib/Cuda entry points come from the shim, the rest from librhmc_amd.so
    und = subprocess.run(["nm", "-u", exe], capture_output=True, text=True).stdout
    assert "cuda_" not in und
    dyn = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True, text=True).stdout
    for sym in ("gpu_create", "gpu_init0", "gpu_shifted_inverter", "gpu_fermion_force", "gpu_sol_sum"):
        assert sym in dyn, sym


def _write_inputs(d, g):
    mg = _load("make_golden", os.path.join(GOLD, "make_golden.py"))
    g["links"].tofile(os.path.join(d, "links.bin"))
    g["mom"].tofile(os.path.join(d, "mom.bin"))
    g["phi"].tofile(os.path.join(d, "phi.bin"))
    apps = [(float(g[n + "_a0"]), g[n + "_a"], g[n + "_b"]) for n in ("first", "md", "last")]
    mg.write_approx(d, float(g["eig_min"]), float(g["eig_max"]), apps)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["t4", "t4o"])
def test_reference_update_on_gpu(name):
    exe = EXE.format(name)
    if not os.path.exists(exe):
        pytest.skip(f"{exe} not built (needs the reference sources at build time)")
    g = dict(np.load(os.path.join(GOLD, f"golden_{name}.npz")))
    with tempfile.TemporaryDirectory() as d:
        _write_inputs(d, g)
        print("running reference update() through the shim:", name, flush=True)
        subprocess.run([exe, d, "traj"], check=True, timeout=120, stdout=subprocess.DEVNULL)
        dh = float(np.fromfile(os.path.join(d, "out_dh.bin"), np.float64)[0])
        V = int(np.prod(g["dims"]))
        links = np.fromfile(os.path.join(d, "out_links.bin"), np.complex128).reshape(4 * V, 3, 3)
    # The reference's own boundary rounds P and ipdot to FP32 and the links to
    # hi/lo float pairs (packer.cc:8-64,443-569), so the GPU
    # build of the reference agrees with its CPU build to that precision.
    ref = float(g["traj_dh"])
    assert abs(dh - ref) < 1e-4 * max(1.0, abs(ref)), (dh, ref)
    assert np.linalg.norm(links - g["traj_links"]) / np.linalg.norm(g["traj_links"]) < 1e-5
