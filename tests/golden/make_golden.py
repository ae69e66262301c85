"""Writes the golden vectors under tests/golden/ from the reference's own FP64
CPU path (oracle/_ref/<cfg>/ref_harness, compiled from the unmodified
reference sources by oracle/build_ref.py).

Run in a container that has /root/reference:
    python tests/golden/make_golden.py [t4 t4o c1]

Each fixture holds the seeded inputs, the normalised rational coefficients
(from the project's Remez), the spectral bounds used for rescaling and the
reference outputs.  Large c1 fields are stored as the reference outputs at a
fixed subset of sites plus full spinors and Delta H.
"""
from __future__ import annotations

import importlib.util
import gzip
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(ROOT, "rhmc-on-gpu_amd"))

from rhmc_amd import fields  # noqa: E402
from rhmc_amd.remez import rhmc_approximations  # noqa: E402


def _load(name, path):
    spec = importlib.util.spec_from_file_location(name, path)
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


build_ref = _load("build_ref", os.path.join(ROOT, "oracle", "build_ref.py"))

SEEDS = {"links": 101, "mom": 202, "phi": 303, "in": 404, "ipdot": 505, "chi": 606}
LINK_EPS = {"t4": None, "t4o": 0.3, "t4m": 0.3, "c1": None}


def write_approx(d, eig_min, eig_max, apps):
    with open(os.path.join(d, "approx.txt"), "w") as fh:
        fh.write(f"{eig_min!r} {eig_max!r}\n")
        for a0, a, b in apps:
            fh.write(f"{len(a)} {a0!r}\n")
            fh.write(" ".join(repr(float(x)) for x in a) + "\n")
            fh.write(" ".join(repr(float(x)) for x in b) + "\n")


def run(exe, d, op, arg=None):
    cmd = [exe, d, op] + ([repr(arg)] if arg is not None else [])
    subprocess.check_call(cmd, stdout=subprocess.DEVNULL)


def rd(d, name, shape, dtype=np.complex128):
    return np.fromfile(os.path.join(d, name), dtype=dtype).reshape(shape)


def make(cfg_name):
    cfg = build_ref.CONFIGS[cfg_name]
    exe = build_ref.build_ref(cfg_name)
    if exe is None:
        raise SystemExit("reference not present")
    dims = (cfg["nx"], cfg["ny"], cfg["nz"], cfg["nt"])
    lat = fields.Lattice(*dims)
    V, Vh, npf = lat.V, lat.Vh, cfg["no_ps"]
    U = fields.random_links(lat, SEEDS["links"], LINK_EPS[cfg_name])
    P = fields.gauss_momenta(lat, SEEDS["mom"])
    phi = fields.gauss_spinor(lat, SEEDS["phi"], npf)
    chi = fields.gauss_spinor(lat, SEEDS["chi"], npf)
    in_e = fields.gauss_spinor(lat, SEEDS["in"], 2)
    ip = fields.gauss_momenta(lat, SEEDS["ipdot"]) * 1j  # anti-hermitian traceless
    first, md, last = rhmc_approximations(cfg["no_flavours"], npf, cfg["approx_metro"], cfg["approx_md"],
                                          cfg["approx_metro"], 1e-5)
    apps = [(x.a0, x.a, x.b) for x in (first, md, last)]
    mass = cfg["mass"]
    eig_min = mass * mass
    eig_max = 4.0 + mass * mass  # upper bound of M^dag M (|D_eo D_oe| <= 4)
    out = dict(dims=np.array(dims), mass=mass, beta=cfg["beta"], no_ps=npf, no_md=cfg["no_md"],
               use_multistep=cfg["use_multistep"], gauge_scale=cfg["gauge_scale"],
               n_flavours=cfg["no_flavours"], res_metro=cfg["residue_metro"], res_md=cfg["residue_md"],
               eig_min=eig_min, eig_max=eig_max, link_seed=SEEDS["links"],
               link_eps=-1.0 if LINK_EPS[cfg_name] is None else LINK_EPS[cfg_name])
    for nm, (a0, a, b) in zip(("first", "md", "last"), apps):
        out[nm + "_a0"] = a0
        out[nm + "_a"] = np.array(a)
        out[nm + "_b"] = np.array(b)
    small = V <= 512
    with tempfile.TemporaryDirectory() as d:
        U.tofile(os.path.join(d, "links.bin"))
        P.tofile(os.path.join(d, "mom.bin"))
        phi.tofile(os.path.join(d, "phi.bin"))
        chi.tofile(os.path.join(d, "chi.bin"))
        ip.tofile(os.path.join(d, "ipdot.bin"))
        in_e[0].tofile(os.path.join(d, "in_e.bin"))
        in_e[1].tofile(os.path.join(d, "in_o.bin"))
        write_approx(d, eig_min, eig_max, apps)
        run(exe, d, "dslash")
        out["dslash_in"] = in_e
        out["doe"] = rd(d, "out_doe.bin", (Vh, 3))
        out["deo"] = rd(d, "out_deo.bin", (Vh, 3))
        # multi-shift with the rescaled MD shifts, used verbatim
        a0r, ar, br, _ = md.rescaled(eig_min, eig_max, -cfg["no_flavours"] / 4.0 / npf)
        write_approx(d, eig_min, eig_max, [apps[0], (a0r, ar, br), apps[2]])
        res_ms = 1e-11 if small else 1e-10
        run(exe, d, "multishift", res_ms)
        out["ms_res"] = res_ms
        out["ms_shifts"] = np.array(br)
        out["ms_x"] = np.stack([np.stack([rd(d, f"out_x_{p}_{k}.bin", (Vh, 3)) for k in range(len(br))])
                                for p in range(npf)])
        write_approx(d, eig_min, eig_max, apps)
        if small:
            for op in ("force_fermion", "force_full", "force_gauge"):
                run(exe, d, op)
                out[op] = rd(d, "out_ipdot.bin", (4 * V, 3, 3))
            run(exe, d, "exp", 0.137)
            out["exp_step"] = 0.137
            out["exp_links"] = rd(d, "out_links.bin", (4 * V, 3, 3))
            run(exe, d, "mom", -0.071)
            out["mom_step"] = -0.071
            out["mom_out"] = rd(d, "out_mom.bin", (4 * V, 3, 3))
            run(exe, d, "action")
            out["action0"] = rd(d, "out_action0.bin", (V,), np.float64)
            out["action1"] = rd(d, "out_action1.bin", (V,), np.float64)
        run(exe, d, "meas_gauge")
        out["meas_gauge"] = np.fromfile(os.path.join(d, "out_meas.bin"), np.float64)
        run(exe, d, "meas_chiral")
        ch = np.fromfile(os.path.join(d, "out_chiral.bin"), np.float64)
        out["meas_chiral"] = ch[0::2] + 1j * ch[1::2]
        out["meas_rnd_e"] = rd(d, "out_rnd_e.bin", (Vh, 3))
        out["meas_rnd_o"] = rd(d, "out_rnd_o.bin", (Vh, 3))
        if cfg_name == "t4":
            # the reference's text configuration writer and reader
            run(exe, d, "conf_write", 7)
            with open(os.path.join(d, "config"), "rb") as fh:
                txt = fh.read()
            with open(os.path.join(HERE, "config_t4.txt.gz"), "wb") as raw, \
                    gzip.GzipFile(fileobj=raw, mode="wb", mtime=0) as fh:
                fh.write(txt)
            run(exe, d, "conf_read")
            out["conf_read_links"] = rd(d, "out_links.bin", (4 * V, 3, 3))
        run(exe, d, "traj")
        out["traj_dh"] = rd(d, "out_dh.bin", (1,), np.float64)[0]
        out["traj_chi"] = rd(d, "out_chi.bin", (npf, Vh, 3))
        out["traj_phinew"] = rd(d, "out_phinew.bin", (npf, Vh, 3))
        tl = rd(d, "out_links.bin", (4 * V, 3, 3))
        tm = rd(d, "out_mom.bin", (4 * V, 3, 3))
        if small:
            out["traj_links"], out["traj_mom"] = tl, tm
            out["links"], out["mom"], out["phi"], out["chi"], out["ipdot"] = U, P, phi, chi, ip
        else:
            sel = np.arange(0, 4 * V, 97)
            out["traj_sel"] = sel
            out["traj_links_sel"], out["traj_mom_sel"] = tl[sel], tm[sel]
            # inputs are regenerated from the seeds; record a checksum of them
            out["input_checksum"] = np.array([np.abs(U).sum(), np.abs(P).sum(), np.abs(phi).sum()])
    path = os.path.join(HERE, f"golden_{cfg_name}.npz")
    np.savez_compressed(path, **out)
    print("wrote", path, os.path.getsize(path))


if __name__ == "__main__":
    for name in sys.argv[1:] or ["t4", "t4o", "c1"]:
        make(name)
