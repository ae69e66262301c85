"""A Markov chain of RHMC trajectories with the reference's checkpoint and
restart rules (lib/Main/rhmc.cc:64-150, lib/Update/update.cc:58-152):

* start from the stored configuration "config" in the work directory when it
  exists (Conf(init=2), continuing its update_iteration), else from the
  given links;
* trajectories with Metropolis accept/reject once update_iteration reaches
  therm_updates; before that, thermalisation updates without the test;
* every save_conf_every iterations the configuration is written, alternating
  between config_0 and config_1 (Conf::write), except after the last
  iteration of the run, which writes config (Conf::write_last);
* a file named "stop" in the work directory ends the run after the current
  trajectory and writes config;
* when the run ends on an iteration that is not a multiple of
  save_conf_every, config is written as well.

All fields stay on the device; configurations go through gpu_write_conf /
gpu_read_conf in the reference's text format.
"""
from __future__ import annotations

import os
import time

import numpy as np


class Chain:
    def __init__(self, gpu, workdir, beta, mass, n_flavours, links=None, therm_updates=0, rng_seed=0):
        self.g, self.dir = gpu, workdir
        self.beta, self.mass, self.nf = beta, mass, n_flavours
        self.therm = therm_updates
        self.rng = np.random.default_rng(rng_seed)
        self.conf_count = 0
        os.makedirs(workdir, exist_ok=True)
        stored = os.path.join(workdir, "config")
        if os.path.exists(stored):
            self.update_iteration, _, _, _ = gpu.read_conf(stored)
            self.resumed = True
        else:
            if links is None:
                raise ValueError("no stored configuration and no start links")
            gpu.init0(links)
            self.update_iteration = 0
            self.resumed = False
        self.log = []

    def _write(self, last):
        if last:
            name = "config"
        else:
            name = f"config_{self.conf_count}"
            self.conf_count = 1 - self.conf_count
        self.g.write_conf(os.path.join(self.dir, name), self.beta, self.mass, self.nf, self.update_iteration)
        return name

    def run(self, run_update_iterations, save_conf_every, traj_kwargs, seed0=1):
        start = self.update_iteration
        while self.update_iteration - start < run_update_iterations:
            therm = self.update_iteration < self.therm
            u = float(self.rng.random())
            t0 = time.perf_counter()
            r = self.g.trajectory(seed=seed0 + self.update_iteration, metropolis=not therm, metropolis_u=u,
                                  **traj_kwargs)
            rec = {"iteration": self.update_iteration, "delta_h": r.delta_h,
                   "accepted": None if therm else bool(r.accepted), "s": time.perf_counter() - t0,
                   "cg_iterations": r.cg_iterations}
            self.update_iteration += 1
            done = self.update_iteration - start == run_update_iterations
            written = []
            if self.update_iteration % save_conf_every == 0:
                written.append(self._write(last=done))
            stop = os.path.exists(os.path.join(self.dir, "stop"))
            if stop:
                written.append(self._write(last=True))
            if self.update_iteration % save_conf_every != 0 and done:
                written.append(self._write(last=True))
            rec["written"] = written
            self.log.append(rec)
            if stop:
                break
        return self.log

    def acceptance(self):
        acc = [r["accepted"] for r in self.log if r["accepted"] is not None]
        return float(np.mean(acc)) if acc else None
