"""Writes config1_dp45_1ns.npz: the FP64 oracle's run of the plain small
relaxation (fullmag_amd.workloads.texture_relax: 200 x 50 x 1 cells,
alpha 0.5, DP45 rtol 1e-6) to 1 ns with the reference runner's loop
(until-clamp, dt carried; crates/fullmag-runner/src/cpu_reference.rs:
448-455,622-639), started at dt 1e-13.  The run takes a few minutes on the
CPU, so the GPU test compares with this file instead of rerunning it.

    python tests/golden/make_config1.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from fullmag_amd.workloads import texture_relax  # noqa: E402
from oracle.oracle import Oracle, Problem  # noqa: E402

T_END = 1e-9
DT0 = 1e-13


def main():
    kw = texture_relax()
    pb = Problem(kw["nx"], kw["ny"], kw["nz"], kw["dx"], kw["dy"], kw["dz"], kw["ms"], kw["a_ex"], kw["alpha"],
                 h_ext=kw["h_ext"], atol=kw["atol"], rtol=kw["rtol"], dt_min=kw["dt_min"], dt_max=kw["dt_max"],
                 headroom=kw["headroom"])
    o = Oracle(pb)
    m = o.prepare(kw["m0"])
    t0 = time.time()
    r = o.run("rk45", m, 0.0, T_END, DT0, 10 ** 7, trace=True)
    ob = o.observables(m)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "config1_dp45_1ns.npz")
    np.savez_compressed(out, m=m, trace=r["trace"], t=r["t"], steps=r["steps"],
                        energies=np.array([ob["e_ex"], ob["e_demag"], ob["e_ext"], ob["e_total"]]),
                        max_rhs=ob["max_rhs"])
    print(f"{r['steps']} steps to t = {r['t']:.6e} s in {time.time() - t0:.1f} s -> {out}")


if __name__ == "__main__":
    main()
