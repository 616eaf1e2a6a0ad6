"""Adjoint gradients of the full-size waveguide workload: 256 x 64 x 2 cells
of 5 nm, a 64 x 64 design Ms map, a 5 mT sinc pulse (20 GHz) on the first 8
columns, RK4 with dt = 50 fs for 20,000 steps, J = mean m_z^2 over the last
8 columns in the last 0.5 ns.

Pinned by central differences (h = 1e-3 relative, the workload's own step)
on the pulse amplitude and one seeded design cell: the FP64 adjoint agrees
with those of the GPU FP64 forward objective to 1e-5 of each value (FD
truncation is O(h^2)).  The FP32 adjoint (field cotangent convolved in
FP64) is compared with the FP64 differences: it carries the FP32 forward's
own error (measured: J differs by 7.8e-4 relative after the 20,000 steps,
and both gradients by 8.9e-4, the same offset), bounded here by 2e-3.
Central differences of the FP32 forward itself are no reference at this
length: FP32 roundoff in J over 20,000 steps is of the size of the
h = 1e-3 perturbation's effect, so they scatter by tens of per cent.  A
different checkpoint interval gives the same FP64 gradient to 1e-12 (the
recompute replays the forward's kernels)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

H = 1e-3
AMP = 5e-3
FP32_TOL = 2e-3


def _problem(precision, interval=0):
    from fullmag_amd.adjoint import WaveguideProblem
    return WaveguideProblem(precision=precision, checkpoint_interval=interval)


def _central_differences(pr, design, cell):
    iy, ix = cell
    fd = []
    for kind in ("cell", "amp"):
        vals = []
        for sgn in (1.0, -1.0):
            d2, a2 = design.copy(), AMP
            if kind == "cell":
                d2[iy, ix] *= 1.0 + sgn * H
            else:
                a2 = AMP * (1.0 + sgn * H)
            jv, be = pr.run(d2, a2)
            be.close()
            vals.append(jv)
        base = design[iy, ix] if kind == "cell" else AMP
        fd.append((vals[0] - vals[1]) / (2.0 * H * base))
    return np.array(fd)


@pytest.fixture(scope="module")
def reference():
    pr = _problem("double")
    rng = np.random.default_rng(2024)
    design = rng.uniform(6e5, 1e6, size=pr.design_shape())
    dy, dx = pr.design_shape()
    iy, ix = int(rng.integers(dy)), int(rng.integers(dx))
    j, gd, ga = pr.value_and_grad(design, AMP)
    fd = _central_differences(pr, design, (iy, ix))
    print(f"waveguide J {j:.6e}; fd cell ({iy},{ix}) {fd[0]:.6e}, amplitude {fd[1]:.6e}", flush=True)
    return {"design": design, "cell": (iy, ix), "j": j, "gd": gd, "ga": ga, "fd": fd}


def test_fp64_adjoint_matches_central_differences(reference):
    r = reference
    iy, ix = r["cell"]
    adj = np.array([r["gd"][iy, ix], r["ga"]])
    rel = np.abs(adj - r["fd"]) / np.abs(r["fd"])
    print(f"fp64 adjoint {adj}, fd {r['fd']}, rel {rel}", flush=True)
    assert np.all(np.abs(r["fd"]) > 0)
    assert rel.max() <= 1e-5


def test_fp32_adjoint_matches_central_differences(reference):
    r = reference
    iy, ix = r["cell"]
    pr = _problem("single")
    j, gd, ga = pr.value_and_grad(r["design"], AMP)
    adj = np.array([gd[iy, ix], ga])
    rel64 = np.abs(adj - r["fd"]) / np.abs(r["fd"])
    print(f"fp32 adjoint {adj}; fp64 fd {r['fd']}, rel {rel64}; J fp32 {j:.6e} vs fp64 {r['j']:.6e} "
          f"(rel {abs(j - r['j']) / abs(r['j']):.3e})", flush=True)
    assert abs(j - r["j"]) <= FP32_TOL * abs(r["j"])
    assert rel64.max() <= FP32_TOL


def test_checkpoint_interval_invariance_full_size(reference):
    # the reference records the whole run (default budget); this one keeps
    # only the last of 40 segments and recomputes the others
    r = reference
    pr = _problem("double", interval=500)
    pr.record_budget_bytes = 1
    j, gd, ga = pr.value_and_grad(r["design"], AMP)
    assert j == r["j"]
    assert np.abs(gd - r["gd"]).max() <= 1e-12 * np.abs(r["gd"]).max()
    assert abs(ga - r["ga"]) <= 1e-12 * abs(r["ga"])
