"""The C++ host classes (the reference's World / Species / PotentialSolver
interfaces over the C ABI) on their own code paths, through
pic_host_test: the PCG-to-Gauss-Seidel fallback of
PotentialSolver::solve() (PotentialSolver.cpp:222-224) and
Species::computeGasProperties (Species.cpp:777-804), each against the
oracle and the reference's fixture."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "engineering-degree-in-plasma-simulations_amd", "pic_host_test")


def _run(args):
    out = subprocess.run([EXE] + args, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr[-2000:]
    return json.loads(out.stdout.strip().splitlines()[-1])


def test_pcg_fallback_to_gauss_seidel(golden, tmp_path):
    """PCG limited to 2 iterations cannot converge; solve() falls back to
    the sweeps and converges to the reference's criterion.  The result
    agrees with the oracle's tight solve and with the reference's own
    fallback run (its loose Newton-step criterion, 1e-3 V)."""
    gr = golden["grid"]
    g = O.make_grid(int(gr[0]), int(gr[1]), int(gr[2]), gr[3:6], gr[6:9])
    gr[:9].astype(np.float64).tofile(tmp_path / "grid.bin")
    rho = golden["rho"]
    rho.tofile(tmp_path / "rho.bin")
    res = _run(["fallback", str(tmp_path / "grid.bin"), str(tmp_path / "rho.bin"),
                str(tmp_path / "phi.bin")])
    assert res["converged"] == 1 and res["ref_l2"] < 1e-6, res
    phi = np.fromfile(tmp_path / "phi.bin")
    nt = np.zeros(O.nn(g), np.int32)
    O.lib().orc_dirichlet_walls(g, nt)
    tight, ok, *_ = O.nrpcg(g, np.zeros(O.nn(g), np.int32), nt, np.zeros(O.nn(g)), rho, 0.0, 0.0,
                            1e20, 5000, 1e-12)
    scale = np.max(np.abs(tight))
    assert np.max(np.abs(phi - tight)) <= 1e-8 * scale
    assert golden["fb_ok"][0] == 1
    assert np.max(np.abs(golden["fb_phi_out"] - tight)) <= 1e-3  # the reference's own accuracy


def test_gas_properties_through_host_classes(golden, tmp_path):
    """getPartRef upload, sampleMoments on the device, computeGasProperties
    on the host: vel and T equal the reference's fixture (FP32 particles)"""
    gr = golden["grid"]
    gr[:9].astype(np.float64).tofile(tmp_path / "grid.bin")
    p = golden["ele_post_long"]
    p.tofile(tmp_path / "parts.bin")
    res = _run(["gas", str(tmp_path / "grid.bin"), str(tmp_path / "parts.bin"), repr(O.ME),
                str(tmp_path / "vel.bin"), str(tmp_path / "T.bin")])
    assert res["particles"] == p.size // 7
    vel = np.fromfile(tmp_path / "vel.bin")
    T = np.fromfile(tmp_path / "T.bin")
    ref_v, ref_T = golden["ele_vel"], golden["ele_T"]
    assert np.max(np.abs(vel - ref_v)) <= 1e-5 * np.max(np.abs(ref_v))
    # T is a difference of moments: relative to the largest, FP32 particle data
    assert np.max(np.abs(T - ref_T)) <= 1e-4 * np.max(np.abs(ref_T))
