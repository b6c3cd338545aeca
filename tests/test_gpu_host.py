"""The C++ host layer end to end on the GPU: tahoe_hip runs input decks
through TotalLagrangianHIPT / HIPMSRMatrixT / NLSolverHIP and writes the
reference's .run layout.

- the reference's own deck Q1P0.3D.b (updated_lagrangian_Q1P0, mapped
  nodes, 100% biaxial stretch) against the reference's committed output
  (13 significant digits);
- the C1 configuration (tl_n16: 16^3 total-Lagrangian, 5 load steps)
  against the converged displacements of the reference run that made the
  fixture, with the device-resident Newton loop, the host-driven
  NLSolver loop and the mixed-precision path (all within 1e-9 relative).
"""
import os

import numpy as np
import pytest

from conftest import load_golden
from host_decks import DECKS, parse_run, run, write_fixture_deck

pytestmark = pytest.mark.gpu


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_reference_deck_q1p0_3d_b(tmp_path):
    out = str(tmp_path / "q.run")
    r = run(["-f", os.path.join(DECKS, "Q1P0.3D.b.xml"), "-o", out])
    assert r.returncode == 0, r.stdout + r.stderr
    ref = np.load(os.path.join(os.path.dirname(DECKS), "q1p0_3d_b_run.npz"))
    got = parse_run(out)
    assert [t for t, _ in got] == list(ref["time"])
    for (t, u), ur in zip(got, ref["u"]):
        assert np.abs(u - ur).max() <= 1e-12 * np.abs(ur).max() + 1e-15, t


@pytest.mark.parametrize("name,args", [
    ("tl_n16", []),
    ("tl_n4", ["-driver", "host"]),
    ("q1p0_n4", []),
    ("tl_n8", ["-precision", "mixed", "-linear_tolerance", "1e-6"]),
])
def test_fixture_deck(tmp_path, name, args):
    g = load_golden(name)
    xml = write_fixture_deck(g, str(tmp_path))
    out = str(tmp_path / "deck.run")
    r = run(["-f", xml, "-o", out, "-digits", "16"] + args, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = parse_run(out)
    assert len(got) == int(g["num_steps"])
    for k, (t, u) in enumerate(got):
        assert rel_err(u, g["step_u"][k]) <= 1e-9, (name, k)
    if not args:
        iters = [int(l.split(":")[1].split()[0]) for l in r.stdout.splitlines() if l.startswith(" step ")]
        assert iters == g["step_iters"].tolist()


@pytest.mark.parametrize("deck,golden", [("beam.Newton.xml", "beam_newton_run.npz"),
                                         ("beam.Newton.TotLag.xml", "beam_newton_totlag_run.npz"),
                                         ("beam.PCG.xml", "beam_pcg_run.npz"),
                                         ("Q1P0.3D.a.xml", "q1p0_3d_a_run.npz")])
def test_reference_beam_decks(tmp_path, deck, golden):
    """The reference's own beam decks (updated_lagrangian / total_lagrangian /
    updated_lagrangian_Q1P0, large_strain_StVenant, force_BC; nonlinear_solver
    with profile_matrix, or PCG_solver -- nonlinear CG with line search --
    with diagonal_matrix) against the reference's committed output (13
    significant digits; Q1P0.3D.a ramps the force over 5 steps, and its
    forward-difference tangent leaves the last Newton iterate within 1e-10)."""
    out = str(tmp_path / "beam.run")
    r = run(["-f", os.path.join(DECKS, deck), "-o", out, "-linear_tolerance", "1e-13"])
    assert r.returncode == 0, r.stdout + r.stderr
    ref = np.load(os.path.join(os.path.dirname(DECKS), golden))
    got = parse_run(out)
    assert [t for t, _ in got] == list(ref["time"])
    for (t, u), ur in zip(got, ref["u"]):
        assert np.abs(u - ur).max() <= 1e-10 * np.abs(ur).max(), t


@pytest.mark.parametrize("lin_tol", ["1.0e-3", "1.0e-10"])
def test_newton_krylov_deck(tmp_path, lin_tol):
    """A fixture deck with <newton_krylov_solver gmres_restart linear_tolerance
    max_linear_iter="0"> (NLSolver_NK): Newton with the device GMRES reaches
    the reference's converged displacements."""
    g = load_golden("tl_n8")
    xml = write_fixture_deck(g, str(tmp_path))
    s = open(xml).read()
    s = s.replace('<nonlinear_solver ', f'<newton_krylov_solver gmres_restart="30" linear_tolerance="{lin_tol}" '
                  'max_linear_iter="0" ').replace("</nonlinear_solver>", "</newton_krylov_solver>")
    open(xml, "w").write(s)
    out = str(tmp_path / "nk.run")
    r = run(["-f", xml, "-o", out, "-digits", "16"], timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = parse_run(out)
    for k, (t, u) in enumerate(got):
        assert rel_err(u, g["step_u"][k]) <= 1e-9, k
