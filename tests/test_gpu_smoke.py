"""The package's one-call smoke check (`__graft_entry__.smoke`) runs green.

It loads two thermal species into a 17x15x13 box (267993 macroparticles each),
deposits, solves Poisson with PCG, pushes and sorts, and checks charge density,
residual and pushed positions against the oracle.
"""
import importlib.util
import os

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.gpu
def test_smoke_entry_point(capsys):
    spec = importlib.util.spec_from_file_location("graft_entry", os.path.join(ROOT, "__graft_entry__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    mod.smoke()
    assert "smoke ok" in capsys.readouterr().out
