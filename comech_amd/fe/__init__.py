"""Drop-ins for the reference package ``fe`` (fe/solver.py, fe/gmg.py, fe/simulate.py)."""
