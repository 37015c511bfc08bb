"""Drop-ins for the reference package ``rb`` (rb/splitfe.py, rb/gpusolve.py, rb/replica3.py, rb/rbsim.py)."""
