"""Drop-in for the reference extension module `infiltrationCalculation` (CHANS-HiPIMS/hipims/cuda/setup.py:43-46).

Put this directory on sys.path and `import infiltrationCalculation` resolves here."""
from hipims_amd.ops import add_infiltration as addinfiltration  # noqa: F401
