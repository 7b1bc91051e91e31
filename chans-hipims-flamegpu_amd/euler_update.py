"""Drop-in for the reference extension module `euler_update` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import euler_update` resolves here."""
from hipims_amd.ops import euler_update as update  # noqa: F401
