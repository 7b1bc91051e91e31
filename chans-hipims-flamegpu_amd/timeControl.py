"""Drop-in for the reference extension module `timeControl` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import timeControl` resolves here."""
from hipims_amd.ops import update_timestep as updateTimestep  # noqa: F401
