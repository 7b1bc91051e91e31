"""Drop-in for the reference extension module `fluxMask` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import fluxMask` resolves here."""
from hipims_amd.ops import flux_mask_update as update  # noqa: F401
