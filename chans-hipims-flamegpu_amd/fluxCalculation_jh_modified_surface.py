"""Drop-in for the reference extension module `fluxCalculation_jh_modified_surface` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import fluxCalculation_jh_modified_surface` resolves here."""
from hipims_amd.ops import add_flux_order1 as addFlux  # noqa: F401
