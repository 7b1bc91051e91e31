"""Drop-in for the reference extension module `fluxCal_2ndOrder_jh_improved` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import fluxCal_2ndOrder_jh_improved` resolves here."""
from hipims_amd.ops import add_flux_order2 as addFlux  # noqa: F401
