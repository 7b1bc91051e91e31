"""Drop-in for the reference extension module `station_PrecipitationCalculation` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import station_PrecipitationCalculation` resolves here."""
from hipims_amd.ops import add_station_precipitation as addStation_Precipitation  # noqa: F401
