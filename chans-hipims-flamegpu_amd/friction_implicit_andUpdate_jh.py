"""Drop-in for the reference extension module `friction_implicit_andUpdate_jh` (CHANS-HiPIMS/hipims/cuda/setup.py:6-64).

Put this directory on sys.path and `import friction_implicit_andUpdate_jh` resolves here."""
from hipims_amd.ops import add_friction_euler_update as addFriction_eulerUpdate  # noqa: F401
