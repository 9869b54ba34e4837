"""chapsim_amd — MI355X-native RK3/Crank-Nicolson time step for CHAPSim's periodic channel.

The package directory is ``chapsim1.0_amd``; it is imported under the name ``chapsim_amd``
(see ``load`` in __graft_entry__.py).  Product code: ``csrc/`` (HIP kernels + C-ABI, built
to ``libchapsim_hip.so``), ``fortran/`` (ISO_C_BINDING module and drop-in substep),
``tg.py`` (Python host mirror of the reference's substep interface).
"""
from .tg import (MIXED, PLAIN_FP64, ChannelTG, ChsError, DeviceArray, Ensemble,  # noqa: F401
                 channel_setup, lib, tdma_lines)
from .init import poiseuille_perturbed  # noqa: F401
