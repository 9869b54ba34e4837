"""chapsim_amd — MI355X-native RK3/Crank-Nicolson time step for CHAPSim's periodic channel.

The package directory is ``chapsim1.0_amd``; it is imported under the name ``chapsim_amd``
(see ``load`` in __graft_entry__.py).  Product code: ``csrc/`` (HIP kernels + C-ABI, built
to ``libchapsim_hip.so``), ``fortran/`` (ISO_C_BINDING module and drop-in substep),
``tg.py`` (Python host mirror of the reference's substep interface).
"""
from .tg import MIXED, PLAIN_FP64, ChannelTG, ChsError, channel_setup, lib  # noqa: F401
from .init import poiseuille_perturbed  # noqa: F401
