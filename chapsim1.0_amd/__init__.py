"""chapsim_amd — MI355X-native RK3/Crank-Nicolson time step for CHAPSim's periodic channel.

The package directory is ``chapsim1.0_amd``; it is imported under the name ``chapsim_amd``
(see ``load`` in __graft_entry__.py).  Product code: ``csrc/`` (HIP kernels + C-ABI, built
to ``libchapsim_hip.so``), ``fortran/`` (ISO_C_BINDING module and drop-in substep),
``tg.py`` (Python host mirror of the reference's substep interfaces: ChannelTG for the
turbulence-generator domain, ChannelIO for the main domain), ``thermo.py`` (property tables
and the energy substep).
"""
from .tg import (MIXED, PLAIN_FP64, ChannelIO, ChannelTG, ChsError, DeviceArray, Ensemble,  # noqa: F401
                 channel_setup, heated_cfl, heated_substep, interp_initial_tg, lib, next_dt, random_field_tg, restart_read_io,
                 restart_read_tg, restart_write_io, restart_write_tg, tdma_lines)
from .init import poiseuille_perturbed  # noqa: F401
