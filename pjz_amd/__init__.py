"""MI355X-native fdtdz() Yee FDTD kernel behind pjz's call surface.

``fdtdz`` is the drop-in for the kernel pjz calls at
``src/pjz/_field.py:254-269`` (``fdtdz_batch`` runs many at once);
``boundaries`` holds the helpers that build its inputs and project its
snapshots (``src/pjz/_field.py:56-150``); ``field``, ``scatter`` and
``SimParams`` mirror pjz's ``field()`` / ``scatter()`` layer over it
(``src/pjz/_field.py:12-53, 171-442``); ``mode``, ``epsilon`` / ``render``
and ``density`` are pjz's set-up steps before it (``src/pjz/_mode.py``,
``_epsilon.py``, ``_density.py``), run on the tensors' device.
"""

from .boundaries import (absorption_mask, frequency_components, output_phases,
                         pml_sigma, ramped_sin, sampling_interval, source_amplitude)
from .fdtdz_torch import fdtdz, fdtdz_batch, fdtdz_batch_value_and_grad
from ._density import density
from ._epsilon import epsilon, render
from ._field import SimParams, field
from ._mode import mode
from ._scatter import scatter

__all__ = ["fdtdz", "fdtdz_batch", "fdtdz_batch_value_and_grad", "field", "scatter", "mode", "epsilon", "render", "density",
           "SimParams", "absorption_mask", "pml_sigma", "ramped_sin", "sampling_interval",
           "output_phases", "frequency_components", "source_amplitude"]
