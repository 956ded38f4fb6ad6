"""p3smi: batched polar Newton-Raphson power flow on one AMD Instinct MI355X.

Modules:
  nr_klu       drop-in for the reference's nr_klu surface (Solver, solve_single, debug_J)
  grids        deterministic synthetic grids and their load cases
  timeseries   DC-power-flow start voltage, usable CPU count
  contingency  N-1 outage cases: stamps, islands, reference buses, served masks
"""

from ._lib import P3smiError, build, device_count  # noqa: F401

__version__ = "0.1.0"
