"""MI355X-native action evaluation for PIMC++ (short-range pair action, kinetic
action, Ewald long-range action and the bisection chain that calls them).

The compute path is the HIP library built from pimcplusplus_amd/csrc; this
package holds its ctypes binding (`abi`, `native`), the pair-table readers and
generators (`tables`) and the synthetic systems of the benchmarks (`workloads`).
"""

from . import abi, tables, workloads  # noqa: F401

__all__ = ["abi", "tables", "workloads"]
