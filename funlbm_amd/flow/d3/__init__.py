from .base import FlowD3, _lattice_class

FlowD3Q13 = _lattice_class("D3Q13")
FlowD3Q15 = _lattice_class("D3Q15")
FlowD3Q19 = _lattice_class("D3Q19")
FlowD3Q27 = _lattice_class("D3Q27")

__all__ = ["FlowD3", "FlowD3Q27", "FlowD3Q19", "FlowD3Q15", "FlowD3Q13"]
