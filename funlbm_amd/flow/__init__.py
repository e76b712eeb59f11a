from .base import Flow, FlowConfig, Param
from .d3 import FlowD3, FlowD3Q13, FlowD3Q15, FlowD3Q19, FlowD3Q27

__all__ = [
    "Flow",
    "FlowD3",
    "FlowD3Q19",
    "FlowD3Q13",
    "FlowConfig",
    "Param",
    "create_flow",
    "FlowD3Q15",
    "FlowD3Q27",
]


def create_flow(flow_config: FlowConfig, *args, **kwargs):
    """Dispatch on param_type (src/funlbm/flow/__init__.py:16-41)."""
    if flow_config.param_type == "D3Q27":
        return FlowD3Q27(config=flow_config, *args, **kwargs)
    elif flow_config.param_type == "D3Q19":
        return FlowD3Q19(config=flow_config, *args, **kwargs)
    elif flow_config.param_type == "D3Q15":
        return FlowD3Q15(config=flow_config, *args, **kwargs)
    elif flow_config.param_type == "D3Q13":
        return FlowD3Q13(config=flow_config, *args, **kwargs)
    else:
        raise ValueError("Unknown parameter type: {}".format(flow_config.param_type))
