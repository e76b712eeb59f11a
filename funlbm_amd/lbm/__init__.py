from .base import Config, LBMBase, create_lbm_config
from .lbm3d import LBMD3, LBMD3Q13, LBMD3Q15, LBMD3Q19, LBMD3Q27

__all__ = ["LBMBase", "LBMD3", "LBMD3Q19", "LBMD3Q27", "LBMD3Q15", "LBMD3Q13", "Config", "create_lbm_config"]
