"""Worker base class (src/funlbm/base/work.py:7-63).

The reference picks cuda > mps > cpu and defaults to "cpu".  This package
computes with HIP kernels only, so the default is the ROCm device and a CPU
device is refused.
"""

import logging

import torch

logger = logging.getLogger("funlbm_amd")


def device_detect(device=None) -> torch.device:
    """The ROCm device to run on (src/funlbm/base/work.py:7-30)."""
    if device is None or device == "auto":
        device = "cuda"
    dev = torch.device(device)
    if dev.type != "cuda":
        raise RuntimeError(
            f"funlbm_amd runs its timestep on the GPU with HIP kernels; device {device!r} is not a ROCm device"
        )
    if dev.index is None:
        dev = torch.device("cuda", torch.cuda.current_device() if torch.cuda.is_available() else 0)
    return dev


class Worker:
    def __init__(self, device=None, *args, **kwargs):
        self.device = device_detect(device)

    def dump_checkpoint(self, group=None, *args, **kwargs):
        pass

    def load_checkpoint(self, group=None, *args, **kwargs):
        pass
