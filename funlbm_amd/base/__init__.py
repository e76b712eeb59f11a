from .work import Worker, device_detect

__all__ = ["Worker", "device_detect"]
