from .server import Request, Server

__all__ = ["Request", "Server"]
