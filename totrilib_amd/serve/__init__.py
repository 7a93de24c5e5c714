"""Continuous-batching serving of confocal renders and backprojections."""
from totrilib_amd.serve.arena import Arena, TensorArena
from totrilib_amd.serve.engine import (
    BackprojectRequest,
    RenderRequest,
    ServingEngine,
    Ticket,
    request_cost,
    request_pairs,
)

__all__ = ["Arena", "TensorArena", "ServingEngine", "RenderRequest", "BackprojectRequest",
           "Ticket", "request_cost", "request_pairs"]
