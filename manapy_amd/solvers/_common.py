"""Shared pieces of the solver classes: the dt ring and host-scalar reads."""
import ctypes

from .. import _lib
from ..device import stream_ptr


class DeviceClock(object):
    """dt ring of the fused steps (see mfv_dt_ring_init): four doubles on the device."""

    def __init__(self, device):
        import torch
        self.ring = torch.zeros(4, dtype=torch.float64, device=device)
        self.dt0 = torch.zeros(1, dtype=torch.float64, device=device)
        self.step = 0

    def start(self):
        _lib.call('mfv_dt_ring_init', self.ring.data_ptr(), self.dt0.data_ptr(), stream_ptr())
        self.step = 0

    def dt_of_step(self, k):
        return self.ring[k % 3]

    @property
    def time(self):
        return float(self.ring[3])

    @property
    def next_dt(self):
        return float(self.ring[self.step % 3])


def host_double():
    return ctypes.c_double(0.)
