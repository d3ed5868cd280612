"""Time-delay amplitude extraction on the GPU (the producer of td00/td90).

Drop-in for pycwb/types/td_batch_inputs.py (``TDBatchInputs`` and its
``extract_td_vecs``) and ``TimeFrequencyMap.prepare_td_inputs``
(pycwb/types/time_frequency_map.py:482-521): the WDM map's 00/90 phases are
zero-padded by ``n_coeffs`` time bins, uploaded once as float32 planes, and
``extract_td_vecs(pixel_indices, K)`` runs the HIP kernel ``pycwb_td_vectors``
(restating batch_get_td_vecs, pycwb/utils/td_vector_kernels.py:131-199).
The TD filters come from ``WDM.set_td_filter`` (cwb-core/WDM.cc:638-700).

``extract_td_into`` writes straight into a likelihood batch's packed
``(P, n_ifo, 2K+1)`` {00, 90} table, so a cluster's amplitudes never leave HBM.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import lib, ptr, require_gpu, stream_ptr


@dataclass
class TDBatchInputs:
    """Padded quadrature planes and TD filter tables on the device."""
    padded00: torch.Tensor   # (n_time + 2 n_coeffs, M+1) float32
    padded90: torch.Tensor
    T0: torch.Tensor         # tap-major (2 n_coeffs+1, 2J+1) float64: the reference's table, transposed
    Tx: torch.Tensor
    M: int
    n_coeffs: int
    J: int
    trig: torch.Tensor = None  # (4J, 2) float64: sin, cos of q pi / (2J)

    @classmethod
    def from_map(cls, data: np.ndarray, td_filters, device=None) -> "TDBatchInputs":
        """``data``: complex (M+1, n_time) WDM map (00 real, 90 imaginary)."""
        device = device or require_gpu()
        data = np.asarray(data, dtype=np.complex128)
        nc = int(td_filters.n_coeffs)
        pad = ((nc, nc), (0, 0))
        p00 = np.pad(np.ascontiguousarray(data.real.T), pad).astype(np.float32)
        p90 = np.pad(np.ascontiguousarray(data.imag.T), pad).astype(np.float32)
        up = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a)).to(device=device, dtype=dt)
        J = int(td_filters.max_delay)
        q = np.arange(4 * J) * (np.pi / (2.0 * J))
        return cls(padded00=up(p00, torch.float32), padded90=up(p90, torch.float32),
                   T0=up(np.asarray(td_filters.T0).T, torch.float64),
                   Tx=up(np.asarray(td_filters.Tx).T, torch.float64),
                   M=int(td_filters.M), n_coeffs=nc, J=J,
                   trig=up(np.stack([np.sin(q), np.cos(q)], axis=1), torch.float64))

    @classmethod
    def from_device(cls, data: torch.Tensor, td_filters) -> "TDBatchInputs":
        """``data``: complex128 (M+1, n_time) device map (a t2w_device output
        slice); the padded planes are built on the device."""
        dev = data.device
        nc = int(td_filters.n_coeffs)
        planes = torch.view_as_real(data.transpose(0, 1).contiguous())   # (n_time, M+1, 2)
        pad = torch.zeros((nc, planes.shape[1], 2), dtype=planes.dtype, device=dev)
        padded = torch.cat([pad, planes, pad], dim=0).to(torch.float32)
        J = int(td_filters.max_delay)
        q = np.arange(4 * J) * (np.pi / (2.0 * J))
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(device=dev,
                                                                  dtype=torch.float64)
        return cls(padded00=padded[..., 0].contiguous(), padded90=padded[..., 1].contiguous(),
                   T0=up(np.asarray(td_filters.T0).T), Tx=up(np.asarray(td_filters.Tx).T),
                   M=int(td_filters.M), n_coeffs=nc, J=J,
                   trig=up(np.stack([np.sin(q), np.cos(q)], axis=1)))

    def _check_pixels(self, pixel_indices) -> torch.Tensor:
        dev = self.padded00.device
        pix = torch.as_tensor(np.asarray(pixel_indices.cpu() if isinstance(pixel_indices,
                                                                          torch.Tensor)
                                         else pixel_indices, dtype=np.int32)).to(dev)
        n_time = self.padded00.shape[0] - 2 * self.n_coeffs
        if pix.numel() and (int(pix.min()) < 0 or int(pix.max()) >= n_time * (self.M + 1)):
            raise ValueError("pixel index outside the time-frequency map")
        return pix.contiguous()

    def _launch(self, pix, K, out, row_stride, interleave, stream=None):
        rc = lib().pycwb_td_vectors(ptr(pix), int(pix.numel()), ptr(self.padded00),
                                    ptr(self.padded90), int(self.padded00.shape[0]), self.M,
                                    ptr(self.T0), ptr(self.Tx), ptr(self.trig), self.n_coeffs,
                                    self.J, int(K),
                                    ptr(out), int(row_stride), int(interleave),
                                    stream_ptr(stream))
        _lib.check(rc, "pycwb_td_vectors")

    def extract_td_vecs(self, pixel_indices, K: int, stream=None) -> torch.Tensor:
        """(n_pixels, 4K+2) float32 device tensor: rows [a00(-K..K), a90(-K..K)]
        (td_batch_inputs.py:45-67)."""
        pix = self._check_pixels(pixel_indices)
        out = torch.empty((pix.numel(), 4 * int(K) + 2), dtype=torch.float32,
                          device=self.padded00.device)
        if pix.numel():
            self._launch(pix, K, out, 4 * int(K) + 2, 0, stream)
        return out

    def extract_td_into(self, pixel_indices, K: int, td: torch.Tensor, td_off: int, ifo: int,
                        n_ifo: int, stream=None) -> None:
        """Write detector ``ifo``'s amplitudes of the listed pixels into a packed
        ``(P, n_ifo, 2K+1, 2)`` float32 table starting at pair ``td_off``."""
        pix = self._check_pixels(pixel_indices)
        nd = 2 * int(K) + 1
        need = 2 * (int(td_off) + pix.numel() * n_ifo * nd)
        if td.dtype != torch.float32 or not td.is_contiguous() or td.numel() < need:
            raise ValueError("td must be a contiguous float32 table large enough for the pixels")
        if not 0 <= ifo < n_ifo:
            raise ValueError("ifo outside [0, n_ifo)")
        view = td.view(-1)[2 * (int(td_off) + ifo * nd):]
        if pix.numel():
            self._launch(pix, K, view, 2 * n_ifo * nd, 1, stream)
