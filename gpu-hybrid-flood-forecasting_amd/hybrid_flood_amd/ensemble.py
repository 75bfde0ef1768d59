"""Batched ensemble of independent scenarios on grids of uneven size.

Each member is a full domain (its own DEM, Manning field, mask, rain series,
boundary codes and controls).  ``EnsembleRunner.run`` advances all members to
``t_end`` in graph-captured launches; every launch takes one CFL
step of every unfinished member, with each member's own dt.  The per-member
arithmetic is the single-domain fused step, so a member's result is identical
to running it alone with ``integrate_to_outputs(state, params, [0, t_end])``.
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .numerics import SWEParams, SWEState, _host_f32, _scalar


@dataclass
class Member:
    params: SWEParams
    state: SWEState  # initial state in, final state out (updated in place by run)


class EnsembleRunner:
    """Owns one ``SweEnsemble*`` over a list of members."""

    def __init__(self, members: list[Member], *, variant: str = "fast"):
        if not members:
            raise ValueError("An ensemble needs at least one member.")
        lib = _lib.load()
        self.members = members
        self._keep = []
        descs = (_lib.SweMemberDesc * len(members))()
        for k, m in enumerate(members):
            p = m.params
            times = _host_f32(p.rainfall_times_s)
            rates = _host_f32(p.rainfall_rates_m_s)
            if times.shape != rates.shape:
                raise ValueError("Rainfall times and rates must have the same length.")
            self._keep += [times, rates, p, m.state]
            ny, nx = p.bed.shape
            for t in m.state:
                if tuple(t.shape) != (ny, nx) or t.dtype != torch.float32 or not t.is_cuda:
                    raise ValueError(f"member {k}: state must be float32 GPU tensors of shape {(ny, nx)}")
            descs[k].params = _lib.SweParamsDesc(
                nx=nx, ny=ny, bed=p.bed.data_ptr(), manning_n=p.manning_n.data_ptr(),
                domain_mask=p.domain_mask.data_ptr(),
                rainfall_multiplier=p.rainfall_multiplier.data_ptr(),
                boundary_types=(ctypes.c_int32 * 4)(*[int(b) for b in p.boundary_types]),
                rainfall_times_s=times.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                rainfall_rates_m_s=rates.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                n_rainfall=len(times), default_rainfall_m_s=_scalar(p.default_rainfall_m_s),
                dx=_scalar(p.dx), dy=_scalar(p.dy), gravity=_scalar(p.gravity), cfl=_scalar(p.cfl),
                dry_tolerance=_scalar(p.dry_tolerance), max_dt=_scalar(p.max_dt))
            descs[k].state = _lib.state_struct(*m.state)
        torch.cuda.current_stream().synchronize()
        handle = ctypes.c_void_p()
        _lib.check(lib.swe_ensemble_create(descs, len(members), _lib.VARIANTS[variant],
                                           ctypes.byref(handle)), "swe_ensemble_create")
        self.handle = handle
        self.variant = variant
        self.cells = int(sum(m.params.bed.numel() for m in members))

    def __del__(self):
        h = getattr(self, "handle", None)
        if h and _lib._lib is not None:
            _lib._lib.swe_ensemble_destroy(h)
            self.handle = None

    def set_tile_rows(self, rows: int) -> None:
        _lib.check(_lib.load().swe_ensemble_set_tile_rows(self.handle, int(rows)), "set_tile_rows")

    def set_graph_chunk(self, chunk: int) -> None:
        _lib.check(_lib.load().swe_ensemble_set_graph_chunk(self.handle, int(chunk)), "set_graph_chunk")

    def work_shape(self) -> tuple[int, int]:
        w, g = ctypes.c_int32(), ctypes.c_int32()
        _lib.check(_lib.load().swe_ensemble_total_waves(self.handle, ctypes.byref(w), ctypes.byref(g)),
                   "total_waves")
        return w.value, g.value

    def set_rain_rounding(self, mode: str) -> None:
        """'nearest' (the reference's) or 'stochastic' rounding of every
        member's rain source (swe_ensemble_set_rain_rounding)."""
        _lib.check(_lib.load().swe_ensemble_set_rain_rounding(self.handle, _lib.rain_rounding_code(mode)),
                   "swe_ensemble_set_rain_rounding")

    def static_layout(self, member: int) -> int:
        """The packed static record member ``member``'s last start chose (fast
        variant): 2 = bed + Manning class, 1 = (bed, n^2), 0 = (bed, n^2, multiplier)."""
        v = ctypes.c_int32()
        _lib.check(_lib.load().swe_ensemble_get_static_layout(self.handle, int(member), ctypes.byref(v)),
                   "swe_ensemble_get_static_layout")
        return v.value

    def _infos(self):
        return (_lib.SweRunInfo * len(self.members))()

    def run(self, t_end: float, *, max_steps: int = 100_000) -> list[dict]:
        """Advance every member from t=0 to ``t_end``; states are updated in place."""
        infos = self._infos()
        rc = _lib.load().swe_ensemble_run(self.handle, float(t_end), int(max_steps), infos,
                                          _lib.stream_handle())
        if rc not in (_lib.SWE_OK, _lib.SWE_EMAXSTEPS, _lib.SWE_ENONFINITE):
            _lib.check(rc, "swe_ensemble_run")
        out = [i.as_dict() for i in infos]
        if rc != _lib.SWE_OK:
            raise RuntimeError(_lib.last_error())
        return out

    # Driven form, for benchmarks: begin / advance / poll / end.
    def begin(self, t_end: float, max_steps: int = 100_000) -> None:
        _lib.check(_lib.load().swe_ensemble_begin(self.handle, float(t_end), int(max_steps),
                                                  _lib.stream_handle()), "swe_ensemble_begin")

    def advance(self, launches: int) -> None:
        _lib.check(_lib.load().swe_ensemble_advance(self.handle, int(launches), _lib.stream_handle()),
                   "swe_ensemble_advance")

    def running(self) -> int:
        r = ctypes.c_int32()
        _lib.check(_lib.load().swe_ensemble_poll(self.handle, ctypes.byref(r), _lib.stream_handle()),
                   "swe_ensemble_poll")
        return r.value

    def compact(self) -> None:
        """Drop finished members from the launch (call between chunks)."""
        _lib.check(_lib.load().swe_ensemble_compact(self.handle, _lib.stream_handle()),
                   "swe_ensemble_compact")

    def end(self) -> tuple[int, list[dict]]:
        infos = self._infos()
        rc = _lib.load().swe_ensemble_end(self.handle, infos, _lib.stream_handle())
        if rc not in (_lib.SWE_OK, _lib.SWE_EMAXSTEPS, _lib.SWE_ENONFINITE):
            _lib.check(rc, "swe_ensemble_end")
        return rc, [i.as_dict() for i in infos]


def member_from_case(case, *, device=None, n_scale: float = 1.0, rain_scale: float = 1.0) -> Member:
    """Build a member from a ``synthetic.Case`` with scaled roughness and rain."""
    from .numerics import make_params

    params = make_params(
        case.bed, (case.manning_n * np.float32(n_scale)).astype(np.float32), case.domain_mask,
        boundary_types=case.boundary_types, rainfall_times_s=case.rain_times_s,
        rainfall_rates_m_s=(case.rain_rates_m_s * np.float32(rain_scale)).astype(np.float32),
        dx=case.dx, cfl=case.cfl, dry_tolerance=case.dry_tolerance, max_dt=case.max_dt,
        device=device)
    dev = params.bed.device
    h = torch.as_tensor(case.h0, device=dev).contiguous()
    return Member(params, SWEState(h, torch.zeros_like(h), torch.zeros_like(h)))
