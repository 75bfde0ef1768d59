"""Step-throughput harness with the reference's protocol
(jax_solver/benchmarking.py:17-97): a fixed-dt loop of ``finite_volume_step``
with time ``index * dt``, warm-up excluded, one synchronised timed pass.

The reference runs the loop on every JAX platform (cpu, gpu); this build has
only the GPU path, so the cpu record reports ``available: False``.
"""

from __future__ import annotations

import json
from datetime import datetime, timezone
from pathlib import Path
from time import perf_counter
from typing import Any

import numpy as np
import torch

from . import _lib
from .numerics import SWEParams, SWEState, context_for


def benchmark_compiled_steps(state: SWEState, params: SWEParams, *, iterations: int = 100,
                             dt_s: float = 1.0, output_directory: str | Path | None = None,
                             variant: str | None = None) -> dict[str, Any]:
    """Benchmark repeated fused steps at fixed dt on the current GPU."""
    if iterations < 1:
        raise ValueError("Benchmark iterations must be positive.")
    records: list[dict[str, Any]] = [
        {"platform": "cpu", "available": False,
         "reason": "The HIP step has no CPU backend."}
    ]
    lib = _lib.load()
    ctx = context_for(params, variant)
    dev = state.h.device
    bufs = [SWEState(*(x.clone() for x in state)),
            SWEState(*(torch.empty_like(x) for x in state))]
    dt32 = np.float32(dt_s)
    stream = _lib.stream_handle(dev)

    def run(n):
        for k in range(n):
            src, dst = bufs[k & 1], bufs[(k + 1) & 1]
            t = float(np.float32(k) * dt32)  # index * dt in the state dtype (benchmarking.py:54)
            _lib.check(lib.swe_step(ctx.handle, _lib.state_struct(*src), _lib.state_struct(*dst),
                                    float(dt32), t, stream), "swe_step")

    run(iterations)  # warm-up (module load, caches), excluded like the JIT warm-up
    torch.cuda.synchronize(dev)
    bufs[0] = SWEState(*(x.clone() for x in state))
    torch.cuda.synchronize(dev)
    start = perf_counter()
    run(iterations)
    torch.cuda.synchronize(dev)
    elapsed = perf_counter() - start
    records.append({
        "platform": "gpu", "available": True, "device": torch.cuda.get_device_name(dev),
        "iterations": iterations, "total_seconds": elapsed,
        "seconds_per_step": elapsed / iterations, "steps_per_second": iterations / elapsed,
        "cell_updates_per_second": state.h.numel() * iterations / elapsed,
        "grid_shape": list(state.h.shape), "dtype": "float32", "kernel_variant": ctx.variant,
        "includes_jit_compilation": False,
    })
    report = {
        "generated_at": datetime.now(timezone.utc).isoformat(),
        "library": lib.swe_version().decode(),
        "devices": [torch.cuda.get_device_name(dev)],
        "gpu_benchmark_available": True,
        "benchmarks": records,
    }
    if output_directory is not None:
        directory = Path(output_directory)
        directory.mkdir(parents=True, exist_ok=True)
        stamp = datetime.now(timezone.utc).strftime("%Y%m%dT%H%M%S%fZ")
        path = directory / f"benchmark_{stamp}.json"
        path.write_text(json.dumps(report, indent=2, sort_keys=True), encoding="utf-8")
        report["output_path"] = str(path)
    return report
