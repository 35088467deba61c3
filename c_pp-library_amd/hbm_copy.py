"""Calibration run for HBM byte counters: one device-to-device copy of a known
size (1 GiB of float64 read, 1 GiB written) done by torch's elementwise copy
kernel, repeated a few times.  Profiling it with rocprofv3 --pmc FETCH_SIZE
(or WRITE_SIZE) gives the counter's reading per dispatch for exactly 1 GiB of
algorithmic traffic, the factor that turns the counters of the solver kernels
into bytes.  Prints the copy bandwidth measured with HIP events."""
import json

import torch

N = (1 << 30) // 8
src = torch.ones(N, dtype=torch.float64, device="cuda")
dst = torch.empty_like(src)
dst.copy_(src)
torch.cuda.synchronize()
e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
reps = 5
e0.record()
for _ in range(reps):
    dst.copy_(src)
e1.record()
torch.cuda.synchronize()
ms = e0.elapsed_time(e1) / reps
print(json.dumps({"bytes_read": N * 8, "bytes_written": N * 8, "ms_per_copy": ms,
                  "copy_GBps": 2 * N * 8 / (ms / 1e3) / 1e9}))
