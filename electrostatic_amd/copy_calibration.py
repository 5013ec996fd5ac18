"""Plain device-to-device copy of a known size (256 MiB, three launches).

Run under `rocprofv3 --pmc FETCH_SIZE` and `--pmc WRITE_SIZE` to calibrate
the traffic counters: bytes per launch divided by the counter's KiB gives
the factor applied to the RGF kernels' counters.
"""
import torch

N = 1 << 25   # doubles: 256 MiB

if __name__ == "__main__":
    a = torch.ones(N, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {N * 8} bytes per launch, 3 launches")
