"""Plain device-to-device copy of a known size, used to calibrate HBM byte counters
(rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE): 256 MiB read and 256 MiB written per copy."""
import torch

NBYTES = 256 << 20

if __name__ == "__main__":
    a = torch.ones(NBYTES, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(f"copied {NBYTES} bytes x3")
