"""Copy kernel of known size (1 GiB read, 1 GiB written per clone) used to
calibrate the FETCH_SIZE / WRITE_SIZE hardware counters against real bytes."""
import torch

if __name__ == "__main__":
    x = torch.ones(2**28, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    for _ in range(3):
        y = x.clone()
    torch.cuda.synchronize()
    print("copy bytes per clone", x.numel() * 4, flush=True)
