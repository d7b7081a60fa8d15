"""Timing probe of the surrogate's training steps (not a test; run under
rocprofv3 to reproduce profiles/mlp_training.md): 65,536 random 7-point
masks, 2 epochs at batches of 1024, 4096 and 8192 rows."""
import sys, time, numpy as np
from pathlib import Path
ROOT = Path(__file__).resolve().parents[1]
sys.path[:0] = [str(ROOT), str(ROOT / "generalized-toric-gpu-search_amd")]
from gtoric.surrogate import Surrogate
from gtoric.evolve import random_masks
rng = np.random.default_rng(1)
m = random_masks(rng, 65536, 7)
d = rng.integers(25, 36, size=65536).astype(np.float64)
sur = Surrogate(0, seed=1)
sur.fit(m[:4096], d[:4096], epochs=1, batch=1024)
for batch in (1024, 4096, 8192):
    t0 = time.perf_counter()
    losses = sur.fit(m, d, epochs=2, batch=batch)
    dt = time.perf_counter() - t0
    steps = 2 * ((65536 + batch - 1) // batch)
    print(f"batch {batch}: {dt*1e3:.2f} ms for {steps} steps = {dt/steps*1e6:.1f} us/step, loss {losses[-1]:.4f}", flush=True)
