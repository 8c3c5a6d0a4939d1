"""Initial parameters of the warm-start network (float32, layout of include/dhip.h): weights
N(0, 1/fan_in), zero biases, output weights scaled by 0.1 so that the first prediction is close
to "no change"."""
import numpy as np

NF, H, NO = 12, 64, 2
SHAPES = [("W1", (H, NF)), ("b1", (H,)), ("W2", (H, H)), ("b2", (H,)), ("W3", (H, H)), ("b3", (H,)),
          ("W4", (NO, H)), ("b4", (NO,))]


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            w = rng.normal(scale=1.0 / np.sqrt(shp[1]), size=shp)
            if name == "W4":
                w *= 0.1
            parts.append(w.ravel())
        else:
            parts.append(np.zeros(shp).ravel())
    return np.concatenate(parts).astype(np.float32)
