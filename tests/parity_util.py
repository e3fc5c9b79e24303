"""Comparison helpers shared by the GPU parity tests."""
import numpy as np


def class_boundary_mask(cls_img: np.ndarray) -> np.ndarray:
    """True where a pixel has a 4-neighbour (or 8) of a different class."""
    c = cls_img
    m = np.zeros(c.shape, bool)
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            if dy == 0 and dx == 0:
                continue
            sh = np.roll(np.roll(c, dy, 0), dx, 1)
            m |= sh != c
    return m


def median_rel_err(a: np.ndarray, b: np.ndarray, floor: float = 0.0) -> float:
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    sel = (np.abs(b) > floor) & (np.abs(a) > floor)
    if not sel.any():
        return 0.0
    return float(np.median(np.abs(a[sel] - b[sel]) / np.abs(b[sel])))
