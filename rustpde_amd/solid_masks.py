"""Mask builders for the solid-fluid interaction (src/navier/solid_masks.rs:34-162).

Host-side set-up: each returns [mask, value], physical [len(x), len(y)] arrays for
`Navier2D.solid` (x, y: the grid points of the temperature field times the axis scales,
e.g. `rustpde_amd.checkpoint.coords`).  The smoothing layer follows the reference's
0.5 (1 - tanh(2 (r - R) / d)) with d = R / 10 (or height / 10).
"""
import math

import numpy as np


def solid_cylinder_inner(x, y, x0, y0, radius):
    """Everything with r < radius is solid (solid_masks.rs:34-58)."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    mask = np.zeros((len(x), len(y)))
    layer = radius / 10.0
    for i, xi in enumerate(x):
        for j, yj in enumerate(y):
            r = math.sqrt((x0 - xi) ** 2.0 + (y0 - yj) ** 2.0)
            if r < radius - layer:
                mask[i, j] = 1.0
            elif r < radius + layer:
                mask[i, j] = 0.5 * (1.0 - math.tanh(2.0 * (r - radius) / layer))
    return [mask, np.zeros_like(mask)]


def solid_roughness_sinusoid(x, y, height, wavenumber):
    """Sinusoidal roughness elements on the bottom (value 0.5) and top (value -0.5) plates
    (solid_masks.rs:61-101)."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    mask = np.zeros((len(x), len(y)))
    value = np.zeros((len(x), len(y)))
    bottom, top = y[0], y[-1]
    layer = height / 10.0
    for i, xi in enumerate(x):
        y_rough = height * (top - bottom) / 2.0 * (math.sin(wavenumber * xi) + 0.5)
        for j, yj in enumerate(y):
            for y_dist, val in ((yj - bottom, 0.5), (top - yj, -0.5)):
                if y_dist <= y_rough:
                    mask[i, j] = 1.0
                    value[i, j] = val
                elif y_dist <= y_rough + layer:
                    mask[i, j] = 0.5 * (1.0 - math.tanh(2.0 * (y_dist - y_rough) / layer))
                    value[i, j] = val
    return [mask, value]


def solid_porosity(x, y, diameter, porosity):
    """A regular array of cylinders that leaves the fraction `porosity` open
    (solid_masks.rs:104-139); overlapping masks add up, as in the reference."""
    x = np.asarray(x, dtype=float)
    y = np.asarray(y, dtype=float)
    mask = np.zeros((len(x), len(y)))
    radius = diameter / 2.0
    length, height = x[-1] - x[0], y[-1] - y[0]
    # f64::round: half away from zero (the counts are non-negative)
    nxc = math.floor(math.sqrt((1.0 - porosity) * 4.0 * length ** 2 / (math.pi * diameter ** 2)) + 0.5)
    nyc = math.floor(math.sqrt((1.0 - porosity) * 4.0 * height ** 2 / (math.pi * diameter ** 2)) + 0.5)
    dx = (length - nxc * diameter) / (nxc + 1.0)
    dy = (height - nyc * diameter) / (nyc + 1.0)
    ox = x[0] + dx + radius
    for _ in range(int(nxc)):
        oy = y[0] + dy + radius
        for _ in range(int(nyc)):
            mask += solid_cylinder_inner(x, y, ox, oy, radius)[0]
            oy += dy + diameter
        ox += dx + diameter
    return [mask, np.zeros_like(mask)]
