"""Standard test phantoms (reference: python/usx/phantoms.py:17-73)."""

from __future__ import annotations

from ._usx import Medium, Phantom, Probe, Vec3


def _speckle(probe: Probe, medium: Medium, x_half: float, z_min: float, z_max: float,
             y_half: float, seed: int, density_per_cell: float) -> Phantom:
    n = Phantom.recommended_scatterer_count(probe, medium, 2 * x_half, z_max - z_min,
                                            2 * y_half, f_number=2.0, per_cell=density_per_cell)
    n = int(min(max(n, 2000), 400_000))
    return Phantom.speckle_box(-x_half, x_half, z_min, z_max, y_half, n, seed)


def wire_targets(depths=(0.010, 0.020, 0.030, 0.040, 0.050),
                 lateral=(-0.006, 0.0, 0.006)) -> Phantom:
    """Isolated point scatterers on a (depth x lateral) grid."""
    return Phantom.point_targets([Vec3(float(x), 0.0, float(z)) for z in depths for x in lateral],
                                 1.0)


def cyst_phantom(probe: Probe, medium: Medium | None = None,
                 cysts=((0.0, 0.020, 0.003, 0.0),
                        (-0.007, 0.035, 0.003, 0.25),
                        (0.007, 0.035, 0.003, 2.5)),
                 z_min: float = 0.005, z_max: float = 0.055,
                 seed: int = 3, density_per_cell: float = 12.0) -> Phantom:
    """Speckle background with (x, z, radius, amplitude_scale) inclusions."""
    half = 0.5 * probe.aperture_width() * 1.15
    ph = _speckle(probe, medium or Medium(), half, z_min, z_max, 2e-4, seed, density_per_cell)
    for x, z, r, scale in cysts:
        ph.scale_amplitude_in_circle(Vec3(float(x), 0.0, float(z)), float(r), float(scale))
    return ph


def resolution_and_contrast(probe: Probe, medium: Medium | None = None,
                            z_min: float = 0.005, z_max: float = 0.055,
                            seed: int = 5) -> Phantom:
    """Cyst for contrast plus four bright wires for resolution."""
    ph = cyst_phantom(probe, medium, cysts=((0.006, 0.030, 0.0035, 0.0),),
                      z_min=z_min, z_max=z_max, seed=seed)
    ph.extend(Phantom.point_targets(
        [Vec3(-0.008, 0.0, float(z)) for z in (0.012, 0.024, 0.036, 0.048)], 40.0))
    return ph
