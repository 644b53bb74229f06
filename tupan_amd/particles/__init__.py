"""Device-resident particle systems (tupan/particles)."""
from .system import ParticleSystem  # noqa: F401
