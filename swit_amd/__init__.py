"""MI355X-native (gfx950) 2-D acoustic forward/adjoint solver with SWIT's
toolbox/solver.py call surface."""

__version__ = "0.1.0"
